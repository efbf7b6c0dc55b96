"""Learned GJK warm start: dataset labelling, training and evaluation.

The MLP (39-128-128-128-4, GELU) runs entirely in the library's own device
kernels (mfcl_mlp_* in include/mfcl.h): feature extraction, BF16 MFMA
forward/backward products and Adam.  Its output direction is passed as the
per-pair GJK guess, the hook of the reference's GJKSolver
(include/fcl/narrowphase/detail/gjk_solver.h:140-163), so the exact GJK/EPA
still decides every result.

Labels come from the exact signed-distance query (GJKSolver::
shapeSignedDistance semantics): for a pair at signed distance d with
witness points pa (on a) and pb (on b), the target direction in a's frame is
sign(d) * R_a^T (pa - pb) / |pa - pb| -- the direction whose support point on
the Minkowski difference a - b is its point nearest the origin, so GJK
started from it needs no further search when separated -- and the target
distance is d.

Usage (on a GPU):
  python -m mfcl.warmstart train --pairs 10000000 --seed 4 --out mind-fcl_amd/mfcl/data/warmstart_mlp.npz
  python -m mfcl.warmstart eval --pairs 1048576 --seed 5
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

from . import CONVEX_RESULT_DTYPE, MODE_SIGNED_DISTANCE, Context, DeviceArray, request
from . import workloads as W

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
DEFAULT_PARAMS = os.path.join(DATA_DIR, "warmstart_mlp.npz")
FEAT = 64


class _View:
    """A device pointer into an allocation (for batch slices)."""

    def __init__(self, base: DeviceArray, offset_bytes: int):
        self.ptr = C.c_void_p(base.ptr.value + int(offset_bytes))


def quat_rot(q: np.ndarray) -> np.ndarray:
    q = np.asarray(q, np.float64)
    q = q / np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((len(q), 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - w * z)
    R[:, 0, 2] = 2 * (x * z + w * y)
    R[:, 1, 0] = 2 * (x * y + w * z)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - w * x)
    R[:, 2, 0] = 2 * (x * z - w * y)
    R[:, 2, 1] = 2 * (y * z + w * x)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def targets_from_results(out: np.ndarray, pose_a: np.ndarray) -> np.ndarray:
    """(n, 4) targets: unit direction in a's frame (0 when undefined) and
    the signed distance."""
    d = out["distance"].astype(np.float64)
    v = out["pa"] - out["pb"]
    R = quat_rot(pose_a[:, 3:7])
    va = np.einsum("nji,nj->ni", R, v)  # R^T v
    nv = np.linalg.norm(va, axis=1)
    ok = (out["ok"] == 1) & (nv > 0) & (d != 0)
    t = np.zeros((len(out), 4), np.float32)
    t[ok, :3] = (np.sign(d[ok])[:, None] * va[ok] / nv[ok, None]).astype(np.float32)
    t[:, 3] = np.where(out["ok"] == 1, d, 0.0).astype(np.float32)
    return t


class PairSet:
    """A hull-pair workload uploaded to one context (shapes + device pairs)."""

    def __init__(self, ctx: Context, n: int, seed: int):
        self.ctx = ctx
        self.n = n
        self.h = W.hull_pairs(n, seed)
        h = self.h
        ctx.upload_shapes(h["types"], h["params"], h["vert_off"], h["vert_cnt"], h["verts"])
        self.pa, self.pb = ctx.to_device(h["pair_a"]), ctx.to_device(h["pair_b"])
        self.qa, self.qb = ctx.to_device(h["pose_a"]), ctx.to_device(h["pose_b"])
        self.out = ctx.alloc(n * CONVEX_RESULT_DTYPE.itemsize)
        self.guess = ctx.alloc(n * 12)
        self.req = request(MODE_SIGNED_DISTANCE)
        self.req_warm = request(MODE_SIGNED_DISTANCE, use_guess=1)

    def query(self, warm: bool):
        if warm:
            self.ctx.mlp_guess_dev(self.pa, self.pb, self.qa, self.qb, self.n, self.guess)
            self.ctx.convex_query_dev(self.pa, self.pb, self.qa, self.qb, self.n, self.req_warm, self.out,
                                      guess=self.guess)
        else:
            self.ctx.convex_query_dev(self.pa, self.pb, self.qa, self.qb, self.n, self.req, self.out)

    def results(self) -> np.ndarray:
        out = np.zeros(self.n, CONVEX_RESULT_DTYPE)
        self.out.download(out)
        return out

    def features(self) -> np.ndarray:
        f = self.ctx.alloc(self.n * FEAT * 4)
        self.ctx.mlp_features_dev(self.pa, self.pb, self.qa, self.qb, self.n, f)
        x = np.zeros((self.n, FEAT), np.float32)
        f.download(x)
        f.free()
        return x

    def free(self):
        for a in (self.pa, self.pb, self.qa, self.qb, self.out, self.guess):
            a.free()


def make_dataset(ctx: Context, n_pairs: int, seed: int, chunk: int = 1 << 20, log=print):
    """Features and exact-query targets for n_pairs hull pairs."""
    X = np.zeros((n_pairs, FEAT), np.float32)
    T = np.zeros((n_pairs, 4), np.float32)
    done = 0
    ci = 0
    while done < n_pairs:
        m = min(chunk, n_pairs - done)
        t0 = time.time()
        ps = PairSet(ctx, m, seed * 100003 + ci)
        ps.query(False)
        out = ps.results()
        X[done:done + m] = ps.features()
        T[done:done + m] = targets_from_results(out, ps.h["pose_a"])
        ps.free()
        done += m
        ci += 1
        log(f"[data] {done}/{n_pairs} pairs labelled ({time.time() - t0:.1f}s for this chunk)")
    return X, T


def train(ctx: Context, X: np.ndarray, T: np.ndarray, epochs: int, batch: int = 65536, lr: float = 2e-3,
          seed: int = 0, log=print):
    """Adam on the device; returns the per-epoch mean losses."""
    n = len(X) // batch * batch
    rng = np.random.default_rng(seed)
    perm = rng.permutation(len(X))[:n]
    dX = ctx.to_device(np.ascontiguousarray(X[perm]))
    dT = ctx.to_device(np.ascontiguousarray(T[perm]))
    nb = n // batch
    losses = []
    step = 0
    total = epochs * nb
    for ep in range(epochs):
        order = rng.permutation(nb)
        acc = 0.0
        t0 = time.time()
        for b in order:
            frac = step / max(1, total - 1)
            cur = lr * (0.02 + 0.98 * 0.5 * (1 + np.cos(np.pi * frac)))  # cosine decay to 2% of lr
            hyper = np.array([cur, 0.9, 0.999, 1e-8, 0.0, 1.0], np.float32)
            acc += ctx.mlp_train_step_dev(_View(dX, b * batch * FEAT * 4), _View(dT, b * batch * 16), batch, hyper)
            step += 1
        losses.append(acc / nb)
        log(f"[train] epoch {ep + 1}/{epochs} loss {losses[-1]:.5f} ({time.time() - t0:.2f}s)")
    dX.free()
    dT.free()
    return losses


def compare(plain: np.ndarray, warm: np.ndarray, tight: np.ndarray, scale: np.ndarray, band: float = 1e-6,
            rel: float = 1e-5, floor: float = 1e-6):
    """Agreement of warm-started vs plain queries on one workload.

    Booleans (intersecting or not) are compared outside the touching band.
    Accuracy is judged against a tighter solve (GJK/EPA tolerance 1e-10):
    the reference's GJK distance stops at tolerance 1e-6 (absolute) or when
    an iteration improves by less than that, so two start directions can end
    a few 1e-6 apart; the warm start must be as accurate as the plain
    solver, i.e. within max(1e-5 |d|, 1e-6) of the tight distance on at
    least as many pairs."""
    dp, dw, dx = plain["distance"], warm["distance"], tight["distance"]
    touching = np.abs(dp) <= band * scale
    sign_mism = int(np.sum(((dp < 0) != (dw < 0)) & ~touching))
    ok_mism = int(np.sum(plain["ok"] != warm["ok"]))
    ok = (plain["ok"] == 1) & (warm["ok"] == 1) & (tight["ok"] == 1)
    bound = np.maximum(rel * np.abs(dx[ok]), floor)
    ep, ew = np.abs(dp[ok] - dx[ok]), np.abs(dw[ok] - dx[ok])
    return dict(sign_mismatch_outside_band=sign_mism, ok_mismatch=ok_mism, touching_band=int(touching.sum()),
                pairs_checked=int(ok.sum()),
                plain_frac_outside_tol=float(np.mean(ep > bound)) if ok.any() else 0.0,
                warm_frac_outside_tol=float(np.mean(ew > bound)) if ok.any() else 0.0,
                plain_max_abs_err=float(ep.max()) if ok.any() else 0.0,
                warm_max_abs_err=float(ew.max()) if ok.any() else 0.0,
                warm_vs_plain_max_abs_diff=float(np.abs(dp[ok] - dw[ok]).max()) if ok.any() else 0.0)


def evaluate(ctx: Context, n: int, seed: int, steps: int = 5, log=print, keep=False):
    """Plain vs warm-started queries on n held-out pairs.  keep: also return
    the result records, the guesses and the workload."""
    ps = PairSet(ctx, n, seed)
    # the tighter solve that accuracy is judged against
    ctx.convex_query_dev(ps.pa, ps.pb, ps.qa, ps.qb, n,
                         request(MODE_SIGNED_DISTANCE, gjk_tolerance=1e-10, epa_tolerance=1e-10), ps.out)
    tight = ps.results()
    ctx.set_timing(True)
    res = {}
    for warm in (False, True):
        ps.query(warm)
        ctx.sync()
        t0 = time.perf_counter()
        kern = []
        for _ in range(steps):
            ps.query(warm)
            kern.append(ctx.last_timing().copy())
        ctx.sync()
        dt = (time.perf_counter() - t0) / steps
        out = ps.results()
        k = np.array(kern).mean(0)
        sep = out["distance"] > 0
        res["warm" if warm else "plain"] = dict(
            ms_per_step=1e3 * dt, pairs_per_s=n / dt, out=out, kernel_ms=dict(order=float(k[0]), gjk=float(k[1]),
                                                                           epa=float(k[2]), mlp=float(k[5])),
            gjk_iters_mean=float(out["gjk_iters"].mean()),
            gjk_iters_separated=float(out["gjk_iters"][sep].mean()) if sep.any() else 0.0,
            gjk_iters_intersecting=float(out["gjk_iters"][~sep].mean()) if (~sep).any() else 0.0,
            epa_iters_mean=float(out["epa_iters"][out["epa_status"] >= 0].mean()) if (out["epa_status"] >= 0).any()
            else 0.0)
    r = W.bounding_radius(ps.h["types"], ps.h["params"], ps.h["vert_cnt"], ps.h["verts"], ps.h["vert_off"])
    scale = r[ps.h["pair_a"]] + r[ps.h["pair_b"]]
    agree = compare(res["plain"]["out"], res["warm"]["out"], tight, scale)
    kept = None
    if keep:
        g = np.zeros((n, 3), np.float32)
        ps.guess.download(g)
        kept = dict(plain=res["plain"]["out"], warm=res["warm"]["out"], tight=tight, guess=g, h=ps.h)
    ps.free()
    for k in ("plain", "warm"):
        res[k].pop("out")
    res["agreement"] = agree
    res["speedup"] = res["warm"]["pairs_per_s"] / res["plain"]["pairs_per_s"]
    log(f"[eval] {json.dumps(res)}")
    if keep:
        res["kept"] = kept
    return res


def save_params(path: str, params: np.ndarray, meta: dict):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, params=params.astype(np.float32), meta=np.frombuffer(json.dumps(meta).encode(), np.uint8))


def load_params(path: str = DEFAULT_PARAMS):
    z = np.load(path, allow_pickle=False)
    return z["params"], json.loads(bytes(z["meta"]).decode())


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--pairs", type=int, default=10_000_000)
    t.add_argument("--seed", type=int, default=4)
    t.add_argument("--epochs", type=int, default=12)
    t.add_argument("--batch", type=int, default=65536)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--init-seed", type=int, default=7)
    t.add_argument("--out", default=DEFAULT_PARAMS)
    t.add_argument("--eval-pairs", type=int, default=1 << 20)
    t.add_argument("--eval-seed", type=int, default=5)
    e = sub.add_parser("eval")
    e.add_argument("--pairs", type=int, default=1 << 20)
    e.add_argument("--seed", type=int, default=5)
    e.add_argument("--params", default=DEFAULT_PARAMS)
    a = ap.parse_args(argv)
    log = lambda s: print(s, file=sys.stderr, flush=True)  # noqa: E731
    ctx = Context(0)
    if a.cmd == "train":
        t0 = time.time()
        X, T = make_dataset(ctx, a.pairs, a.seed, log=log)
        t_data = time.time() - t0
        ctx.mlp_init(a.init_seed)
        t0 = time.time()
        losses = train(ctx, X, T, a.epochs, a.batch, a.lr, seed=a.init_seed, log=log)
        t_train = time.time() - t0
        params = ctx.mlp_get_params()
        meta = dict(pairs=a.pairs, seed=a.seed, epochs=a.epochs, batch=a.batch, lr=a.lr, init_seed=a.init_seed,
                    losses=losses, label_seconds=t_data, train_seconds=t_train)
        if a.eval_pairs:
            meta["eval"] = evaluate(ctx, a.eval_pairs, a.eval_seed, log=log)
        save_params(a.out, params, meta)
        print(json.dumps(meta))
    else:
        params, meta = load_params(a.params)
        ctx.mlp_set_params(params)
        print(json.dumps(evaluate(ctx, a.pairs, a.seed, log=log)))
    ctx.close()


if __name__ == "__main__":
    main()
