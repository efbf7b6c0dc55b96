"""Pins of the oracle's directed MPR penetration (collisionPenetration in
its DirectedPenetration mode; MPR::RunDirectedPenetration, include/fcl/
cvx_collide/mpr.hpp:496-704) against what the geometry fixes:

- two spheres: shape b moved by t along d separates at the larger root of
  |c + t d| = r_a + r_b (closed form);
- two axis-aligned boxes: the overlap along d ends at the first face,
  min_k (sign(d_k) h_k - c_k) / d_k (exact for polytopes);
- two random convex hulls: the largest t with t d in conv{a - b} (a linear
  program over the pairwise vertex differences);
- a separated pair reports no contact and runs no search.
"""
import numpy as np
import pytest
from scipy.optimize import linprog

from mfcl import workloads as W

BOX, SPHERE, CONVEX = 0, 1, 100
IDENT = np.array([0, 0, 0, 1, 0, 0, 0], np.float32)  # (t, w x y z)


def _pose(t, q=(1, 0, 0, 0)):
    p = np.zeros(7, np.float32)
    p[:3] = t
    p[3:] = q
    return p


def _unit(rng, n):
    u = rng.normal(size=(n, 3))
    return (u / np.linalg.norm(u, axis=1, keepdims=True)).astype(np.float32)


def test_spheres_closed_form(oracle):
    rng = np.random.default_rng(60)
    n = 400
    r = rng.uniform(0.1, 0.5, (n, 2))
    types = np.array([SPHERE, SPHERE] * n, np.int32)
    params = np.zeros((2 * n, 3))
    params[:, 0] = r.reshape(-1)
    S = oracle.ShapeSet(types, params)
    ca = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    off = _unit(rng, n) * (rng.uniform(0.05, 0.95, (n, 1)) * r.sum(1, keepdims=True)).astype(np.float32)
    cb = (ca + off).astype(np.float32)
    d = _unit(rng, n)
    pa = np.stack([_pose(c) for c in ca])
    pb = np.stack([_pose(c) for c in cb])
    out, _ = S.directed_penetration(np.arange(0, 2 * n, 2), np.arange(1, 2 * n, 2), oracle.pose12(pa),
                                    oracle.pose12(pb), d.astype(np.float64), threads=8)
    assert np.all(out["hit"] == 1) and np.all(out["epa_status"] == 0)
    c = cb.astype(np.float64) - ca.astype(np.float64)
    dd = d.astype(np.float64)
    cd = np.sum(c * dd, 1)
    t = -cd + np.sqrt(cd ** 2 - np.sum(c * c, 1) + r.sum(1) ** 2)
    err = np.abs(out["depth"] - t)
    print(f"\n[directed spheres] max |depth - closed form| {err.max():.2e}")
    assert err.max() < 1e-5
    assert np.allclose(out["normal"], dd)


def test_boxes_exact(oracle):
    rng = np.random.default_rng(61)
    n = 400
    ha = rng.uniform(0.1, 0.5, (n, 3)).astype(np.float32)
    hb = rng.uniform(0.1, 0.5, (n, 3)).astype(np.float32)
    types = np.array([BOX, BOX] * n, np.int32)
    params = np.zeros((2 * n, 3))
    params[0::2] = 2 * ha
    params[1::2] = 2 * hb
    S = oracle.ShapeSet(types, params)
    ca = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    cb = (ca + rng.uniform(-0.9, 0.9, (n, 3)) * (ha + hb)).astype(np.float32)
    d = _unit(rng, n)
    d[:50] = [1, 0, 0]  # an axis direction, one face only
    out, _ = S.directed_penetration(np.arange(0, 2 * n, 2), np.arange(1, 2 * n, 2),
                                    oracle.pose12(np.stack([_pose(x) for x in ca])),
                                    oracle.pose12(np.stack([_pose(x) for x in cb])), d.astype(np.float64), threads=8)
    assert np.all(out["hit"] == 1) and np.all(out["epa_status"] == 0)
    c = cb.astype(np.float64) - ca.astype(np.float64)
    h = ha.astype(np.float64) + hb.astype(np.float64)
    dd = d.astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        lim = np.where(dd > 0, (h - c) / dd, np.where(dd < 0, (-h - c) / dd, np.inf))
    t = lim.min(1)
    err = np.abs(out["depth"] - t)
    print(f"\n[directed boxes] max |depth - exact| {err.max():.2e}")
    assert err.max() < 1e-9


def test_hulls_linear_program(oracle):
    rng = np.random.default_rng(62)
    n = 60
    v, off, cnt, _ = W.ellipsoid_hulls(rng, 2 * n, 8, 12)
    keep_v, keep_o, keep_c, tris, to, tc = [], [], [], [], [], []
    o = q = 0
    for i in range(2 * n):
        P = v[off[i]:off[i] + cnt[i]]
        k, t = W.hull_faces(P)
        keep_v.append(P[k]); keep_o.append(o); keep_c.append(len(k)); o += len(k)
        tris.append(t); to.append(q); tc.append(len(t)); q += len(t)
    V = np.concatenate(keep_v)
    S = oracle.ShapeSet(np.full(2 * n, CONVEX, np.int32), np.zeros((2 * n, 3)), V, keep_o, keep_c,
                        np.concatenate(tris), to, tc)
    poses = W.poses(rng, 2 * n, 0.0, 0.15)
    d = _unit(rng, n)
    ta, tb = oracle.pose12(poses[0::2]), oracle.pose12(poses[1::2])
    out, _ = S.directed_penetration(np.arange(0, 2 * n, 2), np.arange(1, 2 * n, 2), ta, tb, d.astype(np.float64),
                                    threads=8)
    checked = 0
    for i in range(n):
        if not out["hit"][i]:
            assert out["epa_status"][i] == -1
            continue
        assert out["epa_status"][i] == 0
        Ra, Rb = ta[i, :9].reshape(3, 3), tb[i, :9].reshape(3, 3)
        A = keep_v[2 * i].astype(np.float64) @ Ra.T + ta[i, 9:]
        B = keep_v[2 * i + 1].astype(np.float64) @ Rb.T + tb[i, 9:]
        D = (A[:, None, :] - B[None, :, :]).reshape(-1, 3)
        m = len(D)
        # variables: lambda (m), t; maximise t with sum(lambda p) - t d = 0, sum(lambda) = 1
        A_eq = np.zeros((4, m + 1))
        A_eq[:3, :m] = D.T
        A_eq[:3, m] = -d[i].astype(np.float64)
        A_eq[3, :m] = 1
        b_eq = np.array([0, 0, 0, 1.0])
        c = np.zeros(m + 1)
        c[m] = -1
        res = linprog(c, A_eq=A_eq, b_eq=b_eq, bounds=[(0, None)] * m + [(None, None)], method="highs")
        assert res.status == 0
        assert abs(out["depth"][i] - res.x[m]) < 1e-7, (i, out["depth"][i], res.x[m])
        checked += 1
    print(f"\n[directed hulls] {checked} colliding pairs equal to the LP optimum")
    assert checked > 20


def test_separated_pair(oracle):
    S = oracle.ShapeSet(np.array([SPHERE, BOX], np.int32), np.array([[0.2, 0, 0], [0.4, 0.4, 0.4]]))
    out, _ = S.directed_penetration([0], [1], oracle.pose12(_pose((0, 0, 0))[None]),
                                    oracle.pose12(_pose((1, 0, 0))[None]), np.array([[1.0, 0, 0]]))
    assert out["hit"][0] == 0 and out["n"][0] == 0 and out["epa_status"][0] == -1


# ---- incremental minimum penetration (mpr_incremental_penetration.hpp) ----

def test_incremental_spheres(oracle):
    """From any hint with a positive component along the centre offset, the
    refinement reaches the minimum penetration r_a + r_b - |c| along c."""
    rng = np.random.default_rng(63)
    n = 300
    r = rng.uniform(0.1, 0.5, (n, 2))
    types = np.array([SPHERE, SPHERE] * n, np.int32)
    params = np.zeros((2 * n, 3))
    params[:, 0] = r.reshape(-1)
    S = oracle.ShapeSet(types, params)
    ca = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    off = _unit(rng, n) * (rng.uniform(0.2, 0.9, (n, 1)) * r.sum(1, keepdims=True)).astype(np.float32)
    cb = (ca + off).astype(np.float32)
    c = cb.astype(np.float64) - ca.astype(np.float64)
    cu = c / np.linalg.norm(c, axis=1, keepdims=True)
    hint = cu + rng.normal(scale=0.5, size=(n, 3))
    hint /= np.linalg.norm(hint, axis=1, keepdims=True)
    out, _ = S.incremental_penetration(np.arange(0, 2 * n, 2), np.arange(1, 2 * n, 2),
                                       oracle.pose12(np.stack([_pose(x) for x in ca])),
                                       oracle.pose12(np.stack([_pose(x) for x in cb])), hint, threads=8)
    assert np.all(out["hit"] == 1) and np.all(np.isin(out["epa_status"], [0, 4]))
    t = r.sum(1) - np.linalg.norm(c, axis=1)
    err = np.abs(out["depth"] - t)
    ang = np.arccos(np.clip(np.sum(out["normal"] * cu, 1), -1, 1))
    print(f"\n[incremental spheres] max |depth - (r_a + r_b - |c|)| {err.max():.2e}, max angle {ang.max():.2e}")
    assert err.max() < 1e-4
    assert np.all(out["depth"] >= t - 1e-9)


def test_incremental_boxes(oracle):
    """Axis-aligned boxes: every result is at least the minimum penetration
    min_k (h_k - |c_k|) (no escape is shorter), and a hint near the best
    axis reaches it."""
    rng = np.random.default_rng(64)
    n = 400
    ha = rng.uniform(0.1, 0.5, (n, 3)).astype(np.float32)
    hb = rng.uniform(0.1, 0.5, (n, 3)).astype(np.float32)
    types = np.array([BOX, BOX] * n, np.int32)
    params = np.zeros((2 * n, 3))
    params[0::2] = 2 * ha
    params[1::2] = 2 * hb
    S = oracle.ShapeSet(types, params)
    ca = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    cb = (ca + rng.uniform(-0.9, 0.9, (n, 3)) * (ha + hb)).astype(np.float32)
    c = cb.astype(np.float64) - ca.astype(np.float64)
    h = ha.astype(np.float64) + hb.astype(np.float64)
    pen = h - np.abs(c)
    best = pen.argmin(1)
    tmin = pen.min(1)
    axis = np.zeros((n, 3))
    axis[np.arange(n), best] = np.sign(c[np.arange(n), best])
    near = axis + rng.normal(scale=0.05, size=(n, 3))
    near /= np.linalg.norm(near, axis=1, keepdims=True)
    anyd = _unit(rng, n).astype(np.float64)
    ta = oracle.pose12(np.stack([_pose(x) for x in ca]))
    tb = oracle.pose12(np.stack([_pose(x) for x in cb]))
    pa, pb = np.arange(0, 2 * n, 2), np.arange(1, 2 * n, 2)
    o_near, _ = S.incremental_penetration(pa, pb, ta, tb, near, threads=8)
    o_any, _ = S.incremental_penetration(pa, pb, ta, tb, anyd, threads=8)
    for o in (o_near, o_any):
        ok = np.isin(o["epa_status"], [0, 4])
        assert np.all(o["depth"][ok] >= tmin[ok] - 1e-9)
    ok = np.isin(o_near["epa_status"], [0, 4])
    err = np.abs(o_near["depth"] - tmin)[ok]
    print(f"\n[incremental boxes] near-axis hints: {int(ok.sum())}/{n} converged, max |depth - min| {err.max():.2e}, "
          f"exact on {int((err == 0).sum())}")
    # the refinement stops once a step gains less than the tolerance (1e-6)
    assert ok.sum() > 0.95 * n and err.max() < 2e-6 and (err == 0).mean() > 0.9


def test_incremental_hulls_match_epa(oracle):
    """Hull pairs, hinted with the EPA normal (the minimum-penetration
    direction, EPA pinned separately): the refinement keeps EPA's depth."""
    rng = np.random.default_rng(65)
    n = 80
    v, off, cnt, _ = W.ellipsoid_hulls(rng, 2 * n, 8, 40)
    keep_v, keep_o, keep_c, tris, to, tc = [], [], [], [], [], []
    o = q = 0
    for i in range(2 * n):
        P = v[off[i]:off[i] + cnt[i]]
        k, t = W.hull_faces(P)
        keep_v.append(P[k]); keep_o.append(o); keep_c.append(len(k)); o += len(k)
        tris.append(t); to.append(q); tc.append(len(t)); q += len(t)
    S = oracle.ShapeSet(np.full(2 * n, CONVEX, np.int32), np.zeros((2 * n, 3)), np.concatenate(keep_v), keep_o,
                        keep_c, np.concatenate(tris), to, tc)
    poses = W.poses(rng, 2 * n, 0.0, 0.1)
    ta, tb = oracle.pose12(poses[0::2]), oracle.pose12(poses[1::2])
    pa, pb = np.arange(0, 2 * n, 2), np.arange(1, 2 * n, 2)
    epa, _ = S.collide(pa, pb, ta, tb, 1, True, threads=8)
    hit = epa["hit"] == 1
    out, _ = S.incremental_penetration(pa, pb, ta, tb, epa["normal"], threads=8)
    ok = hit & np.isin(out["epa_status"], [0, 4])
    rel = np.abs(out["depth"] - epa["depth"])[ok] / np.maximum(epa["depth"][ok], 1e-3)
    print(f"\n[incremental hulls] {int(ok.sum())}/{int(hit.sum())} colliding pairs refined, "
          f"max relative |depth - EPA| {rel.max():.2e}")
    assert ok.sum() > 0.9 * hit.sum() and hit.sum() > 30
    assert rel.max() < 1e-4
