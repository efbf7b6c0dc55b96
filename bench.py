"""Benchmark: batched convex-hull GJK/EPA pair queries per second on one MI355X.

Workload (default): 1,048,576 pairs of convex hulls of N points on random
ellipsoids (axes U[0.05, 0.4] m, N log-uniform in [8, 512]), centres
U[0, 2]^3 m, uniform rotations; GJKSolver::shapeSignedDistance semantics
(GJK distance, EPA depth when intersecting).  Inputs are resident in HBM
before the timed region; one step = one batched query over all pairs.

Prints one JSON line.  Also reports the roofline of the dominant kernel
(GJK stage, HBM-bound accounting: 12 B per hull vertex of both shapes +
120 B of pose/index/result traffic per pair) from HIP events on the query
stream, and the oracle's FP64 multithreaded CPU throughput on a bounded
sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mind-fcl_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _faces_worker(args):
    from mfcl import workloads as W

    verts, off, cnt, idx = args
    out = []
    for i in idx:
        pts = verts[off[i]:off[i] + cnt[i]]
        out.append(W.hull_faces(pts))
    return out


def cpu_baseline(h, n_sample, threads, target_cpu_s=10.0):
    """Oracle FP64 shapeSignedDistance over the first n_sample pairs, with
    the reference's hull support search (vertex-adjacency walk above 32
    vertices); repeated until ~target_cpu_s of CPU work has been timed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    idx_hulls = np.arange(2 * n_sample)
    chunks = np.array_split(idx_hulls, max(1, threads))
    sub = (h["verts"][: int(h["vert_off"][2 * n_sample - 1] + h["vert_cnt"][2 * n_sample - 1])],
           h["vert_off"], h["vert_cnt"])
    t0 = time.time()
    if threads > 1:
        from concurrent.futures import ProcessPoolExecutor

        with ProcessPoolExecutor(max_workers=threads) as ex:
            parts = list(ex.map(_faces_worker, [(sub[0], sub[1], sub[2], c) for c in chunks]))
    else:
        parts = [_faces_worker((sub[0], sub[1], sub[2], c)) for c in chunks]
    faces = [f for p in parts for f in p]
    log(f"[cpu] qhull faces for {len(faces)} hulls in {time.time() - t0:.1f}s")
    vs, ts, vo, vc, to, tc = [], [], [], [], [], []
    o = q = 0
    for i, (keep, tri) in enumerate(faces):
        v = h["verts"][h["vert_off"][i]:h["vert_off"][i] + h["vert_cnt"][i]][keep]
        vs.append(v)
        ts.append(tri)
        vo.append(o)
        vc.append(len(v))
        to.append(q)
        tc.append(len(tri))
        o += len(v)
        q += len(tri)
    S = pyoracle.ShapeSet(h["types"][: 2 * n_sample], h["params"][: 2 * n_sample], np.concatenate(vs), vo, vc,
                          np.concatenate(ts), to, tc, threads=threads)
    ta = pyoracle.pose12(h["pose_a"][:n_sample])
    tb = pyoracle.pose12(h["pose_b"][:n_sample])
    pa, pb = h["pair_a"][:n_sample], h["pair_b"][:n_sample]
    ref, secs = S.signed_distance(pa, pb, ta, tb, threads=threads)  # warm-up pass
    reps = max(1, int(np.ceil(target_cpu_s / max(1e-6, secs * threads))))
    total = 0.0
    for _ in range(reps):
        _, s = S.signed_distance(pa, pb, ta, tb, threads=threads)
        total += s
    rate = reps * n_sample / total
    return rate, ref, dict(sample=f"first {n_sample} pairs of the same workload, {reps} timed passes "
                                  f"({reps * n_sample} pair queries, {total:.2f}s wall on {threads} threads)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--pairs", type=int, default=1 << 20)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--cpu-pairs", type=int, default=16384)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-shape autotuner")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a plain device-to-device copy of BYTES bytes 3 times (the known-size "
                         "reference kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        x = torch.ones(args.copy_calibration // 4, dtype=torch.float32, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": args.copy_calibration}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (pairs would shard by "
              "contiguous chunks with replicated geometry); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import mfcl
    from mfcl import workloads as W

    t0 = time.time()
    h = W.hull_pairs(args.pairs, args.seed)
    n = args.pairs
    log(f"[gen] {n} pairs, {len(h['verts'])} hull vertices in {time.time() - t0:.1f}s")
    ctx = mfcl.Context(0)
    t0 = time.time()
    ctx.upload_shapes(h["types"], h["params"], h["vert_off"], h["vert_cnt"], h["verts"])
    dpa, dpb = ctx.to_device(h["pair_a"]), ctx.to_device(h["pair_b"])
    dqa, dqb = ctx.to_device(h["pose_a"]), ctx.to_device(h["pose_b"])
    dout = ctx.alloc(n * mfcl.CONVEX_RESULT_DTYPE.itemsize)
    ctx.sync()
    log(f"[upload] {time.time() - t0:.1f}s")
    req = mfcl.request(mfcl.MODE_SIGNED_DISTANCE)
    tuned = {}
    if not args.no_tune:
        from mfcl import autotune

        def run():
            ctx.convex_query_dev(dpa, dpb, dqa, dqb, n, req, dout)

        t_group = autotune.tune(ctx, autotune.GJK_GROUP, [1, 8, 16], run, reps=2, log=log)
        t_epa = autotune.tune(ctx, autotune.EPA_THREADS, [8192, 32768, 131072], run, reps=2, log=log)
        tuned = {"gjk_group": min(t_group, key=t_group.get), "epa_threads": min(t_epa, key=t_epa.get)}
    ctx.set_timing(True)
    for _ in range(args.warmup):
        ctx.convex_query_dev(dpa, dpb, dqa, dqb, n, req, dout)
    ctx.sync()
    kern = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.convex_query_dev(dpa, dpb, dqa, dqb, n, req, dout)
        kern.append(ctx.last_timing()[:3].copy())
    ctx.sync()
    elapsed = time.perf_counter() - t0
    kern = np.array(kern)
    counters = ctx.last_counters()
    out = np.zeros(n, mfcl.CONVEX_RESULT_DTYPE)
    dout.download(out)
    value = n * args.steps / elapsed
    ms_step = 1e3 * elapsed / args.steps
    nv = h["vert_cnt"].astype(np.int64)
    alg_bytes = 12.0 * float(nv[h["pair_a"]].sum() + nv[h["pair_b"]].sum()) + 120.0 * n
    t_gjk = float(kern[:, 1].mean()) * 1e-3
    achieved = alg_bytes / t_gjk / 1e9
    log(f"[gpu] {ms_step:.3f} ms/step; kernels ms (order, gjk, epa) = {kern.mean(0)}; "
        f"epa pairs {int(counters[2])} (rerun at full capacity {int(counters[3])}); "
        f"vertices scanned {int(counters[1])}")
    ep = out["epa_status"] >= 0
    if ep.any():
        it = out["epa_iters"][ep].astype(np.int64)
        log(f"[gpu] EPA status counts {dict(zip(*np.unique(out['epa_status'][ep], return_counts=True)))}; "
            f"iterations mean {it.mean():.2f} p50 {np.percentile(it, 50):.0f} p99 {np.percentile(it, 99):.0f} "
            f"max {it.max()}")
    gi = out["gjk_iters"].astype(np.int64)
    log(f"[gpu] GJK iterations mean {gi.mean():.2f} p99 {np.percentile(gi, 99):.0f} max {gi.max()}")
    result = {
        "metric": "pair queries/s, 1 MI355X: GJK/EPA signed distance over convex-hull pairs",
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (f32 support scans)",
        "data": "synthetic",
        "config": {
            "workload": f"{n} convex-hull pairs, N log-uniform [8,512] points on ellipsoids (axes U[0.05,0.4] m), "
                        "centres U[0,2]^3, GJKSolver::shapeSignedDistance (GJK distance + EPA)",
            "pairs": n,
            "mean_vertices_per_hull": float(nv.mean()),
            "intersecting_fraction": float((out["distance"] < 0).mean()),
            "launch": tuned,
        },
        "roofline": {
            "kernel": "narrowKernel (GJK stage)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "algorithmic_bytes_per_launch": alg_bytes,
            "kernel_ms": {"order": float(kern[:, 0].mean()), "gjk": float(kern[:, 1].mean()),
                          "epa": float(kern[:, 2].mean())},
        },
    }
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        rate, ref, info = cpu_baseline(h, min(args.cpu_pairs, n), threads)
        m = min(args.cpu_pairs, n)
        sign_mism = int(np.sum((ref["distance"] < 0) != (out["distance"][:m] < 0)))
        result["cpu_baseline"] = {"value": rate, "unit": "pairs/s", "cores": threads, "kind": "port",
                                  "sample": info["sample"], "sample_sign_mismatches_vs_gpu": sign_mism}
        result["speedup_vs_cpu"] = value / rate
    ctx.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
