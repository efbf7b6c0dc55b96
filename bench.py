"""Benchmark: batched convex-hull GJK/EPA pair queries per second on one MI355X.

Workload (default): 1,048,576 pairs of convex hulls of N points on random
ellipsoids (axes U[0.05, 0.4] m, N log-uniform in [8, 512]), every hull
built on the device from its own points (mfcl_upload_convex_points) with
its faces, so hulls above 32 vertices find support points by the vertex
walk as fcl::Convex does; centres U[0, 2]^3 m, uniform rotations;
GJKSolver::shapeSignedDistance semantics (GJK distance, EPA depth when
intersecting).  Inputs are resident in HBM before the timed region; one
step = one batched query over all pairs.

Prints one JSON line.  Also reports the roofline of the GJK stage (HBM
accounting per pair: each FP32 hull vertex read once, 12 (N_a + N_b) B, plus
2 x 28 B of poses, 2 x 8 B of shape offset and count and a 48 B result)
from HIP events on the query stream, and the oracle's FP64 multithreaded CPU
throughput on a bounded sample of the same workload.
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


def host_cpu_info():
    """Threads the CPU baselines use: every CPU this process may run on
    (sched_getaffinity), capped by the cgroup CPU quota when one is set (a
    quota of q CPUs gives no more throughput to more than q threads).  The
    host's hardware thread count is reported beside it."""
    aff = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
        if q != "max":
            quota = float(q) / float(per)
    except (OSError, ValueError):
        pass
    threads = aff if quota is None else max(1, min(aff, int(np.ceil(quota))))
    return {"threads": threads, "affinity_cpus": aff, "cgroup_cpu_quota": quota,
            "host_hardware_threads": os.cpu_count()}


def timed_median(fn, n_units, target_s=10.0, runs=5):
    """Warm-up call, then `runs` timed runs of k calls each (k sized so the
    runs total about target_s); returns the median rate in units/s, k and
    the total seconds.  fn() returns its own wall seconds."""
    first = fn()
    k = max(1, int(np.ceil(target_s / runs / max(first, 1e-6))))
    rates, total = [], 0.0
    for _ in range(runs):
        secs = sum(fn() for _ in range(k))
        total += secs
        rates.append(k * n_units / secs)
    return float(np.median(rates)), k, total


def cpu_baseline(h, geom, n_sample, threads, target_cpu_s=10.0):
    """Oracle FP64 shapeSignedDistance over the first n_sample pairs, on
    qhull's hulls of their point sets (geom: vertices and triangles, the
    vertex-adjacency walk above 32 vertices, as fcl::Convex): one warm-up
    pass, then the median of 5 timed runs totalling about target_cpu_s of
    wall time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    m = 2 * n_sample
    v, vo, vc, t, to, tc = geom
    S = pyoracle.ShapeSet(h["types"][:m], h["params"][:m], v, vo, vc, t, to, tc, threads=threads)
    ta = pyoracle.pose12(h["pose_a"][:n_sample])
    tb = pyoracle.pose12(h["pose_b"][:n_sample])
    pa, pb = h["pair_a"][:n_sample], h["pair_b"][:n_sample]
    ref, _ = S.signed_distance(pa, pb, ta, tb, threads=threads)
    rate, k, total = timed_median(lambda: S.signed_distance(pa, pb, ta, tb, threads=threads)[1], n_sample,
                                  target_cpu_s)
    return rate, ref, dict(sample=f"first {n_sample} pairs of the same workload; median of 5 runs of {k} passes "
                                  f"each after a warm-up ({5 * k * n_sample} pair queries, {total:.2f}s wall on "
                                  f"{threads} threads)")


def measured_traffic(tuned):
    """HBM bytes per launch of the lane GJK kernel on this workload, from the
    rocprofv3 passes summarised in profiles/ (read bytes from the L2's read
    requests by size, written bytes from WRITE_SIZE); None when the group
    kernel was chosen (not profiled)."""
    if tuned.get("gjk_group", 1) != 1:
        return None
    path = os.path.join(ROOT, "profiles", "lane_kernel_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)["traffic_bytes_per_launch"]


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
    ap.add_argument("--no-warm", action="store_true", help="skip the learned warm-start measurement")
    ap.add_argument("--no-mesh", action="store_true", help="skip the mesh-mesh measurement")
    ap.add_argument("--no-arm", action="store_true", help="skip the robot-arm configuration measurement")
    ap.add_argument("--arm-configs", type=int, default=65536)
    ap.add_argument("--arm-cpu-configs", type=int, default=4096)
    ap.add_argument("--mesh-pool", type=int, default=256)
    ap.add_argument("--mesh-cpu-pairs", type=int, default=512)
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
    h = W.hull_point_pairs(args.pairs, args.seed)
    n = args.pairs
    log(f"[gen] {n} pairs, {len(h['verts'])} points in {time.time() - t0:.1f}s")
    cpu = host_cpu_info()
    if args.cpu_threads:
        cpu["threads"] = args.cpu_threads
    geom = None
    if not args.no_cpu:
        # qhull's hulls of the CPU sample's point sets, in forked workers
        # before anything touches the GPU
        t0 = time.time()
        m = min(args.cpu_pairs, n)
        geom = W.hulls_of_sets(h["verts"], h["vert_off"], h["vert_cnt"], 2 * m, workers=min(16, cpu["threads"]))
        log(f"[cpu] qhull hulls of the {2 * m} sample shapes in {time.time() - t0:.1f}s")
    ctx = mfcl.Context(0)
    t0 = time.time()
    # every shape is the hull of its own point set, built on the device
    hull_cnt = ctx.upload_convex_points(h["verts"], h["vert_off"], h["vert_cnt"])
    log(f"[hulls] {n * 2} hulls, {int(hull_cnt.sum())} vertices ({int((hull_cnt < h['vert_cnt']).sum())} sets "
        f"with points inside their hull) built on the device in {time.time() - t0:.2f}s")
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

        t_group = autotune.tune(ctx, autotune.GJK_GROUP, [1, 32], run, reps=2, log=log)
        t_epa = autotune.tune(ctx, autotune.EPA_THREADS, [32768, 131072], run, reps=2, log=log)
        t_lds = autotune.tune(ctx, autotune.EPA_LDS_FACES, [48, 64], run, reps=3, log=log)
        tuned = {"gjk_group": min(t_group, key=t_group.get), "epa_threads": min(t_epa, key=t_epa.get),
                 "epa_lds_faces": min(t_lds, key=t_lds.get)}
        if tuned["gjk_group"] == 1:
            t_refill = autotune.tune(ctx, autotune.LANE_REFILL, [40, 48], run, reps=3, log=log)
            tuned["lane_refill"] = min(t_refill, key=t_refill.get)
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
    nv = hull_cnt.astype(np.int64)
    # GJK stage, per launch: every FP32 hull vertex of both shapes read once
    # (12 B each) and per pair 56 B of poses, 16 B of shape offsets and
    # counts and a 48 B result
    pair_verts = float(nv[h["pair_a"]].sum() + nv[h["pair_b"]].sum())
    alg_bytes = 12.0 * pair_verts + 120.0 * n
    # secondary: the bytes of the vertex records the support searches
    # actually request (16 B per vertex read, counted by the library, EPA's
    # share excluded, plus a 4 B neighbour id) and 184 B of pair I/O
    gjk_vertices = float(counters[1]) - float(counters[4])
    support_bytes = 20.0 * gjk_vertices + 184.0 * n
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
            "workload": f"{n} convex-hull pairs, every hull built on the device from its own N points "
                        "(N log-uniform [8,512], on ellipsoids with axes U[0.05,0.4] m), with hull faces "
                        "(support by the vertex walk above 32 vertices, as fcl::Convex), "
                        "centres U[0,2]^3, GJKSolver::shapeSignedDistance (GJK distance + EPA)",
            "pairs": n,
            "mean_vertices_per_hull": float(nv.mean()),
            "intersecting_fraction": float((out["distance"] < 0).mean()),
            "launch": tuned,
        },
        "roofline": {
            "kernel": "laneKernelRSd (GJK stage, lanes refilled)" if tuned.get("gjk_group", 1) == 1
                      else "narrowKernel (GJK stage)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(tuned),
            "algorithmic_bytes_per_launch": alg_bytes,
            "support_read_bytes_per_launch": support_bytes,
            "kernel_ms": {"order": float(kern[:, 0].mean()), "gjk": float(kern[:, 1].mean()),
                          "epa": float(kern[:, 2].mean())},
        },
    }
    params_path = os.path.join(ROOT, "mind-fcl_amd", "mfcl", "data", "warmstart_mlp.npz")
    if not args.no_warm and os.path.exists(params_path):
        result["warm_start"] = bench_warm(ctx, args, h, dpa, dpb, dqa, dqb, dout, n, out, value, params_path)
    if not args.no_cpu:
        threads = cpu["threads"]
        rate, ref, info = cpu_baseline(h, geom, min(args.cpu_pairs, n), threads)
        m = min(args.cpu_pairs, n)
        sign_mism = int(np.sum((ref["distance"] < 0) != (out["distance"][:m] < 0)))
        not_exact = int(np.sum((ref["distance"] != out["distance"][:m]) | (ref["ok"] != out["ok"][:m])))
        result["cpu_baseline"] = {"value": rate, "unit": "pairs/s", "cores": threads, "kind": "port",
                                  "host": cpu, "sample": info["sample"], "sample_sign_mismatches_vs_gpu": sign_mism,
                                  "sample_pairs_not_bit_identical_to_gpu": not_exact}
        result["speedup_vs_cpu"] = value / rate
    del h
    if not args.no_mesh:
        result["mesh"] = bench_mesh(ctx, args, threads=cpu["threads"])
    if not args.no_arm:
        result["arm"] = bench_arm(ctx, args, threads=cpu["threads"])
    ctx.close()
    print(json.dumps(result), flush=True)


def bench_warm(ctx, args, h, dpa, dpb, dqa, dqb, dout, n, out_plain, value_plain, params_path):
    """The same workload with the learned warm start: MLP guess (features +
    BF16 MFMA forward) then GJK/EPA from that guess; compared with the plain
    results and with a tighter solve."""
    import mfcl
    from mfcl import warmstart as WS

    params, meta = WS.load_params(params_path)
    ctx.mlp_set_params(params)
    dguess = ctx.alloc(n * 12)
    req_w = mfcl.request(mfcl.MODE_SIGNED_DISTANCE, use_guess=1)

    def run():
        ctx.mlp_guess_dev(dpa, dpb, dqa, dqb, n, dguess)
        ctx.convex_query_dev(dpa, dpb, dqa, dqb, n, req_w, dout, guess=dguess)

    for _ in range(args.warmup):
        run()
    ctx.sync()
    kern = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run()
        kern.append(ctx.last_timing().copy())
    ctx.sync()
    el = time.perf_counter() - t0
    kern = [float(k) for k in np.array(kern).mean(0)]
    out_w = np.zeros(n, mfcl.CONVEX_RESULT_DTYPE)
    dout.download(out_w)
    ctx.convex_query_dev(dpa, dpb, dqa, dqb, n, mfcl.request(mfcl.MODE_SIGNED_DISTANCE, gjk_tolerance=1e-10,
                                                             epa_tolerance=1e-10), dout)
    tight = np.zeros(n, mfcl.CONVEX_RESULT_DTYPE)
    dout.download(tight)
    from mfcl import workloads as W

    r = W.bounding_radius(h["types"], h["params"], h["vert_cnt"], h["verts"], h["vert_off"])
    agree = WS.compare(out_plain, out_w, tight, r[h["pair_a"]] + r[h["pair_b"]])
    v = n * args.steps / el
    # MLP roofline: 2 * (39*128 + 128*128*2 + 128*4) = 76,544 FLOP per pair on the BF16 MFMA peak
    mlp_flop = 76544.0 * n
    res = {"value": v, "unit": "pairs/s", "ms_per_step": 1e3 * el / args.steps, "speedup_vs_plain": v / value_plain,
           "kernel_ms": {"order": float(kern[0]), "gjk": float(kern[1]), "epa": float(kern[2]), "mlp": float(kern[5])},
           "gjk_iters_mean": {"plain": float(out_plain["gjk_iters"].mean()), "warm": float(out_w["gjk_iters"].mean())},
           "agreement": agree,
           "mlp_roofline": {"bound": "mfma", "achieved": mlp_flop / max(kern[5], 1e-9) / 1e9, "peak": 2500.0,
                            "unit": "TFLOP/s", "frac": mlp_flop / max(kern[5], 1e-9) / 1e9 / 2500.0,
                            "note": "features + 4 products + normalisation, per launch over all pairs"},
           "trained_on": {k: meta[k] for k in ("pairs", "seed", "epochs") if k in meta}}
    log(f"[warm] {json.dumps(res)}")
    dguess.free()
    return res


def bench_arm(ctx, args, threads, margin=0.05, cap=64):
    """Robot-arm configurations: 65,536 joint vectors of a 7-link chain of
    convex links against 512 oriented boxes and a 256 x 256 fBm heightmap
    (1 cm pixels) -- kinematics, AABB-tree broadphase, signed distance of
    every box candidate pair within the margin, heightmap candidate pixels
    per link with a boolean narrowphase, per-configuration collision flag
    and minimum distance (to the boxes)."""
    import mfcl
    from mfcl import workloads as W

    scene = W.arm_scene(args.arm_configs, seed=6, heightmap=True)
    L, nb = scene["n_links"], scene["n_boxes"]
    ctx.upload_shapes(scene["types"], scene["params"], scene["vert_off"], scene["vert_cnt"], scene["verts"],
                      faces=mfcl.faces_from_triangles(scene["tris"], scene["tri_off"], scene["tri_cnt"]))
    mn, mx = W.box_aabbs(scene["box_pose"], scene["box_half"])
    ctx.aabbtree_build(mn, mx, np.arange(nb))
    ctx.upload_heightmap(scene["hm_heights"], scene["hm_res"], scene["hm_res"])
    n = len(scene["q"])
    hm = dict(box_shape0=scene["hm_box_shape0"], pose7=scene["hm_pose7"], cap=4096)
    arm = mfcl.make_arm(np.arange(L), scene["link_len"], scene["link_radius"], L, nb, broadphase_cap=cap,
                        max_pairs=16 * n * L, heightmap=hm)
    dq, dbox = ctx.to_device(scene["q"]), ctx.to_device(scene["box_pose"])
    dany, dmin = ctx.alloc(4 * n), ctx.alloc(4 * n)
    ctx.reserve_arm(arm, n)
    for _ in range(max(1, args.warmup)):
        npairs = ctx.arm_query_dev(arm, dq, n, margin, dbox, dany, dmin)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        npairs = ctx.arm_query_dev(arm, dq, n, margin, dbox, dany, dmin)
    ctx.sync()
    el = time.perf_counter() - t0
    any_hit, min_d = np.zeros(n, np.int32), np.zeros(n, np.float32)
    dany.download(any_hit)
    dmin.download(min_d)
    value = n * args.steps / el
    res = {"value": value, "unit": "configurations/s", "ms_per_step": 1e3 * el / args.steps,
           "pairs_per_step": npairs, "heightmap_pairs_per_step": ctx.arm_last_heightmap_pairs(),
           "colliding_fraction": float(np.mean(any_hit & 1)),
           "failed_fraction": float(np.mean((any_hit & 2) > 0)),
           "workload": f"{n} configurations of a {L}-link arm (hull links of 64..128 vertices with faces) vs {nb} "
                       "oriented boxes in a 2 m cube 0.5 m in front of the base (candidates within 0.05 m) and a "
                       "256 x 256 fBm heightmap of 1 cm pixels (4 octaves, 0.1 m) 0.45 m below the base"}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import arm_ref
        import pyoracle

        m = min(args.arm_cpu_configs, n)
        sub = dict(scene)
        S = arm_ref.shape_set(scene)
        tree = pyoracle.AABBTree(mn, mx, np.arange(nb))
        def cpu_pass():
            pose = arm_ref.fk(scene["q"][:m], scene["link_len"])
            fl, md, _ = arm_ref.pipeline(sub, scene["q"][:m], margin, cap, S, tree, threads=threads, pose=pose)
            hit, _, _ = arm_ref.hm_collide(sub, S, pose, threads=threads)
            return fl | hit.reshape(m, L).any(1).astype(np.int32), md

        last = {}

        def timed_pass():
            t0 = time.perf_counter()
            last["r"] = cpu_pass()
            return time.perf_counter() - t0

        rate, k, total = timed_median(timed_pass, m, 10.0)
        flags, mind = last["r"]
        band = np.abs(mind) <= 1e-5
        res["cpu_baseline"] = {"value": rate, "unit": "configurations/s", "cores": threads, "kind": "port",
                               "sample": f"first {m} configurations; median of 5 runs of {k} passes each after a "
                                         f"warm-up ({total:.2f}s wall)",
                               "sample_hit_mismatches_vs_gpu": int(np.sum(((flags & 1) != (any_hit[:m] & 1)) & ~band))}
        res["speedup_vs_cpu"] = value / rate
    log(f"[arm] {json.dumps(res)}")
    for d in (dq, dbox, dany, dmin):
        d.free()
    return res


def bench_mesh(ctx, args, threads):
    """Mesh-mesh OBB-tree collision: 4096 pairs per launch over a pool of
    icosphere meshes (1e3..1e5 triangles, log-uniform), CollisionRequest(16)."""
    import mfcl
    from mfcl import workloads as W

    t0 = time.time()
    pool = W.mesh_pool(args.mesh_pool, 3)
    pairs = W.mesh_pairs(pool, 4096, 30)
    ctx.upload_meshes(pool["verts"], pool["vert_off"], pool["vert_cnt"], pool["tris"], pool["tri_off"],
                      pool["tri_cnt"])
    log(f"[mesh] pool {args.mesh_pool} meshes, {int(pool['tri_cnt'].sum())} triangles, built in "
        f"{time.time() - t0:.1f}s")
    n = len(pairs["pair_a"])
    dpa, dpb = ctx.to_device(pairs["pair_a"]), ctx.to_device(pairs["pair_b"])
    dqa, dqb = ctx.to_device(pairs["pose_a"]), ctx.to_device(pairs["pose_b"])
    dres = ctx.alloc(n * mfcl.MESH_RESULT_DTYPE.itemsize)
    dcon = ctx.alloc(n * 16 * 8)
    tuned = {}
    if not args.no_tune:
        from mfcl import autotune

        def run():
            ctx.mesh_collide_dev(dpa, dpb, dqa, dqb, n, 16, dres, dcon)

        t_bpc = autotune.tune(ctx, autotune.MESH_BLOCKS_PER_CU, [2, 3], run, reps=5, log=log)
        tuned = {"mesh_blocks_per_cu": min(t_bpc, key=t_bpc.get)}
    for _ in range(2):
        ctx.mesh_collide_dev(dpa, dpb, dqa, dqb, n, 16, dres, dcon)
    ctx.sync()
    km = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.mesh_collide_dev(dpa, dpb, dqa, dqb, n, 16, dres, dcon)
        km.append(float(ctx.last_timing()[3]))
    ctx.sync()
    el = time.perf_counter() - t0
    res = np.zeros(n, mfcl.MESH_RESULT_DTYPE)
    dres.download(res)
    bv, lf = res["bv_tests"].astype(np.float64).sum(), res["leaf_tests"].astype(np.float64).sum()
    alg = 2 * 64 * bv + 2 * 36 * lf + 64 * n
    kms = float(np.mean(km))
    out = {"value": n * args.steps / el, "unit": "pairs/s", "pairs_per_launch": n, "kernel_ms": kms, "launch": tuned,
           "hit_fraction": float((res["hit"] == 1).mean()), "overflow_pairs": int((res["hit"] < 0).sum()),
           "bv_tests_mean": bv / n, "bv_tests_max": int(res["bv_tests"].max()), "leaf_tests_mean": lf / n,
           "roofline": {"bound": "hbm", "achieved": alg / (kms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": alg / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS, "algorithmic_bytes_per_launch": alg},
           "workload": f"{n} mesh pairs over a pool of {args.mesh_pool} noisy icospheres "
                       "(1e3..1e5 triangles, radius U[0.1,0.5]), ~40% overlapping, first 16 contacts"}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle

        M = pyoracle.MeshSet(pool["verts"], pool["vert_off"], pool["vert_cnt"], pool["tris"], pool["tri_off"],
                             pool["tri_cnt"], threads=threads)
        m = min(args.mesh_cpu_pairs, n)
        ta, tb = pyoracle.pose12(pairs["pose_a"][:m]), pyoracle.pose12(pairs["pose_b"][:m])
        ref, _ = M.collide(pairs["pair_a"][:m], pairs["pair_b"][:m], ta, tb, 16, threads=threads)
        rate, k, tot = timed_median(lambda: M.collide(pairs["pair_a"][:m], pairs["pair_b"][:m], ta, tb, 16,
                                                      threads=threads)[1], m, 10.0)
        out["cpu_baseline"] = {"value": rate, "unit": "pairs/s", "cores": threads, "kind": "port",
                               "sample": f"first {m} pairs; median of 5 runs of {k} passes each after a warm-up "
                                         f"({tot:.2f}s wall)",
                               "sample_bool_mismatches_vs_gpu": int(np.sum(ref["hit"] != res["hit"][:m]))}
        out["speedup_vs_cpu"] = out["value"] / rate
    log(f"[mesh] {json.dumps(out)}")
    for a in (dpa, dpb, dqa, dqb, dres, dcon):
        a.free()
    return out


if __name__ == "__main__":
    main()
