"""GPU: edge cases of the convex query against the oracle, field for field.

* ragged batch sizes (1, 63, 65, 257 pairs: partial wavefronts and blocks);
* the solver limits the reference exposes (GJKSolver, gjk_solver-inl.h:
  1124-1127): GJK iteration caps of 0, 1 and 2 (IterLimit paths and GJK
  giving up before separation is proven), EPA iteration caps of 1 and 3
  (IterationLimit), and EPA face capacities of 8 (MallocFailed inside the
  on-chip polytope) and 96 (pairs that outgrow the on-chip polytope rerun at
  full capacity);
* an empty batch.
Both GJK kernels, hulls with faces (vertex walk) and without.
"""
import numpy as np
import pytest

from mfcl import autotune
from mfcl import workloads as W
import mfcl

FIELDS = ("ok", "gjk_status", "epa_status", "gjk_iters", "epa_iters", "distance", "pa", "pb")


@pytest.fixture(scope="module")
def setup(oracle):
    h = W.hull_pairs(1024, 12, nmin=8, nmax=256, faces=True)
    S_walk = oracle.ShapeSet(h["types"], h["params"], h["verts"], h["vert_off"], h["vert_cnt"], h["tris"],
                             h["tri_off"], h["tri_cnt"])
    S_naive = oracle.ShapeSet(h["types"], h["params"], h["verts"], h["vert_off"], h["vert_cnt"])
    faces = mfcl.faces_from_triangles(h["tris"], h["tri_off"], h["tri_cnt"])
    return h, S_walk, S_naive, faces


def _compare(out, ref):
    return {f: int((out[f] != ref[f]).reshape(len(out), -1).any(1).sum()) for f in FIELDS}


def _run(ctx, h, sel, req, group):
    autotune.set_knob(ctx, autotune.GJK_GROUP, group)
    try:
        return ctx.convex_query(h["pair_a"][sel], h["pair_b"][sel], h["pose_a"][sel], h["pose_b"][sel], req)
    finally:
        autotune.set_knob(ctx, autotune.GJK_GROUP, 1)


@pytest.mark.gpu
@pytest.mark.parametrize("group", [1, 32])
@pytest.mark.parametrize("n", [1, 63, 65, 257])
def test_ragged_sizes(oracle, gpu_ctx, setup, n, group):
    h, S, _, faces = setup
    gpu_ctx.upload_shapes(h["types"], h["params"], h["vert_off"], h["vert_cnt"], h["verts"], faces=faces)
    sel = np.arange(100, 100 + n)
    ta, tb = oracle.pose12(h["pose_a"][sel]), oracle.pose12(h["pose_b"][sel])
    ref, _ = S.signed_distance(h["pair_a"][sel], h["pair_b"][sel], ta, tb)
    out = _run(gpu_ctx, h, sel, mfcl.request(mfcl.MODE_SIGNED_DISTANCE), group)
    diff = _compare(out, ref)
    print(f"\n[ragged n={n} group={group}] differing {diff}")
    assert all(v == 0 for v in diff.values())


LIMITS = [  # gjk_max_iterations, epa_max_faces, epa_max_iterations
    (0, 256, 255), (1, 256, 255), (2, 256, 255),
    (128, 256, 1), (128, 256, 3),
    (128, 8, 255), (128, 96, 255),
]


@pytest.mark.gpu
@pytest.mark.parametrize("walk", [True, False])
@pytest.mark.parametrize("group", [1, 32])
@pytest.mark.parametrize("gjk_it,epa_faces,epa_it", LIMITS)
def test_solver_limits(oracle, gpu_ctx, setup, gjk_it, epa_faces, epa_it, group, walk):
    h, S_walk, S_naive, faces = setup
    S = S_walk if walk else S_naive
    gpu_ctx.upload_shapes(h["types"], h["params"], h["vert_off"], h["vert_cnt"], h["verts"],
                          faces=faces if walk else None)
    sel = np.arange(len(h["pair_a"]))
    ta, tb = oracle.pose12(h["pose_a"]), oracle.pose12(h["pose_b"])
    sp = np.array([gjk_it, 1e-6, epa_faces, epa_it, 1e-6, 0], np.float64)
    ref, _ = S.signed_distance(h["pair_a"], h["pair_b"], ta, tb, solver=sp, threads=8)
    req = mfcl.request(mfcl.MODE_SIGNED_DISTANCE, gjk_max_iterations=gjk_it, epa_max_faces=epa_faces,
                       epa_max_iterations=epa_it)
    out = _run(gpu_ctx, h, sel, req, group)
    diff = _compare(out, ref)
    st = np.unique(out["epa_status"], return_counts=True)
    print(f"\n[limits gjk_it={gjk_it} epa_faces={epa_faces} epa_it={epa_it} group={group} walk={walk}] "
          f"epa statuses {dict(zip(st[0].tolist(), st[1].tolist()))} differing {diff}")
    assert all(v == 0 for v in diff.values())


@pytest.mark.gpu
@pytest.mark.parametrize("lds_faces,epa_threads", [(16, 131072), (24, 524288), (48, 131072)])
def test_epa_continuation_and_rerun(oracle, gpu_ctx, setup, lds_faces, epa_threads):
    """Pairs that outgrow the LDS polytope continue in a device-memory
    polytope of full capacity (the first 2048 wavefronts of the EPA grid) or
    rerun at full capacity (the others): results equal the oracle's at the
    request's capacity, field for field."""
    h, S, _, faces = setup
    gpu_ctx.upload_shapes(h["types"], h["params"], h["vert_off"], h["vert_cnt"], h["verts"], faces=faces)
    ta, tb = oracle.pose12(h["pose_a"]), oracle.pose12(h["pose_b"])
    ref, _ = S.signed_distance(h["pair_a"], h["pair_b"], ta, tb, threads=8)
    autotune.set_knob(gpu_ctx, autotune.EPA_LDS_FACES, lds_faces)
    autotune.set_knob(gpu_ctx, autotune.EPA_THREADS, epa_threads)
    try:
        out = gpu_ctx.convex_query(h["pair_a"], h["pair_b"], h["pose_a"], h["pose_b"],
                                   mfcl.request(mfcl.MODE_SIGNED_DISTANCE))
        c = gpu_ctx.last_counters()
    finally:
        autotune.set_knob(gpu_ctx, autotune.EPA_LDS_FACES, 64)
        autotune.set_knob(gpu_ctx, autotune.EPA_THREADS, 131072)
    diff = _compare(out, ref)
    print(f"\n[epa lds_faces={lds_faces} threads={epa_threads}] EPA pairs {int(c[2])} continued in device memory "
          f"{int(c[5])} rerun {int(c[3])} differing {diff}")
    assert all(v == 0 for v in diff.values())
    if lds_faces < 48:
        assert c[5] + c[3] > 0


@pytest.mark.gpu
def test_empty_batch(gpu_ctx, setup):
    h, _, _, faces = setup
    gpu_ctx.upload_shapes(h["types"], h["params"], h["vert_off"], h["vert_cnt"], h["verts"], faces=faces)
    out = gpu_ctx.convex_query(h["pair_a"][:0], h["pair_b"][:0], h["pose_a"][:0], h["pose_b"][:0],
                               mfcl.request(mfcl.MODE_SIGNED_DISTANCE))
    assert len(out) == 0
