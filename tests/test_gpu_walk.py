"""GPU: convex hulls with faces -- the reference's vertex-adjacency support
walk (convex-inl.h:202-290) -- against the oracle, which restates the same
walk.

Hulls of more than 32 vertices with watertight, connected faces walk; the
rest (32 or fewer vertices, faces withheld, or a face removed so the mesh is
not watertight) scan every vertex.  The device walk visits the same
neighbours in the same order with the same FP64 comparisons, so results are
compared field by field for exact equality, in signed-distance and
penetration modes, for both GJK kernels.
"""
import numpy as np
import pytest

from mfcl import autotune
from mfcl import workloads as W
import mfcl


def _setup(oracle, n_pairs, seed):
    h = W.hull_pairs(n_pairs, seed, nmin=8, nmax=512, faces=True)
    n = len(h["types"])
    tri_cnt = h["tri_cnt"].copy()
    rng = np.random.default_rng(seed + 100)
    kind = rng.integers(0, 4, size=n)  # 0, 1: faces; 2: faces minus one triangle; 3: no faces
    tri_cnt[kind == 2] -= 1
    tri_cnt[kind == 3] = 0
    S = oracle.ShapeSet(h["types"], h["params"], h["verts"], h["vert_off"], h["vert_cnt"], h["tris"],
                        h["tri_off"], tri_cnt)
    faces = mfcl.faces_from_triangles(h["tris"], h["tri_off"], tri_cnt)
    return h, S, faces, kind


def _run(ctx, h, req, group):
    autotune.set_knob(ctx, autotune.GJK_GROUP, group)
    try:
        return ctx.convex_query(h["pair_a"], h["pair_b"], h["pose_a"], h["pose_b"], req)
    finally:
        autotune.set_knob(ctx, autotune.GJK_GROUP, 1)


FIELDS_SD = ("ok", "gjk_status", "epa_status", "gjk_iters", "epa_iters", "distance", "pa", "pb")
FIELDS_PEN = ("hit", "ok", "gjk_status", "epa_status", "gjk_iters", "epa_iters", "depth", "normal", "pos")


@pytest.mark.gpu
@pytest.mark.parametrize("group", [1, 32])
def test_walk_signed_distance_bit_exact(oracle, gpu_ctx, group):
    h, S, faces, kind = _setup(oracle, 4096, 7)
    walks = np.array([S.walks(i) for i in range(len(h["types"]))])
    assert walks[(kind <= 1) & (h["vert_cnt"] > 32)].all() and not walks[kind >= 2].any()
    assert not walks[h["vert_cnt"] <= 32].any()
    ta, tb = oracle.pose12(h["pose_a"]), oracle.pose12(h["pose_b"])
    ref, _ = S.signed_distance(h["pair_a"], h["pair_b"], ta, tb, threads=8)
    gpu_ctx.upload_shapes(h["types"], h["params"], h["vert_off"], h["vert_cnt"], h["verts"], faces=faces)
    out = _run(gpu_ctx, h, mfcl.request(mfcl.MODE_SIGNED_DISTANCE), group)
    diff = {f: int((out[f] != ref[f]).reshape(len(out), -1).any(1).sum()) for f in FIELDS_SD}
    print(f"\n[walk sd, group {group}] pairs={len(out)} walking hulls={int(walks.sum())} "
          f"intersecting={int((ref['distance'] < 0).sum())} differing per field {diff}")
    assert all(v == 0 for v in diff.values())


@pytest.mark.gpu
@pytest.mark.parametrize("group", [1, 32])
def test_walk_penetration_bit_exact(oracle, gpu_ctx, group):
    h, S, faces, _ = _setup(oracle, 4096, 8)
    ta, tb = oracle.pose12(h["pose_a"]), oracle.pose12(h["pose_b"])
    ref, _ = S.collide(h["pair_a"], h["pair_b"], ta, tb, 1, True, threads=8)
    gpu_ctx.upload_shapes(h["types"], h["params"], h["vert_off"], h["vert_cnt"], h["verts"], faces=faces)
    out = _run(gpu_ctx, h, mfcl.request(mfcl.MODE_COLLIDE, max_contacts=1, penetration=1), group)
    fields = [f for f in FIELDS_PEN if f in ref.dtype.names]
    diff = {f: int((out[f] != ref[f]).reshape(len(out), -1).any(1).sum()) for f in fields}
    print(f"\n[walk collide, group {group}] pairs={len(out)} hits={int(ref['hit'].sum())} differing per field {diff}")
    assert all(v == 0 for v in diff.values())


@pytest.mark.gpu
def test_walk_faces_rejected(gpu_ctx):
    h = W.hull_pairs(64, 3, faces=True)
    gpu_ctx.upload_shapes(h["types"], h["params"], h["vert_off"], h["vert_cnt"], h["verts"])
    faces, off, cnt = mfcl.faces_from_triangles(h["tris"], h["tri_off"], h["tri_cnt"])
    bad = faces.copy()
    bad[off[5] + 1] = int(h["vert_cnt"][5])  # a vertex index one past the hull's last vertex
    with pytest.raises(RuntimeError, match="outside the hull"):
        gpu_ctx.upload_convex_faces(bad, off, cnt)
    bad = faces.copy()
    bad[off[7]] = 2  # a two-vertex face
    with pytest.raises(RuntimeError, match="invalid vertex count"):
        gpu_ctx.upload_convex_faces(bad, off, cnt)
    gpu_ctx.upload_convex_faces(faces, off, cnt)  # the valid faces are accepted afterwards
