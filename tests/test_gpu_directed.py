"""GPU parity: collisionPenetration in its MPR modes -- directed
(mfcl_convex_directed_penetration; mpr.hpp:496-704) and incremental minimum
penetration (mfcl_convex_incremental_penetration;
mpr_incremental_penetration.hpp) -- against the oracle, whose restatements
are pinned by closed forms, a linear program and EPA
(test_oracle_directed.py).

Hull pairs (walking and scanning hulls) and the primitive mix: the boolean
collide, the MPR status and iteration count, the depth along the
direction, the contact position and the normal must be identical, bit for
bit.
"""
import numpy as np
import pytest

import mfcl
from mfcl import workloads as W

FIELDS = ("hit", "n_contacts", "epa_status", "epa_iters", "depth", "pos", "normal")


def _dirs(rng, n):
    u = rng.normal(size=(n, 3))
    return (u / np.linalg.norm(u, axis=1, keepdims=True)).astype(np.float32)


def _compare(out, ref):
    diff = {
        "hit": int((out["hit"] != ref["hit"]).sum()),
        "n_contacts": int((out["n_contacts"] != ref["n"]).sum()),
    }
    hit = ref["n"] > 0
    for f in ("epa_status", "epa_iters", "depth"):
        diff[f] = int((out[f][hit] != ref[f][hit]).sum())
    for f in ("pos", "normal"):
        diff[f] = int((out[f][hit] != ref[f][hit]).any(1).sum())
    return diff, hit


@pytest.mark.gpu
@pytest.mark.parametrize("incremental", [False, True])
def test_directed_hulls_bit_exact(oracle, gpu_ctx, incremental):
    h = W.hull_pairs(4096, 73, faces=True)
    rng = np.random.default_rng(74)
    n = len(h["pair_a"])
    h["pose_b"][:, :3] = h["pose_a"][:, :3] + rng.uniform(-0.25, 0.25, (n, 3)).astype(np.float32)
    dirs = _dirs(rng, n)
    S = oracle.ShapeSet(h["types"], h["params"], h["verts"], h["vert_off"], h["vert_cnt"], h["tris"],
                        h["tri_off"], h["tri_cnt"])
    gpu_ctx.upload_shapes(h["types"], h["params"], h["vert_off"], h["vert_cnt"], h["verts"],
                          faces=mfcl.faces_from_triangles(h["tris"], h["tri_off"], h["tri_cnt"]))
    req = mfcl.request(mfcl.MODE_COLLIDE, max_contacts=1)
    out = gpu_ctx.convex_directed_penetration(h["pair_a"], h["pair_b"], h["pose_a"], h["pose_b"], dirs, req,
                                              incremental=incremental)
    ref, _ = S.directed_penetration(h["pair_a"], h["pair_b"], oracle.pose12(h["pose_a"]), oracle.pose12(h["pose_b"]),
                                    dirs.astype(np.float64), 1, req.epa_tolerance, threads=8, incremental=incremental)
    diff, hit = _compare(out, ref)
    st = np.unique(ref["epa_status"][hit], return_counts=True)
    print(f"\n[{'incremental' if incremental else 'directed'} hulls] pairs={n} colliding={int(hit.sum())} MPR statuses {dict(zip(*st))} "
          f"iterations mean {ref['epa_iters'][hit].mean():.2f}; differing per field {diff}")
    assert hit.sum() > n // 3
    assert all(v == 0 for v in diff.values())


@pytest.mark.gpu
@pytest.mark.parametrize("incremental", [False, True])
def test_directed_primitives_bit_exact(oracle, gpu_ctx, incremental):
    w = W.primitive_pairs(4096, 75)
    rng = np.random.default_rng(76)
    n = len(w["pair_a"])
    dirs = _dirs(rng, n)
    S = oracle.ShapeSet(w["types"], w["params"][:, :3])
    gpu_ctx.upload_shapes(w["types"], w["params"])
    req = mfcl.request(mfcl.MODE_COLLIDE, max_contacts=1)
    out = gpu_ctx.convex_directed_penetration(w["pair_a"], w["pair_b"], w["pose_a"], w["pose_b"], dirs, req,
                                              incremental=incremental)
    ref, _ = S.directed_penetration(w["pair_a"], w["pair_b"], oracle.pose12(w["pose_a"]), oracle.pose12(w["pose_b"]),
                                    dirs.astype(np.float64), 1, req.epa_tolerance, threads=8, incremental=incremental)
    diff, hit = _compare(out, ref)
    print(f"\n[{'incremental' if incremental else 'directed'} primitives] pairs={n} colliding={int(hit.sum())}; differing per field {diff}")
    assert hit.sum() > 1000
    assert all(v == 0 for v in diff.values())
