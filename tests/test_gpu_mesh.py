"""GPU parity: OBB-tree mesh-mesh collision vs the FP64 oracle.

Bar (CollisionRequest(16), penetration disabled): the collision boolean is
identical for every pair; every returned (b1, b2) belongs to the oracle's
full colliding triangle-pair set; when that set has at most 16 members the
two sets are equal.  (Which 16 of a larger set come back depends on the
traversal order, which is not specified by the boolean/contact contract.)
"""
import numpy as np
import pytest

from mfcl import workloads as W


def _pool(n_meshes, seed, tmin, tmax):
    return W.mesh_pool(n_meshes, seed, tmin, tmax)


@pytest.mark.gpu
def test_mesh_pairs_first16(oracle, gpu_ctx):
    pool = _pool(24, 11, 200, 6000)
    pairs = W.mesh_pairs(pool, 384, 12)
    M = oracle.MeshSet(pool["verts"], pool["vert_off"], pool["vert_cnt"], pool["tris"], pool["tri_off"],
                       pool["tri_cnt"])
    ta, tb = oracle.pose12(pairs["pose_a"]), oracle.pose12(pairs["pose_b"])
    gpu_ctx.upload_meshes(pool["verts"], pool["vert_off"], pool["vert_cnt"], pool["tris"], pool["tri_off"],
                          pool["tri_cnt"])
    out, cont = gpu_ctx.mesh_collide(pairs["pair_a"], pairs["pair_b"], pairs["pose_a"], pairs["pose_b"], 16)
    ref, _ = M.collide(pairs["pair_a"], pairs["pair_b"], ta, tb, 16, threads=8)
    assert np.all(out["hit"] >= 0), "traversal stack overflow / bad ids"
    mism = out["hit"] != ref["hit"]
    print(f"\n[mesh first-16] pairs={len(out)} hits={int(ref['hit'].sum())} bool_mismatch={int(mism.sum())} "
          f"gpu bv tests mean {out['bv_tests'].mean():.0f} (oracle {ref['bv_tests'].mean():.0f}), "
          f"leaf tests mean {out['leaf_tests'].mean():.0f} (oracle {ref['leaf_tests'].mean():.0f})")
    assert mism.sum() == 0, np.nonzero(mism)[0][:20]
    checked_full = 0
    for i in np.nonzero(ref["hit"])[0]:
        full, nfull = M.collide_all(int(pairs["pair_a"][i]), int(pairs["pair_b"][i]), ta[i], tb[i])
        fullset = set(map(tuple, full.tolist()))
        got = set(map(tuple, cont[i, :out["n_contacts"][i]].tolist()))
        assert out["n_contacts"][i] == min(16, nfull)
        assert got <= fullset
        if nfull <= 16:
            assert got == fullset
            checked_full += 1
    print(f"  pairs with full-set equality checked: {checked_full}")


@pytest.mark.gpu
def test_mesh_full_sets_unbounded(oracle, gpu_ctx):
    """With a contact cap above the set size the GPU returns the complete
    colliding set, equal to the oracle's (and independent of tree shape)."""
    pool = _pool(8, 21, 500, 3000)
    pairs = W.mesh_pairs(pool, 64, 22, spread=1.2)
    M = oracle.MeshSet(pool["verts"], pool["vert_off"], pool["vert_cnt"], pool["tris"], pool["tri_off"],
                       pool["tri_cnt"])
    ta, tb = oracle.pose12(pairs["pose_a"]), oracle.pose12(pairs["pose_b"])
    gpu_ctx.upload_meshes(pool["verts"], pool["vert_off"], pool["vert_cnt"], pool["tris"], pool["tri_off"],
                          pool["tri_cnt"])
    cap = 1024
    out, cont = gpu_ctx.mesh_collide(pairs["pair_a"], pairs["pair_b"], pairs["pose_a"], pairs["pose_b"], cap)
    n_eq = 0
    for i in range(len(out)):
        full, nfull = M.collide_all(int(pairs["pair_a"][i]), int(pairs["pair_b"][i]), ta[i], tb[i])
        if nfull >= cap:
            continue
        got = set(map(tuple, cont[i, :out["n_contacts"][i]].tolist()))
        assert got == set(map(tuple, full.tolist())), i
        n_eq += 1
    print(f"\n[mesh full sets] equal sets on {n_eq}/{len(out)} pairs")
    assert n_eq > len(out) // 2
