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


@pytest.mark.gpu
def test_mesh_config3_sizes(oracle, gpu_ctx):
    """The mesh-mesh configuration at its stated sizes: 4,096 pairs per
    launch over noisy icospheres of 1e3..1e5 triangles (log-uniform), about
    40 % overlapping, CollisionRequest(16).  Every pair's boolean equals the
    oracle's, no pair overflows the traversal stack, returned contacts are
    colliding triangle pairs of the oracle's full set."""
    pool = W.mesh_pool(256, 3)
    pairs = W.mesh_pairs(pool, 4096, 30)
    M = oracle.MeshSet(pool["verts"], pool["vert_off"], pool["vert_cnt"], pool["tris"], pool["tri_off"],
                       pool["tri_cnt"], threads=8)
    ta, tb = oracle.pose12(pairs["pose_a"]), oracle.pose12(pairs["pose_b"])
    gpu_ctx.upload_meshes(pool["verts"], pool["vert_off"], pool["vert_cnt"], pool["tris"], pool["tri_off"],
                          pool["tri_cnt"])
    out, cont = gpu_ctx.mesh_collide(pairs["pair_a"], pairs["pair_b"], pairs["pose_a"], pairs["pose_b"], 16)
    ref, _ = M.collide(pairs["pair_a"], pairs["pair_b"], ta, tb, 16, threads=8)
    overflow = int((out["hit"] < 0).sum())
    mism = out["hit"] != ref["hit"]
    print(f"\n[mesh cfg3] pairs={len(out)} triangles {int(pool['tri_cnt'].min())}..{int(pool['tri_cnt'].max())} "
          f"hits={int(ref['hit'].sum())} ({ref['hit'].mean():.3f}) overflow={overflow} bool_mismatch={int(mism.sum())} "
          f"bv tests mean {out['bv_tests'].mean():.0f} max {out['bv_tests'].max()}")
    assert overflow == 0
    assert mism.sum() == 0, np.nonzero(mism)[0][:20]
    assert np.array_equal(out["n_contacts"][ref["hit"] == 1] > 0, np.ones(int(ref["hit"].sum()), bool))
    rng = np.random.default_rng(0)
    hits = np.nonzero(ref["hit"])[0]
    for i in rng.choice(hits, size=min(48, len(hits)), replace=False):
        full, nfull = M.collide_all(int(pairs["pair_a"][i]), int(pairs["pair_b"][i]), ta[i], tb[i])
        got = set(map(tuple, cont[i, :out["n_contacts"][i]].tolist()))
        assert out["n_contacts"][i] == min(16, nfull)
        assert got <= set(map(tuple, full.tolist()))


@pytest.mark.gpu
@pytest.mark.parametrize("cap", [16, 1024])
def test_mesh_penetration_contacts(oracle, gpu_ctx, cap):
    """Penetration-enabled mesh-mesh collide: every contact's triangle ids,
    position, normal and depth equal one of the oracle's (bit for bit);
    with a cap above the pair's contact count the two contact sets are
    equal; the boolean (some contact) equals the oracle's."""
    pool = _pool(16, 31, 300, 4000)
    pairs = W.mesh_pairs(pool, 256, 32, spread=1.3)
    M = oracle.MeshSet(pool["verts"], pool["vert_off"], pool["vert_cnt"], pool["tris"], pool["tri_off"],
                       pool["tri_cnt"], threads=8)
    ta, tb = oracle.pose12(pairs["pose_a"]), oracle.pose12(pairs["pose_b"])
    gpu_ctx.upload_meshes(pool["verts"], pool["vert_off"], pool["vert_cnt"], pool["tris"], pool["tri_off"],
                          pool["tri_cnt"])
    out, cont = gpu_ctx.mesh_collide_contacts(pairs["pair_a"], pairs["pair_b"], pairs["pose_a"], pairs["pose_b"],
                                              cap)
    nref, _, _ = M.collide_pen(pairs["pair_a"], pairs["pair_b"], ta, tb, cap, threads=8)
    assert np.all(out["hit"] >= 0)
    assert np.array_equal(out["hit"], (nref > 0).astype(np.int32))

    def key(c):
        return (int(c["b1"]), int(c["b2"]), *c["pos"].tolist(), *c["normal"].tolist(), float(c["depth"]))

    n_eq = n_sub = 0
    for i in np.nonzero(out["hit"])[0]:
        full, nfull = M.collide_all_pen(int(pairs["pair_a"][i]), int(pairs["pair_b"][i]), ta[i], tb[i])
        fullset = [key(c) for c in full]
        got = [key(c) for c in cont[i, :out["n_contacts"][i]]]
        assert out["n_contacts"][i] == min(cap, nfull)
        assert set(got) <= set(fullset), (i, got[:2], fullset[:2])
        if nfull <= cap:
            assert sorted(got) == sorted(fullset)
            n_eq += 1
        else:
            n_sub += 1
    print(f"\n[mesh penetration cap={cap}] hits={int(out['hit'].sum())} equal sets {n_eq} subsets {n_sub} "
          f"contacts {int(out['n_contacts'].sum())}")
    assert n_eq + n_sub > 20
