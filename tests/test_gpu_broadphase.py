"""GPU parity: flattened AABB-tree broadphase vs the oracle's tree and brute force.

Bar: for every query box the set of overlapping leaf ids equals the
oracle's (binary_AABB_tree-inl.h:400-440) and the brute-force all-pairs
set (the reference's own cross-check, test/broadphase/
test_binary_AABB_tree_collision.cpp:255-292).
"""
import numpy as np
import pytest


def _boxes(rng, n, lo, hi, hmin, hmax):
    c = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    h = rng.uniform(hmin, hmax, size=(n, 3)).astype(np.float32)
    return c - h, c + h


@pytest.mark.gpu
def test_aabb_tree_queries(oracle, gpu_ctx):
    rng = np.random.default_rng(6)
    mn, mx = _boxes(rng, 512, 0.0, 2.0, 0.02, 0.3)
    ids = np.arange(512, dtype=np.int64) * 7 + 3
    qmn, qmx = _boxes(rng, 4096, 0.0, 2.0, 0.01, 0.2)
    gpu_ctx.aabbtree_build(mn, mx, ids)
    cap = 128
    counts, got = gpu_ctx.aabbtree_query(qmn, qmx, cap)
    T = oracle.AABBTree(mn, mx, ids)
    rc, rids, _, _ = T.query(qmn, qmx, cap)
    brute = ~((qmn[:, None, :] > mx[None, :, :]).any(2) | (qmx[:, None, :] < mn[None, :, :]).any(2))
    assert np.all(counts >= 0)
    assert np.array_equal(counts, rc)
    assert np.array_equal(counts, brute.sum(1))
    for i in range(len(qmn)):
        k = min(cap, counts[i])
        assert set(got[i, :k].tolist()) == set(rids[i, :k].tolist()) == set(ids[brute[i]].tolist())
    print(f"\n[aabb tree] queries={len(qmn)} mean hits={counts.mean():.2f}")
