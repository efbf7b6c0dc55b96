"""GPU parity: batched convex narrowphase vs the FP64 oracle.

Bar: collision booleans identical except pairs inside the touching band
(|signed distance| <= 1e-6 * scale, scale = sum of the two bounding radii),
which are counted and reported; depth / distance within 1e-5 relative to
max(|reference|, 1e-3 * scale); contact normals within 1e-5 (1 - cos).
"""
import numpy as np
import pytest

from common import hull_shapes_for_oracle, scale_of
from mfcl import workloads as W
import mfcl

REL = 1e-5
BAND = 1e-6


def _prim_setup(oracle, n=4096, seed=1):
    w = W.primitive_pairs(n, seed)
    S = oracle.ShapeSet(w["types"], w["params"])
    ta, tb = oracle.pose12(w["pose_a"]), oracle.pose12(w["pose_b"])
    scale = scale_of(w["types"], w["params"], w["pair_a"], w["pair_b"])
    sd, _ = S.signed_distance(w["pair_a"], w["pair_b"], ta, tb, threads=8)
    touching = np.abs(sd["distance"]) <= BAND * scale
    return w, S, ta, tb, scale, touching


@pytest.mark.gpu
@pytest.mark.parametrize("penetration", [1, 0])
def test_primitive_pairs_collide(oracle, gpu_ctx, penetration):
    w, S, ta, tb, scale, touching = _prim_setup(oracle)
    ref, _ = S.collide(w["pair_a"], w["pair_b"], ta, tb, 1, bool(penetration), threads=8)
    gpu_ctx.upload_shapes(w["types"], w["params"])
    req = mfcl.request(mfcl.MODE_COLLIDE, max_contacts=1, penetration=penetration)
    out = gpu_ctx.convex_query(w["pair_a"], w["pair_b"], w["pose_a"], w["pose_b"], req)
    mism = (out["hit"] != ref["hit"]) & ~touching
    print(f"\n[primitive collide pen={penetration}] pairs={len(out)} hits={int(ref['hit'].sum())} "
          f"touching_band={int(touching.sum())} bool_mismatch_outside_band={int(mism.sum())}")
    assert mism.sum() == 0, np.nonzero(mism)[0][:20]
    if penetration:
        both = (out["hit"] == 1) & (ref["hit"] == 1) & ~touching
        tol = REL * np.maximum(np.abs(ref["depth"]), 1e-3 * scale)
        derr = np.abs(out["depth"] - ref["depth"])
        cosd = 1.0 - np.sum(out["normal"] * ref["normal"], axis=1)
        perr = np.linalg.norm(out["pos"] - ref["pos"], axis=1)
        print(f"  depth max abs err {derr[both].max():.3e}, max err/tol {np.max(derr[both] / tol[both]):.3e}, "
              f"normal max(1-cos) {cosd[both].max():.3e}, pos max err {perr[both].max():.3e}")
        bad = both & (derr > tol)
        assert bad.sum() == 0, np.nonzero(bad)[0][:20]
        assert np.all(cosd[both] <= 1e-5)
        assert np.all(perr[both] <= REL * scale[both] + 1e-9)
        # closed-form pairs match the reference's arithmetic almost exactly
        cf = both & (ref["method"] == 0)
        assert np.all(derr[cf] <= 1e-9)


def _hull_setup(oracle, n_pairs, seed):
    h = W.hull_pairs(n_pairs, seed)
    idx = np.arange(2 * n_pairs)
    v, vo, vc, t, to, tc = hull_shapes_for_oracle(h, idx)
    S = oracle.ShapeSet(h["types"], h["params"], v, vo, vc, t, to, tc)
    ta, tb = oracle.pose12(h["pose_a"]), oracle.pose12(h["pose_b"])
    scale = scale_of(h["types"], h["params"], h["pair_a"], h["pair_b"], h["verts"], h["vert_off"], h["vert_cnt"])
    return h, S, ta, tb, scale


@pytest.mark.gpu
def test_hull_pairs_signed_distance(oracle, gpu_ctx):
    h, S, ta, tb, scale = _hull_setup(oracle, 4096, 2)
    ref, _ = S.signed_distance(h["pair_a"], h["pair_b"], ta, tb, threads=8)
    gpu_ctx.upload_shapes(h["types"], h["params"], h["vert_off"], h["vert_cnt"], h["verts"])
    req = mfcl.request(mfcl.MODE_SIGNED_DISTANCE)
    out = gpu_ctx.convex_query(h["pair_a"], h["pair_b"], h["pose_a"], h["pose_b"], req)
    touching = np.abs(ref["distance"]) <= BAND * scale
    ref_hit = ref["distance"] < 0
    gpu_hit = out["distance"] < 0
    mism = (ref_hit != gpu_hit) & ~touching
    okm = (ref["ok"] != out["ok"])
    both = (ref["ok"] == 1) & (out["ok"] == 1) & ~touching
    tol = REL * np.maximum(np.abs(ref["distance"]), 1e-3 * scale)
    derr = np.abs(out["distance"] - ref["distance"])
    werr = np.maximum(np.linalg.norm(out["pa"] - ref["pa"], axis=1), np.linalg.norm(out["pb"] - ref["pb"], axis=1))
    print(f"\n[hull signed distance] pairs={len(out)} intersecting={int(ref_hit.sum())} "
          f"touching_band={int(touching.sum())} sign_mismatch={int(mism.sum())} ok_mismatch={int(okm.sum())} "
          f"dist max abs err {derr[both].max():.3e} max err/tol {np.max(derr[both] / tol[both]):.3e} "
          f"witness max err {werr[both].max():.3e} gjk iters mean {out['gjk_iters'].mean():.2f}")
    assert mism.sum() == 0, np.nonzero(mism)[0][:20]
    assert okm.sum() == 0, np.nonzero(okm)[0][:20]
    bad = both & (derr > tol)
    assert bad.sum() == 0, np.nonzero(bad)[0][:20]
    assert np.all(werr[both] <= REL * scale[both] * 10)
