"""GPU: the allocation-free convex query (mfcl_convex_query_static_dev, a
static per-GPU workspace) against mfcl_convex_query_dev (a reserved
workspace) through the C ABI: identical result records on two calls in a
row, the second smaller, and a clear error beyond the static capacity.
"""
import ctypes as C

import numpy as np
import pytest

import mfcl
from mfcl import workloads as W


def _dev_arrays(ctx, h, lo, hi, jitter=None):
    pa = np.ascontiguousarray(h["pair_a"][lo:hi], np.uint32)
    pb = np.ascontiguousarray(h["pair_b"][lo:hi], np.uint32)
    qa = np.ascontiguousarray(h["pose_a"][lo:hi], np.float32)
    qb = np.ascontiguousarray(h["pose_b"][lo:hi], np.float32).copy()
    if jitter is not None:
        qb[:, :3] += jitter[: hi - lo]
    return [ctx.to_device(x) for x in (pa, pb, qa, qb)]


def _run(ctx, fn, dev, n, req):
    out = ctx.alloc(n * mfcl.CONVEX_RESULT_DTYPE.itemsize)
    out.upload(np.full(n, 0x5A, np.uint8).view(np.uint8).repeat(mfcl.CONVEX_RESULT_DTYPE.itemsize))
    rc = fn(ctx.h, dev[0].ptr, dev[1].ptr, dev[2].ptr, dev[3].ptr, None, n, C.byref(req), out.ptr, None)
    assert rc == 0, ctx.lib.mfcl_last_error(ctx.h)
    ctx.sync()
    res = out.download(np.zeros(n, mfcl.CONVEX_RESULT_DTYPE))
    out.free()
    return res


@pytest.mark.gpu
def test_static_workspace_query_matches_reserved(gpu_ctx):
    lib = gpu_ctx.lib
    cap = C.c_uint64(0)
    assert lib.mfcl_convex_query_static_dev_max_pairs(C.byref(cap)) == 0
    cap = int(cap.value)
    print(f"\n[static workspace] capacity {cap} pairs")
    assert cap >= 65536
    h = W.hull_pairs(24576, 9, faces=True)
    gpu_ctx.upload_shapes(h["types"], h["params"], h["vert_off"], h["vert_cnt"], h["verts"],
                          faces=mfcl.faces_from_triangles(h["tris"], h["tri_off"], h["tri_cnt"]))
    rng = np.random.default_rng(3)
    for n, lo, mode, jit in ((24576, 0, mfcl.MODE_SIGNED_DISTANCE, None),
                             (5000, 7000, mfcl.MODE_COLLIDE, rng.normal(scale=0.05, size=(5000, 3)))):
        req = mfcl.request(mode, max_contacts=1, penetration=1)
        dev = _dev_arrays(gpu_ctx, h, lo, lo + n, jit)
        gpu_ctx.reserve(mfcl.WS_CONVEX, n)
        ref = _run(gpu_ctx, lib.mfcl_convex_query_dev, dev, n, req)
        got = _run(gpu_ctx, lib.mfcl_convex_query_static_dev, dev, n, req)
        epa = int((ref["epa_status"] >= 0).sum())
        print(f"[static workspace] n={n} mode={mode} EPA pairs={epa} hits={int(ref['hit'].sum())}")
        assert epa > 50
        assert got.tobytes() == ref.tobytes()
        for d in dev:
            d.free()
    # beyond the static capacity: a clear error, nothing launched
    d = gpu_ctx.alloc(64)
    req = mfcl.request(mfcl.MODE_SIGNED_DISTANCE)
    rc = lib.mfcl_convex_query_static_dev(gpu_ctx.h, d.ptr, d.ptr, d.ptr, d.ptr, None, cap + 1, C.byref(req), d.ptr,
                                          None)
    assert rc == -1  # MFCL_E_INVALID
    assert b"at most" in lib.mfcl_last_error(gpu_ctx.h)
    d.free()
