"""GPU: query sequences recorded into a HIP graph (mfcl_graph_*) and
replayed: every replay equals the eager run of the same calls, with eager
work and synchronisations between replays, and inputs changed between
replays are picked up."""
import numpy as np
import pytest

import mfcl
from mfcl import workloads as W


def _records(a):
    return a.view(np.uint8).reshape(len(a), -1)


@pytest.mark.gpu
def test_graph_convex_warm_start_sequence(gpu_ctx):
    """MLP guesses + the warm-started signed-distance query + a boolean
    collide query, recorded once, replayed three times."""
    h = W.hull_pairs(20000, 13, faces=True)
    gpu_ctx.upload_shapes(h["types"], h["params"], h["vert_off"], h["vert_cnt"], h["verts"],
                          faces=mfcl.faces_from_triangles(h["tris"], h["tri_off"], h["tri_cnt"]))
    from mfcl import warmstart as WS

    params, _ = WS.load_params()
    gpu_ctx.mlp_set_params(params)
    n = len(h["pair_a"])
    dev = [gpu_ctx.to_device(h[k]) for k in ("pair_a", "pair_b", "pose_a", "pose_b")]
    dg = gpu_ctx.alloc(n * 12)
    out_sd, out_b = gpu_ctx.alloc(n * 128), gpu_ctx.alloc(n * 128)
    req_w = mfcl.request(mfcl.MODE_SIGNED_DISTANCE, use_guess=1)
    req_b = mfcl.request(mfcl.MODE_COLLIDE, max_contacts=1, penetration=0)
    gpu_ctx.reserve(mfcl.WS_CONVEX, n)

    def calls():
        gpu_ctx.mlp_guess_dev(*dev, n, dg)
        gpu_ctx.convex_query_dev(*dev, n, req_w, out_sd, guess=dg)
        gpu_ctx.convex_query_dev(*dev, n, req_b, out_b)

    calls()
    gpu_ctx.sync()
    ref_sd = out_sd.download(np.zeros(n, mfcl.CONVEX_RESULT_DTYPE))
    ref_b = out_b.download(np.zeros(n, mfcl.CONVEX_RESULT_DTYPE))
    gpu_ctx.graph_begin()
    calls()
    g = gpu_ctx.graph_end()
    with pytest.raises(RuntimeError, match="recorded"):
        gpu_ctx.graph_begin()
        try:
            gpu_ctx.sync()
        finally:
            gpu_ctx.graph_end().destroy()
    for rep in range(3):
        out_sd.upload(np.zeros(n * 128, np.uint8))
        out_b.upload(np.zeros(n * 128, np.uint8))
        g.launch()
        gpu_ctx.sync()
        got_sd = out_sd.download(np.zeros(n, mfcl.CONVEX_RESULT_DTYPE))
        got_b = out_b.download(np.zeros(n, mfcl.CONVEX_RESULT_DTYPE))
        assert np.array_equal(_records(got_sd), _records(ref_sd)), rep
        assert np.array_equal(_records(got_b), _records(ref_b)), rep
        # eager work between replays
        gpu_ctx.convex_query_dev(*dev, n // 2, req_b, out_b)
        gpu_ctx.sync()
    # new poses in the same buffers: a replay computes with them
    pose_b2 = h["pose_b"].copy()
    pose_b2[:, :3] += 0.05
    dev[3].upload(pose_b2)
    g.launch()
    gpu_ctx.sync()
    got = out_sd.download(np.zeros(n, mfcl.CONVEX_RESULT_DTYPE))
    gpu_ctx.mlp_guess_dev(*dev, n, dg)
    gpu_ctx.convex_query_dev(*dev, n, req_w, out_sd, guess=dg)
    gpu_ctx.sync()
    eager = out_sd.download(np.zeros(n, mfcl.CONVEX_RESULT_DTYPE))
    assert np.array_equal(_records(got), _records(eager))
    assert not np.array_equal(_records(got), _records(ref_sd))
    g.destroy()
    for a in dev + [dg, out_sd, out_b]:
        a.free()


@pytest.mark.gpu
def test_graph_mesh_query(gpu_ctx):
    pool = W.mesh_pool(12, 41, 500, 4000)
    pairs = W.mesh_pairs(pool, 512, 42)
    gpu_ctx.upload_meshes(pool["verts"], pool["vert_off"], pool["vert_cnt"], pool["tris"], pool["tri_off"],
                          pool["tri_cnt"])
    n = len(pairs["pair_a"])
    dev = [gpu_ctx.to_device(pairs[k]) for k in ("pair_a", "pair_b", "pose_a", "pose_b")]
    dres, dcon = gpu_ctx.alloc(n * 16), gpu_ctx.alloc(n * 16 * 8)
    gpu_ctx.mesh_collide_dev(*dev, n, 16, dres, dcon)
    gpu_ctx.sync()
    ref = dres.download(np.zeros(n, mfcl.MESH_RESULT_DTYPE))
    gpu_ctx.graph_begin()
    gpu_ctx.mesh_collide_dev(*dev, n, 16, dres, dcon)
    g = gpu_ctx.graph_end()
    for rep in range(3):
        dres.upload(np.zeros(n * 16, np.uint8))
        g.launch()
        gpu_ctx.sync()
        got = dres.download(np.zeros(n, mfcl.MESH_RESULT_DTYPE))
        assert np.array_equal(got["hit"], ref["hit"]), rep
        assert np.array_equal(got["n_contacts"], ref["n_contacts"]), rep
    g.destroy()
    for a in dev + [dres, dcon]:
        a.free()
