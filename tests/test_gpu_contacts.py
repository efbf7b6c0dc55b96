"""GPU parity: every contact of a collide query (mfcl_convex_collide_contacts)
against the oracle's CollisionResult contact list, contact by contact.

Box pairs produce up to four contacts (boxBox2, box_box-inl.h:214-803);
CollisionRequest(k) keeps the k deepest through the reference's
std::partial_sort (shape_pair_intersect-inl.h:91-113), so equal depths
(axis-aligned stacked boxes) also test that the device reproduces that
sort's order.  Bar: bit-identical normal, position and depth of every kept
contact, identical counts.
"""
import numpy as np
import pytest

import mfcl
from mfcl import workloads as W

BOX = 0


def _rot(rng):
    q = rng.normal(size=4)
    return q / np.linalg.norm(q)


def _box_workload(n, seed):
    """Random box pairs (about half colliding, mostly face contacts with up
    to four points) plus axis-aligned stacks whose four contacts tie."""
    rng = np.random.default_rng(seed)
    sides = rng.uniform(0.1, 1.0, size=(2 * n, 3)).astype(np.float32)
    pose_a = np.zeros((n, 7), np.float32)
    pose_b = np.zeros((n, 7), np.float32)
    pose_a[:, :3] = rng.uniform(0, 1, (n, 3))
    pose_b[:, :3] = rng.uniform(0, 1, (n, 3))
    pose_a[:, 3:] = np.stack([_rot(rng) for _ in range(n)])
    pose_b[:, 3:] = np.stack([_rot(rng) for _ in range(n)])
    # every fourth pair: an axis-aligned box resting in the top face of another
    k = np.arange(0, n, 4)
    pose_a[k, 3:] = pose_b[k, 3:] = (1, 0, 0, 0)
    sides[2 * k] = (1.0, 1.0, 0.4)
    sides[2 * k + 1] = (0.3, 0.25, 0.2)
    pose_b[k, :3] = pose_a[k, :3] + np.array([0.05, -0.02, 0.2 + 0.1 - 0.01], np.float32)
    types = np.full(2 * n, BOX, np.int32)
    params = np.zeros((2 * n, 4), np.float32)
    params[:, :3] = sides
    return dict(types=types, params=params, pair_a=np.arange(0, 2 * n, 2, dtype=np.uint32),
                pair_b=np.arange(1, 2 * n, 2, dtype=np.uint32), pose_a=pose_a, pose_b=pose_b)


def _compare(oracle, w, out, cont, max_contacts):
    S = oracle.ShapeSet(w["types"], w["params"][:, :3])
    ta, tb = oracle.pose12(w["pose_a"]), oracle.pose12(w["pose_b"])
    ref, rc = S.collide_contacts(w["pair_a"], w["pair_b"], ta, tb, max_contacts, True, threads=8)
    assert np.array_equal(out["hit"], ref["hit"])
    assert np.array_equal(out["n_contacts"], ref["n"])
    stride = min(max_contacts, 4)
    n_multi = 0
    for i in np.nonzero(ref["hit"])[0]:
        m = int(ref["n"][i])
        n_multi += m > 1
        for k in range(m):
            g, r = cont[i, k], rc[i, k]
            assert np.array_equal(g["normal"], r["normal"]), (i, k, g, r)
            assert np.array_equal(g["pos"], r["pos"]), (i, k, g, r)
            assert g["depth"] == r["depth"], (i, k, g["depth"], r["depth"])
        # the result record holds contact 0
        assert out["depth"][i] == rc[i, 0]["depth"]
        assert np.array_equal(out["normal"][i], rc[i, 0]["normal"])
        assert np.array_equal(out["pos"][i], rc[i, 0]["pos"])
    assert cont.shape[1] == stride
    return n_multi


@pytest.mark.gpu
@pytest.mark.parametrize("max_contacts", [4, 3, 2, 1, 8])
def test_box_box_all_contacts(oracle, gpu_ctx, max_contacts):
    w = _box_workload(4096, 40 + max_contacts)
    gpu_ctx.upload_shapes(w["types"], w["params"])
    req = mfcl.request(mfcl.MODE_COLLIDE, max_contacts=max_contacts, penetration=1)
    out, cont = gpu_ctx.convex_collide_contacts(w["pair_a"], w["pair_b"], w["pose_a"], w["pose_b"], req)
    n_multi = _compare(oracle, w, out, cont, max_contacts)
    print(f"\n[box contacts k={max_contacts}] hits={int(out['hit'].sum())} pairs with >1 contact={n_multi} "
          f"contacts={int(out['n_contacts'].sum())}")
    if max_contacts > 1:
        assert n_multi > 500
    # ties: the stacked pairs' four contacts have equal depths
    k = np.arange(0, len(out), 4)
    assert np.all(out["hit"][k] == 1)


@pytest.mark.gpu
def test_mixed_contacts_dev(oracle, gpu_ctx):
    """The device-pointer entry on the mixed primitive workload (box, sphere,
    capsule, cylinder): GJK/EPA pairs and closed forms write their contact
    into the array too."""
    w = W.primitive_pairs(4096, 1)
    gpu_ctx.upload_shapes(w["types"], w["params"])
    n = len(w["pair_a"])
    req = mfcl.request(mfcl.MODE_COLLIDE, max_contacts=4, penetration=1)
    dev = [gpu_ctx.to_device(np.ascontiguousarray(x)) for x in
           (w["pair_a"].astype(np.uint32), w["pair_b"].astype(np.uint32), w["pose_a"].astype(np.float32),
            w["pose_b"].astype(np.float32))]
    dout = gpu_ctx.alloc(n * mfcl.CONVEX_RESULT_DTYPE.itemsize)
    dcon = gpu_ctx.alloc(n * 4 * mfcl.CONTACT_DTYPE.itemsize)
    dcon.upload(np.zeros((n, 4), mfcl.CONTACT_DTYPE))
    gpu_ctx.convex_collide_contacts_dev(*dev, n, req, dout, dcon)
    out = dout.download(np.zeros(n, mfcl.CONVEX_RESULT_DTYPE))
    cont = dcon.download(np.zeros((n, 4), mfcl.CONTACT_DTYPE))
    S = oracle.ShapeSet(w["types"], w["params"][:, :3])
    ta, tb = oracle.pose12(w["pose_a"]), oracle.pose12(w["pose_b"])
    ref, rc = S.collide_contacts(w["pair_a"], w["pair_b"], ta, tb, 4, True, threads=8)
    # outside the touching band the booleans agree; where both hit, contacts
    # match bit for bit (the GJK/EPA path is bit-exact with the oracle)
    both = (out["hit"] == 1) & (ref["hit"] == 1)
    assert both.sum() > 1000
    assert np.array_equal(out["n_contacts"][both], ref["n"][both])
    for i in np.nonzero(both)[0]:
        for k in range(int(ref["n"][i])):
            assert np.array_equal(cont[i, k]["normal"], rc[i, k]["normal"]), (i, k)
            assert np.array_equal(cont[i, k]["pos"], rc[i, k]["pos"]), (i, k)
            assert cont[i, k]["depth"] == rc[i, k]["depth"], (i, k)
    for a in dev + [dout, dcon]:
        a.free()


@pytest.mark.gpu
def test_contacts_need_penetration_and_reserve(gpu_ctx):
    w = _box_workload(64, 7)
    gpu_ctx.upload_shapes(w["types"], w["params"])
    with pytest.raises(RuntimeError, match="penetration"):
        gpu_ctx.convex_collide_contacts(w["pair_a"], w["pair_b"], w["pose_a"], w["pose_b"],
                                        mfcl.request(mfcl.MODE_COLLIDE, max_contacts=4, penetration=0))
    # a device query beyond the reserved workspace fails instead of allocating
    import ctypes as C

    big = gpu_ctx.reserved_pairs() + 1
    d = gpu_ctx.alloc(64)
    req = mfcl.request(mfcl.MODE_COLLIDE, max_contacts=1, penetration=1)
    rc = gpu_ctx.lib.mfcl_convex_query_dev(gpu_ctx.h, d.ptr, d.ptr, d.ptr, d.ptr, None, big, C.byref(req), d.ptr,
                                          None)
    assert rc == -5  # MFCL_E_STATE
    assert b"mfcl_reserve" in gpu_ctx.lib.mfcl_last_error(gpu_ctx.h)
    d.free()
