"""CPU: pin the oracle against the reference's own known-answer tests.

Every case here comes from a test in the reference's test tree (file:line
given per test; the numeric inputs are stored in
tests/golden/reference_kats.json and the two OBJ meshes are the reference's
test resources).  Each check is one the reference itself makes, at the
reference's tolerance or tighter.
"""
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
BOX, SPHERE, CAPSULE, CYLINDER, CONVEX = 0, 1, 3, 5, 100
GJK_INTERSECT, GJK_SEPARATED = 0, 1
MPR_INTERSECT, MPR_SEPARATED = 0, 1


@pytest.fixture(scope="module")
def kats():
    with open(os.path.join(GOLDEN, "reference_kats.json")) as f:
        return json.load(f)


def _tf(R_rows, t):
    out = np.zeros(12)
    out[:9] = np.asarray(R_rows, np.float64).reshape(9)
    out[9:] = t
    return out


def _ident(t=(0.0, 0.0, 0.0)):
    return _tf(np.eye(3), t)


def _compose(A, B):
    """A * B for 12-float (row-major R, t) transforms."""
    Ra, Rb = A[:9].reshape(3, 3), B[:9].reshape(3, 3)
    return _tf(Ra @ Rb, Ra @ B[9:] + A[9:])


def _random_tf(rng, extent=3.0):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    return _tf(R, rng.uniform(-extent, extent, 3))


def _prims(oracle, specs):
    """ShapeSet of primitives from ("box", sx, sy, sz) style specs."""
    code = {"box": BOX, "sphere": SPHERE, "capsule": CAPSULE, "cylinder": CYLINDER}
    types = np.array([code[s[0]] for s in specs], np.int32)
    params = np.zeros((len(specs), 3))
    for i, s in enumerate(specs):
        params[i, :len(s) - 1] = s[1:]
    return oracle.ShapeSet(types, params)


def test_box_box_fixed_instance(oracle, kats):
    """test/cvx_collide/test_gjk2.cpp:130-223: GJK(1000, 1e-6) and MPR agree
    with the boxBox2 closed form on a near-touching configuration."""
    k = kats["box_box_fixed_instance"]
    S = _prims(oracle, [("box", *k["box1_side"]), ("box", *k["box2_side"])])
    tf1, tf2 = _tf(k["pose1_R_rows"], k["pose1_t"]), _tf(k["pose2_R_rows"], k["pose2_t"])
    closed, _ = S.collide([0], [1], tf1[None], tf2[None], max_contacts=4, penetration=True)
    assert closed["method"][0] == 0  # closed form
    gjk, _, _, _ = S.pair_algorithm(0, 1, tf1, tf2, 0, k["gjk_max_iterations"], k["tolerance"])
    mpr, _, _, _ = S.pair_algorithm(0, 1, tf1, tf2, 1, k["gjk_max_iterations"], k["tolerance"])
    if closed["hit"][0]:
        assert gjk == GJK_INTERSECT and mpr == MPR_INTERSECT
    else:
        assert mpr == MPR_SEPARATED


def test_capsule_capsule_fixed_instance(oracle, kats):
    """test/cvx_collide/test_gjk2_distance.cpp:69-150: GJK2(1000, 1e-7)
    distance on a capsule pair vs capsuleCapsuleDistance (the reference
    prints an error above 5e-3; we require 1e-6)."""
    k = kats["capsule_capsule_fixed_instance"]
    S = _prims(oracle, [("capsule", k["radius"], k["lz"]), ("capsule", k["radius"], k["lz"])])
    tf1, tf2 = _tf(k["pose1_R_rows"], k["pose1_t"]), _tf(k["pose2_R_rows"], k["pose2_t"])
    st, dist, _, _ = S.pair_algorithm(0, 1, tf1, tf2, 2, k["gjk_max_iterations"], k["tolerance"])
    ref = oracle.capsule_capsule_distance(k["radius"], k["lz"], tf1, k["radius"], k["lz"], tf2)
    assert st == GJK_SEPARATED
    assert ref > 0
    assert abs(dist - ref) < 1e-6, (dist, ref)


def test_sphere_separation_random(oracle, kats):
    """test/cvx_collide/test_gjk2_distance.cpp:18-59 (1e6 trials there, 2000
    here): spheres of radius 1 at centre distance 2.2, random orientations;
    GJK separation distance 0.2 within 1e-3."""
    k = kats["sphere_separation"]
    r = k["radius"]
    S = _prims(oracle, [("sphere", r), ("sphere", r)])
    rng = np.random.default_rng(0)
    errs = []
    for _ in range(2000):
        d = _random_tf(rng)[:9].reshape(3, 3)[:, 0]
        tf1, tf2 = _random_tf(rng), _random_tf(rng)
        tf2[9:] = tf1[9:] + d * (2 * r + k["expected_distance"])
        st, dist, _, _ = S.pair_algorithm(0, 1, tf1, tf2, 2, k["gjk_max_iterations"], k["gjk_tolerance"])
        assert st == GJK_SEPARATED
        errs.append(abs(dist - k["expected_distance"]))
    assert max(errs) < k["tolerance_reference"], max(errs)


def test_sphere_penetration_epa_random(oracle, kats):
    """test/cvx_collide/test_epa2_with_gjk2.cpp:21-73: unit spheres
    penetrating by 0.2; GJK(1000, 1e-6) + EPA depth within 1e-4."""
    k = kats["sphere_penetration"]
    r = k["radius"]
    # As convex objects the spheres go through GJK/EPA (no closed form):
    # use the signed-distance path, which has no sphere-sphere shortcut.
    S = _prims(oracle, [("sphere", r), ("sphere", r)])
    rng = np.random.default_rng(1)
    n = 2000
    ta = np.zeros((n, 12))
    tb = np.zeros((n, 12))
    for i in range(n):
        d = _random_tf(rng)[:9].reshape(3, 3)[:, 0]
        ta[i], tb[i] = _random_tf(rng), _random_tf(rng)
        tb[i, 9:] = ta[i, 9:] + d * (2 * r - k["expected_depth"])
    solver = np.array([k["gjk_max_iterations"], k["gjk_tolerance"], 256, 255, 1e-6])
    out, _ = S.signed_distance(np.zeros(n), np.ones(n), ta, tb, solver=solver)
    assert np.all(out["ok"] == 1)
    err = np.abs(-out["distance"] - k["expected_depth"])
    assert err.max() < k["tolerance_reference"], err.max()


def test_gjk_shape_intersection_kats(oracle, kats):
    """test/test_fcl_geometric_shapes.cpp shapeIntersectionGJK_* cases through
    collide() with penetration enabled and disabled: the hit boolean, the
    contact count and, where the reference checks it, the normal."""
    rng = np.random.default_rng(2)
    for case in kats["gjk_shape_intersection"]["cases"]:
        S = _prims(oracle, [tuple(case["s1"]), tuple(case["s2"])])
        for fname in case["frames"]:
            frame = _ident() if fname == "identity" else _random_tf(rng, 1000.0)
            tf1 = frame
            tf2 = _compose(frame, _ident(case["t2"]))
            for pen in (True, False):
                out, _ = S.collide([0], [1], tf1[None], tf2[None], max_contacts=1000, penetration=pen)
                assert bool(out["hit"][0]) == case["hit"], (case, fname, pen)
            if not case["hit"]:
                continue
            out, _ = S.collide([0], [1], tf1[None], tf2[None], max_contacts=1000, penetration=True)
            if "n_contacts" in case:
                assert out["n"][0] == case["n_contacts"], case
            if "normal" in case and fname in case.get("normal_frames", case["frames"]):
                want = frame[:9].reshape(3, 3) @ np.array(case["normal"], np.float64)
                tol = case.get("normal_tol", 1e-6)
                assert np.allclose(out["normal"][0], want, atol=tol), (case, out["normal"][0], want)


def _load_obj(path):
    v, f = [], []
    with open(path) as fh:
        for line in fh:
            p = line.split()
            if not p:
                continue
            if p[0] == "v":
                v.append([float(x) for x in p[1:4]])
            elif p[0] == "f":
                f.append([int(x.split("/")[0]) - 1 for x in p[1:4]])
    return np.array(v, np.float32), np.array(f, np.int32)


def test_mesh_env_rob_bvh_equals_brute_force(oracle):
    """test/test_fcl_collision.cpp:915-1059: collision of the env.obj and
    rob.obj test meshes under random poses; the OBBRSS traversal's complete
    colliding triangle-pair set equals the brute-force all-pairs set."""
    ve, fe = _load_obj(os.path.join(GOLDEN, "env.obj"))
    vr, fr = _load_obj(os.path.join(GOLDEN, "rob.obj"))
    assert len(fe) == 2180 and len(fr) == 216
    M = oracle.MeshSet(np.concatenate([ve, vr]), [0, len(ve)], [len(ve), len(vr)], np.concatenate([fe, fr]),
                       [0, len(fe)], [len(fe), len(fr)])
    rng = np.random.default_rng(3)
    lo, hi = ve.min(0), ve.max(0)
    n_hit = 0
    for i in range(24):
        tf_env = _ident()
        tf_rob = _random_tf(rng)
        tf_rob[9:] = rng.uniform(lo, hi)
        bvh, nb = M.collide_all(0, 1, tf_env, tf_rob)
        brute, nr = M.collide_brute(0, 1, tf_env, tf_rob)
        assert nb == nr
        assert set(map(tuple, bvh.tolist())) == set(map(tuple, brute.tolist()))
        n_hit += nb > 0
        # CollisionRequest(num_max_contacts = 1): first contact only, same boolean.
        out, _ = M.collide([0], [1], tf_env[None], tf_rob[None], max_contacts=1)
        assert bool(out["hit"][0]) == (nb > 0)
    assert n_hit >= 4, n_hit
