"""Pins of the oracle's mesh-shape leaf tests (GJKSolver::
shapeTriangleIntersect, include/fcl/narrowphase/detail/gjk_solver-inl.h:
545-563) against what the geometry fixes, away from touching contact:

- sphere: the exact distance from the centre to the triangle (closest
  point on a triangle, Ericson, Real-Time Collision Detection 5.1.5) below
  the radius;
- box: feasibility of "a point of the triangle inside the box" (a linear
  program over barycentric weights);
- convex hull (MPR with the triangle as a shape): feasibility of "a point of
  both" (a linear program over both sets of weights);
- capsule: the segment-triangle distance below the radius.
Cylinder-triangle runs the same MPR as the hull, on the cylinder support
pinned in test_oracle_pins.py.
"""
import numpy as np
import pytest
from scipy.optimize import linprog

from mfcl import workloads as W

BOX, SPHERE, CAPSULE, CONVEX = 0, 1, 3, 100


def _closest_on_triangle(p, a, b, c):
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = ab @ ap, ac @ ap
    if d1 <= 0 and d2 <= 0:
        return a
    bp = p - b
    d3, d4 = ab @ bp, ac @ bp
    if d3 >= 0 and d4 <= d3:
        return b
    vc = d1 * d4 - d3 * d2
    if vc <= 0 and d1 >= 0 and d3 <= 0:
        return a + ab * (d1 / (d1 - d3))
    cp = p - c
    d5, d6 = ab @ cp, ac @ cp
    if d6 >= 0 and d5 <= d6:
        return c
    vb = d5 * d2 - d1 * d6
    if vb <= 0 and d2 >= 0 and d6 <= 0:
        return a + ac * (d2 / (d2 - d6))
    va = d3 * d6 - d5 * d4
    if va <= 0 and (d4 - d3) >= 0 and (d5 - d6) >= 0:
        return b + (c - b) * ((d4 - d3) / ((d4 - d3) + (d5 - d6)))
    denom = 1.0 / (va + vb + vc)
    return a + ab * (vb * denom) + ac * (vc * denom)


def _seg_seg(p1, q1, p2, q2):
    d1, d2, r = q1 - p1, q2 - p2, p1 - p2
    a, e, f = d1 @ d1, d2 @ d2, d2 @ r
    c, b = d1 @ r, d1 @ d2
    den = a * e - b * b
    s = np.clip((b * f - c * e) / den, 0, 1) if den > 1e-300 else 0.0
    t = (b * s + f) / e
    if t < 0:
        t, s = 0.0, np.clip(-c / a, 0, 1)
    elif t > 1:
        t, s = 1.0, np.clip((b - c) / a, 0, 1)
    return np.linalg.norm((p1 + d1 * s) - (p2 + d2 * t))


def _seg_tri(p, q, a, b, c):
    n = np.cross(b - a, c - a)
    dp, dq = n @ (p - a), n @ (q - a)
    if dp * dq <= 0 and dp != dq:  # crosses the plane: inside the triangle?
        x = p + (q - p) * (dp / (dp - dq))
        if np.linalg.norm(_closest_on_triangle(x, a, b, c) - x) < 1e-12:
            return 0.0
    d = min(np.linalg.norm(_closest_on_triangle(p, a, b, c) - p), np.linalg.norm(_closest_on_triangle(q, a, b, c) - q))
    for u, v in ((a, b), (b, c), (c, a)):
        d = min(d, _seg_seg(p, q, u, v))
    return d


def _rot(q):
    w, x, y, z = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _tf(R, t):
    out = np.zeros(12)
    out[:9] = R.reshape(9)
    out[9:] = t
    return out


def _cases(rng, n):
    for _ in range(n):
        Rs, Rm = _rot(rng.normal(size=4)), _rot(rng.normal(size=4))
        ts, tm = rng.uniform(-0.3, 0.3, 3), rng.uniform(-0.3, 0.3, 3)
        tri = rng.uniform(-0.4, 0.4, (3, 3))
        yield Rs, ts, Rm, tm, tri


def test_sphere_triangle(oracle):
    rng = np.random.default_rng(80)
    S = oracle.ShapeSet(np.array([SPHERE], np.int32), np.array([[0.25, 0, 0]]))
    n_hit = n = 0
    for Rs, ts, Rm, tm, tri in _cases(rng, 3000):
        P = tri @ Rm.T + tm
        d = np.linalg.norm(_closest_on_triangle(ts, *P) - ts)
        if abs(d - 0.25) < 1e-9:
            continue
        got = S.triangle_hit(0, _tf(Rs, ts), tri.reshape(9), _tf(Rm, tm))
        assert got == (d < 0.25), (d, got)
        n_hit += got
        n += 1
    assert n_hit > 300 and n - n_hit > 300


def _lp_feasible(A_eq, b_eq, A_ub, b_ub, nvar):
    res = linprog(np.zeros(nvar), A_ub=A_ub, b_ub=b_ub, A_eq=A_eq, b_eq=b_eq, bounds=[(0, None)] * nvar,
                  method="highs")
    return res.status == 0


def _box_tri_lp(h, Rs, ts, P):
    # weights w (3): sum w = 1; |Rs^T (P^T w - ts)| <= h
    M = Rs.T @ P.T  # 3 x 3: column j = Rs^T P_j
    off = Rs.T @ ts
    A_ub = np.vstack([M, -M])
    b_ub = np.concatenate([h + off, h - off])
    return _lp_feasible(np.ones((1, 3)), np.array([1.0]), A_ub, b_ub, 3)


def test_box_triangle(oracle):
    rng = np.random.default_rng(81)
    side = np.array([0.5, 0.3, 0.2])
    S = oracle.ShapeSet(np.array([BOX], np.int32), side[None])
    n_hit = n = 0
    for Rs, ts, Rm, tm, tri in _cases(rng, 3000):
        P = tri @ Rm.T + tm
        inner = _box_tri_lp(side / 2 * (1 - 1e-7), Rs, ts, P)
        outer = _box_tri_lp(side / 2 * (1 + 1e-7), Rs, ts, P)
        if inner != outer:
            continue
        got = S.triangle_hit(0, _tf(Rs, ts), tri.reshape(9), _tf(Rm, tm))
        assert got == inner
        n_hit += got
        n += 1
    assert n_hit > 300 and n - n_hit > 300


def _hull_tri_lp(V, P):
    # weights a (len V), b (3): V^T a = P^T b, sum a = 1, sum b = 1
    m = len(V)
    A_eq = np.zeros((5, m + 3))
    A_eq[:3, :m] = V.T
    A_eq[:3, m:] = -P.T
    A_eq[3, :m] = 1
    A_eq[4, m:] = 1
    return _lp_feasible(A_eq, np.array([0, 0, 0, 1.0, 1.0]), None, None, m + 3)


def test_hull_triangle_mpr(oracle):
    rng = np.random.default_rng(82)
    v, off, cnt, _ = W.ellipsoid_hulls(rng, 6, 12, 60)
    hv, ho, hc, ht, hto, htc = [], [], [], [], [], []
    o = q = 0
    for i in range(6):
        Pp = v[off[i]:off[i] + cnt[i]]
        keep, t = W.hull_faces(Pp)
        hv.append(Pp[keep]); ho.append(o); hc.append(len(keep)); o += len(keep)
        ht.append(t); hto.append(q); htc.append(len(t)); q += len(t)
    S = oracle.ShapeSet(np.full(6, CONVEX, np.int32), np.zeros((6, 3)), np.concatenate(hv), ho, hc,
                        np.concatenate(ht), hto, htc)
    n_hit = n = 0
    for k, (Rs, ts, Rm, tm, tri) in enumerate(_cases(rng, 1200)):
        h = k % 6
        Vl = hv[h].astype(np.float64)
        c = Vl.mean(0)
        P = tri @ Rm.T + tm
        inner = _hull_tri_lp(((Vl - c) * (1 - 1e-6) + c) @ Rs.T + ts, P)
        outer = _hull_tri_lp(((Vl - c) * (1 + 1e-6) + c) @ Rs.T + ts, P)
        if inner != outer:
            continue
        got = S.triangle_hit(h, _tf(Rs, ts), tri.reshape(9), _tf(Rm, tm))
        assert got == inner, k
        n_hit += got
        n += 1
    assert n_hit > 100 and n - n_hit > 100


def test_capsule_triangle_mpr(oracle):
    rng = np.random.default_rng(83)
    r, lz = 0.1, 0.4
    S = oracle.ShapeSet(np.array([CAPSULE], np.int32), np.array([[r, lz, 0]]))
    n_hit = n = 0
    for Rs, ts, Rm, tm, tri in _cases(rng, 2000):
        P = tri @ Rm.T + tm
        a = ts + Rs @ np.array([0, 0, lz / 2])
        b = ts - Rs @ np.array([0, 0, lz / 2])
        d = _seg_tri(a, b, *P)
        if abs(d - r) < 1e-6:
            continue
        got = S.triangle_hit(0, _tf(Rs, ts), tri.reshape(9), _tf(Rm, tm))
        assert got == (d < r), (d, got)
        n_hit += got
        n += 1
    assert n_hit > 200 and n - n_hit > 200


# ---------------------------------------------------------------------------
# Penetration enabled (GJKSolver::shapeTriangleIntersect with contact
# outputs): the sphere's closed-form contact, and GJK + EPA on every other
# shape, pinned by what the geometry fixes.

def test_sphere_triangle_contact(oracle):
    """Sphere contact (sphere_triangle-inl.h:107-187): a hit when the
    exact centre-to-triangle distance is below the radius; pos is a point of
    the triangle within the radius of the centre; normal = (pos - centre)
    normalised and depth = -(r - |pos - centre|); when the centre projects
    inside the triangle, pos is the exact closest point.  A centre lying on
    the triangle gives the flipped plane normal and depth -r."""
    rng = np.random.default_rng(84)
    r = 0.25
    S = oracle.ShapeSet(np.array([SPHERE], np.int32), np.array([[r, 0, 0]]))
    n_hit = n_in = 0
    for Rs, ts, Rm, tm, tri in _cases(rng, 3000):
        P = tri @ Rm.T + tm
        q = _closest_on_triangle(ts, *P)
        d = np.linalg.norm(q - ts)
        if abs(d - r) < 1e-9:
            continue
        hit, pos, nrm, depth, _, _ = S.triangle_contact(0, _tf(Rs, ts), tri.reshape(9), _tf(Rm, tm))
        assert hit == (d < r)
        if not hit:
            continue
        n_hit += 1
        assert np.linalg.norm(_closest_on_triangle(pos, *P) - pos) < 1e-12
        dist = np.linalg.norm(pos - ts)
        assert dist < r + 1e-12
        assert abs(depth + (r - dist)) < 1e-12
        assert np.allclose(nrm, (pos - ts) / dist, atol=1e-12)
        n = np.cross(P[1] - P[0], P[2] - P[0])
        n /= np.linalg.norm(n)
        proj = ts - n * (n @ (ts - P[0]))
        if np.linalg.norm(_closest_on_triangle(proj, *P) - proj) < 1e-13:
            n_in += 1
            assert np.linalg.norm(pos - q) < 1e-12
    assert n_hit > 300 and n_in > 100
    tri0 = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float64)
    eye = np.eye(3)
    hit, pos, nrm, depth, _, _ = S.triangle_contact(0, _tf(eye, np.array([0.25, 0.25, 0.0])), tri0.reshape(9),
                                                    _tf(eye, np.zeros(3)))
    assert hit and depth == -r and np.array_equal(nrm, [0, 0, -1.0]) and np.array_equal(pos, [0.25, 0.25, 0.0])


def _mtd(A, B, axes):
    """Minimum overlap of point sets A and B over the unit axes (the
    penetration depth of their hulls when the axes are all facet normals of
    the Minkowski difference) and the axis that attains it."""
    best, arg = np.inf, None
    for a in axes:
        l = np.linalg.norm(a)
        if l < 1e-9:
            continue
        a = a / l
        pa, pb = A @ a, B @ a
        o = min(pa.max() - pb.min(), pb.max() - pa.min())
        if o < best:
            best, arg = o, a
    return best, arg


def _overlap_along(A, B, a):
    pa, pb = A @ a, B @ a
    return min(pa.max() - pb.min(), pb.max() - pa.min())


def _check_epa_contact(S, shape, Rs, ts, Rm, tm, tri, A, faces_n, edges):
    """EPA contact of shape (world vertices A, face normals, edge vectors)
    against the triangle: depth equals the minimum overlap over the facet
    normals of the Minkowski difference (shape faces, the triangle's normal,
    edge x edge), the contact normal is a direction of that overlap, pos is
    the witness midpoint between the two sets.  Returns whether it hit."""
    P = tri @ Rm.T + tm
    hit, pos, nrm, depth, gs, es = S.triangle_contact(shape, _tf(Rs, ts), tri.reshape(9), _tf(Rm, tm))
    if not hit:
        return False
    te = [P[1] - P[0], P[2] - P[1], P[0] - P[2]]
    axes = list(faces_n) + [np.cross(te[0], te[1])] + [np.cross(e, f) for e in edges for f in te]
    mtd, _ = _mtd(A, P, axes)
    scale = max(np.abs(A).max(), np.abs(P).max())
    assert mtd > -1e-9
    assert abs(depth - mtd) <= 1e-5 * scale + 1e-9, (depth, mtd)
    assert abs(np.linalg.norm(nrm) - 1) < 1e-12
    assert abs(_overlap_along(A, P, nrm) - depth) <= 1e-5 * scale + 1e-9
    return True


def test_box_triangle_contact_depth(oracle):
    rng = np.random.default_rng(85)
    side = np.array([0.5, 0.3, 0.2])
    S = oracle.ShapeSet(np.array([BOX], np.int32), side[None])
    corners = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)]) * side / 2
    n_hit = 0
    for Rs, ts, Rm, tm, tri in _cases(rng, 1500):
        A = corners @ Rs.T + ts
        n_hit += _check_epa_contact(S, 0, Rs, ts, Rm, tm, tri, A, list(Rs.T), list(Rs.T))
    assert n_hit > 200


def test_hull_triangle_contact_depth(oracle):
    rng = np.random.default_rng(86)
    v, off, cnt, _ = W.ellipsoid_hulls(rng, 6, 12, 48)
    hv, ho, hc, ht, hto, htc = [], [], [], [], [], []
    o = q = 0
    for i in range(6):
        Pp = v[off[i]:off[i] + cnt[i]]
        keep, t = W.hull_faces(Pp)
        hv.append(Pp[keep]); ho.append(o); hc.append(len(keep)); o += len(keep)
        ht.append(t); hto.append(q); htc.append(len(t)); q += len(t)
    S = oracle.ShapeSet(np.full(6, CONVEX, np.int32), np.zeros((6, 3)), np.concatenate(hv), ho, hc,
                        np.concatenate(ht), hto, htc)
    n_hit = 0
    for k, (Rs, ts, Rm, tm, tri) in enumerate(_cases(rng, 600)):
        h = k % 6
        Vw = hv[h].astype(np.float64) @ Rs.T + ts
        F = ht[h]
        fn = [np.cross(Vw[f[1]] - Vw[f[0]], Vw[f[2]] - Vw[f[0]]) for f in F]
        ed = {tuple(sorted((int(f[i]), int(f[(i + 1) % 3])))) for f in F for i in range(3)}
        edges = [Vw[b] - Vw[a] for a, b in ed]
        n_hit += _check_epa_contact(S, h, Rs, ts, Rm, tm, tri, Vw, fn, edges)
    assert n_hit > 60
