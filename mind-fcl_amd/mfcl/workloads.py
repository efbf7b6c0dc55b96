"""Seeded synthetic workloads for the five benchmark configurations.

All geometry and poses are generated in FP32 (the batched API's input
format); the FP64 oracle consumes exactly the same FP32 values.
Poses are (tx, ty, tz, qw, qx, qy, qz); rotations are uniform (Shoemake).
"""
from __future__ import annotations

import numpy as np

BOX, SPHERE, CAPSULE, CYLINDER, CONVEX = 0, 1, 3, 5, 100


def random_quats(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform unit quaternions (Shoemake 1992), returned as (w, x, y, z)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    x = a * np.sin(2 * np.pi * u2)
    y = a * np.cos(2 * np.pi * u2)
    z = b * np.sin(2 * np.pi * u3)
    w = b * np.cos(2 * np.pi * u3)
    return np.stack([w, x, y, z], axis=1)


def poses(rng: np.random.Generator, n: int, lo: float, hi: float) -> np.ndarray:
    p = np.empty((n, 7), np.float32)
    p[:, 0:3] = rng.uniform(lo, hi, size=(n, 3))
    p[:, 3:7] = random_quats(rng, n)
    return p


def primitive_pairs(n_pairs: int = 4096, seed: int = 1):
    """4096 pairs over {box, sphere, capsule, cylinder}; half-extents, radii
    and half-lengths U[0.05, 0.5] m, centres U[0, 1]^3 m, uniform rotations.
    Shape 2i is pair i's first shape, 2i+1 its second."""
    rng = np.random.default_rng(seed)
    n = 2 * n_pairs
    kinds = np.array([BOX, SPHERE, CAPSULE, CYLINDER], np.int32)
    types = kinds[rng.integers(0, 4, size=n)]
    half = rng.uniform(0.05, 0.5, size=(n, 3)).astype(np.float32)
    params = np.zeros((n, 4), np.float32)
    box = types == BOX
    params[box, 0:3] = 2.0 * half[box]
    sph = types == SPHERE
    params[sph, 0] = half[sph, 0]
    cap = (types == CAPSULE) | (types == CYLINDER)
    params[cap, 0] = half[cap, 0]          # radius
    params[cap, 1] = 2.0 * half[cap, 1]    # full length lz
    pa = np.arange(0, n, 2, dtype=np.uint32)
    pb = np.arange(1, n, 2, dtype=np.uint32)
    posea = poses(rng, n_pairs, 0.0, 1.0)
    poseb = poses(rng, n_pairs, 0.0, 1.0)
    return dict(types=types, params=params, pair_a=pa, pair_b=pb, pose_a=posea, pose_b=poseb)


def bounding_radius(types, params, vert_cnt=None, verts=None, vert_off=None):
    """Radius of a ball around each shape's local origin containing it."""
    t = np.asarray(types)
    p = np.asarray(params, np.float64)
    r = np.zeros(len(t))
    r[t == BOX] = 0.5 * np.linalg.norm(p[t == BOX, 0:3], axis=1)
    r[t == SPHERE] = p[t == SPHERE, 0]
    m = t == CAPSULE
    r[m] = p[m, 0] + 0.5 * p[m, 1]
    m = t == CYLINDER
    r[m] = np.hypot(p[m, 0], 0.5 * p[m, 1])
    if verts is not None:
        for i in np.nonzero(t == CONVEX)[0]:
            v = verts[vert_off[i]:vert_off[i] + vert_cnt[i]]
            r[i] = np.sqrt((v.astype(np.float64) ** 2).sum(1).max())
    return r


def ellipsoid_hulls(rng: np.random.Generator, n_hulls: int, nmin: int = 8, nmax: int = 512,
                    amin: float = 0.05, amax: float = 0.4, chunk: int = 1 << 20):
    """Vertex sets of hulls of N points on random ellipsoids: axes U[amin,
    amax], N = round(exp(U[ln nmin, ln nmax])), x = diag(a) u with u uniform
    on the unit sphere (Archimedes: z ~ U[-1, 1], phi ~ U[0, 2 pi)), so every
    point lies on the ellipsoid surface."""
    counts = np.rint(np.exp(rng.uniform(np.log(nmin), np.log(nmax), size=n_hulls))).astype(np.uint32)
    axes = rng.uniform(amin, amax, size=(n_hulls, 3)).astype(np.float32)
    off = np.zeros(n_hulls, np.uint64)
    off[1:] = np.cumsum(counts[:-1], dtype=np.uint64)
    total = int(counts.sum(dtype=np.uint64))
    verts = np.empty((total, 3), np.float32)
    owner = np.repeat(np.arange(n_hulls, dtype=np.int32), counts.astype(np.intp))
    # work buffers reused across chunks (fresh large temporaries are slow to fault in)
    z = np.empty(chunk, np.float32)
    phi = np.empty(chunk, np.float32)
    rho = np.empty(chunk, np.float32)
    t = np.empty(chunk, np.float32)
    a = np.empty((chunk, 3), np.float32)
    for s in range(0, total, chunk):
        e = min(total, s + chunk)
        m = e - s
        zz, pp, rr, tt, aa = z[:m], phi[:m], rho[:m], t[:m], a[:m]
        rng.random(out=zz, dtype=np.float32)
        rng.random(out=pp, dtype=np.float32)
        np.multiply(zz, np.float32(2.0), out=zz)
        np.subtract(zz, np.float32(1.0), out=zz)
        np.multiply(pp, np.float32(2.0 * np.pi), out=pp)
        np.multiply(zz, zz, out=rr)
        np.subtract(np.float32(1.0), rr, out=rr)
        np.maximum(rr, np.float32(0.0), out=rr)
        np.sqrt(rr, out=rr)
        np.take(axes, owner[s:e], axis=0, out=aa)
        np.cos(pp, out=tt)
        np.multiply(tt, rr, out=tt)
        np.multiply(tt, aa[:, 0], out=verts[s:e, 0])
        np.sin(pp, out=tt)
        np.multiply(tt, rr, out=tt)
        np.multiply(tt, aa[:, 1], out=verts[s:e, 1])
        np.multiply(zz, aa[:, 2], out=verts[s:e, 2])
    return verts, off, counts, axes


def morton_order(u: np.ndarray) -> np.ndarray:
    """Permutation sorting points of [-1, 1]^3 along a Morton (Z-order)
    curve of their 10-bit quantised coordinates."""
    q = np.clip(((u + 1.0) * 511.5).astype(np.int64), 0, 1023)
    code = np.zeros(len(u), np.int64)
    for b in range(10):
        for a in range(3):
            code |= ((q[:, a] >> b) & 1) << (3 * b + a)
    return np.argsort(code, kind="stable")


def template_hulls(rng: np.random.Generator, n_hulls: int, nmin: int = 8, nmax: int = 512, amin: float = 0.05,
                   amax: float = 0.4, variants: int = 0, vertex_order: str = "random"):
    """Convex hulls with faces: N = round(exp(U[ln nmin, ln nmax])) points
    uniform on the unit sphere, taken from `variants` template point sets per
    N (0: about n_hulls / (4 x distinct N), between 1 and 16) whose hull
    triangles come from qhull, then scaled by axes U[amin,
    amax] (a linear map keeps the hull's faces), so every point lies on an
    ellipsoid and is a hull vertex.  Hulls of one template share its
    triangles (tri_off / tri_cnt point into one array)."""
    counts = np.rint(np.exp(rng.uniform(np.log(nmin), np.log(nmax), size=n_hulls))).astype(np.uint32)
    axes = rng.uniform(amin, amax, size=(n_hulls, 3)).astype(np.float32)
    sizes = np.unique(counts)
    if variants <= 0:
        variants = int(min(16, max(1, n_hulls // (4 * len(sizes)))))
    t_pts, t_off, t_tris, t_toff, t_tcnt = [], [], [], [], []
    po = to = 0
    t_index = {}
    for c in sizes:
        for v in range(variants):
            u = rng.normal(size=(int(c), 3))
            u = (u / np.linalg.norm(u, axis=1, keepdims=True)).astype(np.float32)
            if vertex_order == "morton":
                u = u[morton_order(u)]
            from scipy.spatial import ConvexHull

            tri = ConvexHull(u.astype(np.float64)).simplices.astype(np.int32)
            t_index[(int(c), v)] = len(t_off)
            t_pts.append(u)
            t_off.append(po)
            t_tris.append(tri)
            t_toff.append(to)
            t_tcnt.append(len(tri))
            po += int(c)
            to += len(tri)
    t_pts = np.concatenate(t_pts)
    t_off = np.array(t_off, np.int64)
    which = rng.integers(0, variants, size=n_hulls)
    tid = np.array([t_index[(int(c), int(w))] for c, w in zip(counts, which)], np.int64)
    off = np.zeros(n_hulls, np.uint64)
    off[1:] = np.cumsum(counts[:-1], dtype=np.uint64)
    total = int(counts.sum(dtype=np.uint64))
    owner = np.repeat(np.arange(n_hulls, dtype=np.int64), counts.astype(np.intp))
    local = np.arange(total, dtype=np.int64) - np.repeat(off.astype(np.int64), counts.astype(np.intp))
    verts = t_pts[t_off[tid][owner] + local] * axes[owner]
    tri_off = np.array(t_toff, np.int64)[tid]
    tri_cnt = np.array(t_tcnt, np.int64)[tid]
    return verts.astype(np.float32), off, counts, axes, np.concatenate(t_tris), tri_off, tri_cnt


def hull_pairs(n_pairs: int = 1 << 20, seed: int = 2, nmin: int = 8, nmax: int = 512, faces: bool = False,
               vertex_order: str = "random"):
    """Batched convex-hull workload: 2 hulls per pair (shape 2i, 2i+1),
    centres U[0, 2]^3 m, uniform rotations.  faces=True: hulls from
    template_hulls, with triangles (keys tris, tri_off, tri_cnt)."""
    rng = np.random.default_rng(seed)
    n = 2 * n_pairs
    extra = {}
    if faces:
        verts, off, cnt, axes, tris, toff, tcnt = template_hulls(rng, n, nmin, nmax, vertex_order=vertex_order)
        extra = dict(tris=tris, tri_off=toff, tri_cnt=tcnt)
    else:
        verts, off, cnt, axes = ellipsoid_hulls(rng, n, nmin, nmax)
    types = np.full(n, CONVEX, np.int32)
    params = np.zeros((n, 4), np.float32)
    pa = np.arange(0, n, 2, dtype=np.uint32)
    pb = np.arange(1, n, 2, dtype=np.uint32)
    posea = poses(rng, n_pairs, 0.0, 2.0)
    poseb = poses(rng, n_pairs, 0.0, 2.0)
    return dict(types=types, params=params, verts=verts, vert_off=off, vert_cnt=cnt, axes=axes,
                pair_a=pa, pair_b=pb, pose_a=posea, pose_b=poseb, **extra)


def hull_point_pairs(n_pairs: int = 1 << 20, seed: int = 2, nmin: int = 8, nmax: int = 512):
    """The batched convex-hull workload with an independent point set per
    hull (shape 2i, 2i+1 of pair i): N log-uniform in [nmin, nmax] points on
    an ellipsoid with axes U[0.05, 0.4] m, centres U[0, 2]^3 m, uniform
    rotations.  The hulls (vertices and faces) are built from the points by
    the library (mfcl_upload_convex_points); verts/vert_off/vert_cnt hold
    the points."""
    rng = np.random.default_rng(seed)
    n = 2 * n_pairs
    verts, off, cnt, axes = ellipsoid_hulls(rng, n, nmin, nmax)
    types = np.full(n, CONVEX, np.int32)
    params = np.zeros((n, 4), np.float32)
    pa = np.arange(0, n, 2, dtype=np.uint32)
    pb = np.arange(1, n, 2, dtype=np.uint32)
    posea = poses(rng, n_pairs, 0.0, 2.0)
    poseb = poses(rng, n_pairs, 0.0, 2.0)
    return dict(types=types, params=params, verts=verts, vert_off=off, vert_cnt=cnt, axes=axes, pair_a=pa,
                pair_b=pb, pose_a=posea, pose_b=poseb)


def _hull_faces_range(args):
    verts, off, cnt, lo, hi = args
    return [hull_faces(verts[int(off[i]):int(off[i]) + int(cnt[i])]) for i in range(lo, hi)]


def hulls_of_sets(verts, off, cnt, n_sets: int, workers: int = 1):
    """qhull hulls of the first n_sets point sets, as shape arrays:
    (vertices, vert_off, vert_cnt, triangles, tri_off, tri_cnt), vertices in
    ascending point order per set.  workers > 1 runs qhull in that many
    forked processes (call before anything has touched the GPU)."""
    jobs = []
    step = max(1, (n_sets + 4 * max(workers, 1) - 1) // (4 * max(workers, 1)))
    for lo in range(0, n_sets, step):
        hi = min(n_sets, lo + step)
        o0 = int(off[lo])
        o1 = int(off[hi - 1]) + int(cnt[hi - 1])
        jobs.append((verts[o0:o1], np.asarray(off[lo:hi], np.int64) - o0, cnt[lo:hi], 0, hi - lo))
    if workers > 1:
        import multiprocessing as mp

        with mp.get_context("fork").Pool(workers) as pool:
            parts = pool.map(_hull_faces_range, jobs)
    else:
        parts = [_hull_faces_range(j) for j in jobs]
    hv, hvc, ht, htc = [], [], [], []
    k = 0
    for part in parts:
        for keep, tris in part:
            o = int(off[k])
            hv.append(verts[o + keep])
            hvc.append(len(keep))
            ht.append(tris)
            htc.append(len(tris))
            k += 1
    hvc = np.array(hvc, np.uint32)
    htc = np.array(htc, np.int64)
    hvo = np.zeros(n_sets, np.uint64)
    hvo[1:] = np.cumsum(hvc[:-1], dtype=np.uint64)
    hto = np.zeros(n_sets, np.int64)
    hto[1:] = np.cumsum(htc[:-1])
    return np.concatenate(hv), hvo, hvc, np.concatenate(ht), hto, htc


def hull_faces(v: np.ndarray):
    """Convex hull of a point set: (indices of hull vertices, triangles over
    those indices renumbered 0..k-1).  Uses qhull through scipy."""
    from scipy.spatial import ConvexHull

    h = ConvexHull(v.astype(np.float64))
    keep = np.unique(h.simplices)
    remap = -np.ones(len(v), np.int64)
    remap[keep] = np.arange(len(keep))
    return keep, remap[h.simplices].astype(np.int32)


def icosphere(freq: int):
    """Geodesic sphere: each icosahedron face split into freq^2 triangles,
    vertices projected to the unit sphere; T = 20 freq^2."""
    t = (1.0 + 5 ** 0.5) / 2
    V = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t],
                  [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    F = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
                  [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
                  [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]])
    n = freq
    ij = [(i, j) for i in range(n + 1) for j in range(n + 1 - i)]
    idx = {p: k for k, p in enumerate(ij)}
    loc_tris = []
    for i in range(n):
        for j in range(n - i):
            loc_tris.append((idx[(i, j)], idx[(i + 1, j)], idx[(i, j + 1)]))
            if j < n - i - 1:
                loc_tris.append((idx[(i + 1, j)], idx[(i + 1, j + 1)], idx[(i, j + 1)]))
    loc_tris = np.array(loc_tris, np.int64)
    bary = np.array(ij, np.float64) / n
    pts, tris = [], []
    for f, (a, b, c) in enumerate(F):
        P = V[a][None] * (1 - bary[:, 0:1] - bary[:, 1:2]) + V[b][None] * bary[:, 0:1] + V[c][None] * bary[:, 1:2]
        pts.append(P)
        tris.append(loc_tris + f * len(ij))
    P = np.concatenate(pts)
    T = np.concatenate(tris)
    P /= np.linalg.norm(P, axis=1, keepdims=True)
    key = np.round(P * 1e9).astype(np.int64)
    _, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
    inv = inv.reshape(-1)
    return P[first], inv[T].astype(np.int32)


def mesh_pool(n_meshes: int = 256, seed: int = 3, tmin: float = 1e3, tmax: float = 1e5):
    """Icospheres with T* log-uniform in [tmin, tmax] (freq = round(sqrt(T*/20))),
    radius U[0.1, 0.5] m, radial Gaussian noise with sigma = 2% of radius."""
    rng = np.random.default_rng(seed)
    tstar = np.exp(rng.uniform(np.log(tmin), np.log(tmax), size=n_meshes))
    freqs = np.maximum(1, np.rint(np.sqrt(tstar / 20.0))).astype(int)
    radii = rng.uniform(0.1, 0.5, size=n_meshes)
    cache = {}
    vs, ts = [], []
    for k in range(n_meshes):
        f = int(freqs[k])
        if f not in cache:
            cache[f] = icosphere(f)
        P, T = cache[f]
        rr = radii[k] * (1.0 + 0.02 * rng.standard_normal(len(P)))
        vs.append((P * rr[:, None]).astype(np.float32))
        ts.append(T)
    vcnt = np.array([len(v) for v in vs], np.uint32)
    tcnt = np.array([len(t) for t in ts], np.uint32)
    voff = np.zeros(n_meshes, np.uint64)
    toff = np.zeros(n_meshes, np.uint64)
    voff[1:] = np.cumsum(vcnt[:-1], dtype=np.uint64)
    toff[1:] = np.cumsum(tcnt[:-1], dtype=np.uint64)
    return dict(verts=np.concatenate(vs), tris=np.concatenate(ts), vert_off=voff, vert_cnt=vcnt,
                tri_off=toff, tri_cnt=tcnt, radius=radii, freq=freqs)


def mesh_pairs(pool: dict, n_pairs: int = 4096, seed: int = 30, spread: float = 1.75):
    """Pairs of pool meshes; centre distance d = U[0, spread] * (r1 + r2)
    along a random direction, so that about 40% of the pairs have
    intersecting surfaces (spheres intersect iff |r1 - r2| < d < r1 + r2)."""
    rng = np.random.default_rng(seed)
    m = len(pool["vert_cnt"])
    a = rng.integers(0, m, size=n_pairs).astype(np.uint32)
    b = rng.integers(0, m, size=n_pairs).astype(np.uint32)
    r = pool["radius"]
    d = rng.uniform(0.0, spread, size=n_pairs) * (r[a] + r[b])
    u = rng.standard_normal((n_pairs, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    pa = np.empty((n_pairs, 7), np.float32)
    pb = np.empty((n_pairs, 7), np.float32)
    pa[:, 0:3] = rng.uniform(0.0, 1.0, size=(n_pairs, 3))
    pa[:, 3:7] = random_quats(rng, n_pairs)
    pb[:, 0:3] = pa[:, 0:3] + u * d[:, None]
    pb[:, 3:7] = random_quats(rng, n_pairs)
    return dict(pair_a=a, pair_b=b, pose_a=pa, pose_b=pb)


def fbm_heightmap(rng, half: int = 128, octaves: int = 4, amplitude: float = 0.1) -> np.ndarray:
    """(2 half, 2 half) uint16 millimetre heights: fractal Brownian motion of
    value noise (octave o: a (4 2^o + 1)^2 lattice of U[0, 1] values,
    smoothstep-interpolated, weight 2^-o), normalised to [0, 1] and scaled by
    the amplitude in metres."""
    n = 2 * half
    acc = np.zeros((n, n))
    wsum = 0.0
    u = (np.arange(n) + 0.5) / n
    for o in range(octaves):
        g = 4 * 2 ** o
        lat = rng.uniform(0.0, 1.0, size=(g + 1, g + 1))
        x = u * g
        i = np.minimum(x.astype(np.int64), g - 1)
        f = x - i
        f = f * f * (3 - 2 * f)
        a = lat[i][:, i] * (1 - f)[None, :] + lat[i][:, i + 1] * f[None, :]
        b = lat[i + 1][:, i] * (1 - f)[None, :] + lat[i + 1][:, i + 1] * f[None, :]
        acc += (a * (1 - f)[:, None] + b * f[:, None]) * 0.5 ** o
        wsum += 0.5 ** o
    acc = (acc - acc.min()) / (acc.max() - acc.min())
    return np.rint(acc * amplitude * 1000.0).astype(np.uint16)


def heightmap_box_params(heights_mm: np.ndarray, res_x: float, res_y: float) -> np.ndarray:
    """Box shape parameters (full sides res_x, res_y, height) of every pixel,
    in pixel order y * full_x + x (the map's pixel boxes, constructBox)."""
    h = np.asarray(heights_mm).reshape(-1).astype(np.float64) * 0.001
    out = np.zeros((len(h), 4), np.float32)
    out[:, 0], out[:, 1], out[:, 2] = np.float32(res_x), np.float32(res_y), h.astype(np.float32)
    return out


def arm_scene(n_configs: int = 65536, seed: int = 6, n_links: int = 7, n_boxes: int = 512, heightmap: bool = False):
    """Robot-arm configuration workload: a serial chain of n_links convex
    links (64..128 hull vertices each, with faces; link length U[0.1, 0.4] m,
    hull radius U[0.03, 0.08] m about the link's long axis), n_boxes
    obstacle boxes (half-extents U[0.02, 0.3] m, centres uniform in a 2 m
    cube whose near face is 0.5 m in front of the arm's base, i.e.
    [0.5, 2.5] x [-1, 1] x [-1, 1] m, uniform rotations) and n_configs joint
    vectors U[-pi, pi].  With the cube around the base every configuration
    would collide; in front of it about a third do.  Shapes: the links first
    (ids 0..n_links-1), then the boxes.

    heightmap=True adds a 256 x 256 map of 1 cm pixels with fBm heights (4
    octaves, 0.1 m amplitude) whose frame sits 0.45 m below the base, shifted
    0.3 m along x and turned 0.3 rad about z, so the terrain's top lies
    between 0.45 m and 0.35 m below the base; its 65,536 pixel boxes follow
    the obstacle boxes in the shape table."""
    rng = np.random.default_rng(seed)
    lens = rng.uniform(0.1, 0.4, size=n_links).astype(np.float32)
    rad = rng.uniform(0.03, 0.08, size=n_links).astype(np.float32)
    verts, off, cnt, taxes, tris, toff, tcnt = template_hulls(rng, n_links, 64, 128, variants=1)
    # template_hulls scales unit-sphere template points by random axes (a
    # linear map, so the faces stay those of the hull); map them back and on
    # to the link's own axes instead
    axes = np.stack([lens / 2, rad, rad], 1).astype(np.float64)
    owner = np.repeat(np.arange(n_links), cnt.astype(np.intp))
    verts = (verts.astype(np.float64) / taxes[owner].astype(np.float64) * axes[owner]).astype(np.float32)
    link_radius = np.array([np.nextafter(np.float32(np.sqrt((verts[owner == i].astype(np.float64) ** 2).sum(1).max())),
                                         np.float32(np.inf)) for i in range(n_links)], np.float32)
    half = rng.uniform(0.02, 0.3, size=(n_boxes, 3))
    centres = rng.uniform(-1.0, 1.0, size=(n_boxes, 3)) + np.array([1.5, 0.0, 0.0])
    quats = random_quats(rng, n_boxes)
    box_pose = np.concatenate([centres, quats], 1).astype(np.float32)
    types = np.concatenate([np.full(n_links, CONVEX), np.full(n_boxes, BOX)]).astype(np.int32)
    params = np.zeros((n_links + n_boxes, 4), np.float32)
    params[n_links:, :3] = 2.0 * half
    vert_off = np.concatenate([off, np.zeros(n_boxes, np.uint64)])
    vert_cnt = np.concatenate([cnt, np.zeros(n_boxes, np.uint32)])
    tri_off = np.concatenate([toff, np.zeros(n_boxes, np.int64)])
    tri_cnt = np.concatenate([tcnt, np.zeros(n_boxes, np.int64)])
    q = rng.uniform(-np.pi, np.pi, size=(n_configs, n_links)).astype(np.float32)
    out = dict(types=types, params=params, verts=verts, vert_off=vert_off, vert_cnt=vert_cnt, tris=tris,
               tri_off=tri_off, tri_cnt=tri_cnt, link_len=lens, link_radius=link_radius, box_pose=box_pose,
               box_half=half.astype(np.float32), q=q, n_links=n_links, n_boxes=n_boxes)
    if heightmap:
        hm = fbm_heightmap(np.random.default_rng(seed + 1000))
        res = 0.01
        npix = hm.size
        out["hm_heights"] = hm
        out["hm_res"] = res
        out["hm_box_shape0"] = n_links + n_boxes
        out["hm_pose7"] = np.array([0.3, 0.0, -0.45, np.cos(0.15), 0.0, 0.0, np.sin(0.15)], np.float32)
        out["types"] = np.concatenate([types, np.full(npix, BOX, np.int32)])
        out["params"] = np.concatenate([params, heightmap_box_params(hm, res, res)])
        out["vert_off"] = np.concatenate([vert_off, np.zeros(npix, np.uint64)])
        out["vert_cnt"] = np.concatenate([vert_cnt, np.zeros(npix, np.uint32)])
        out["tri_off"] = np.concatenate([tri_off, np.zeros(npix, np.int64)])
        out["tri_cnt"] = np.concatenate([tri_cnt, np.zeros(npix, np.int64)])
    return out


def box_aabbs(box_pose: np.ndarray, half: np.ndarray):
    """World AABBs of oriented boxes (FP32 corners of |R| h around the centre,
    rounded outward by one ulp)."""
    q = box_pose[:, 3:7].astype(np.float64)
    q = q / np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
                  2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                  2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], 1).reshape(-1, 3, 3)
    ext = np.abs(R) @ half.astype(np.float64)[:, :, None]
    c = box_pose[:, :3].astype(np.float64)
    mn = np.nextafter((c - ext[:, :, 0]).astype(np.float32), np.float32(-np.inf))
    mx = np.nextafter((c + ext[:, :, 0]).astype(np.float32), np.float32(np.inf))
    return mn, mx
