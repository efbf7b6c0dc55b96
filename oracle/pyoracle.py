"""ctypes wrapper of the CPU oracle (oracle/_build/liborc.so).

TEST INFRASTRUCTURE ONLY: the FP64 restatement of mind-fcl's narrowphase,
OBBRSS traversal and broadphase used as the parity checker by tests/,
__graft_entry__.smoke() and the cpu_baseline leg of bench.py.  Product code
never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liborc.so")

MESH_CONTACT_REC = np.dtype([("b1", np.int32), ("b2", np.int32), ("pos", np.float64, 3), ("normal", np.float64, 3),
                             ("depth", np.float64)])
COLLIDE_REC = np.dtype(
    [("hit", np.int32), ("n", np.int32), ("method", np.int32), ("gjk_status", np.int32), ("epa_status", np.int32),
     ("gjk_iters", np.int32), ("epa_iters", np.int32), ("pad", np.int32), ("depth", np.float64),
     ("normal", np.float64, 3), ("pos", np.float64, 3)]
)
CONTACT_REC = np.dtype([("normal", np.float64, 3), ("pos", np.float64, 3), ("depth", np.float64)])
SD_REC = np.dtype(
    [("ok", np.int32), ("gjk_status", np.int32), ("epa_status", np.int32), ("gjk_iters", np.int32),
     ("epa_iters", np.int32), ("support_calls", np.int32), ("distance", np.float64), ("pa", np.float64, 3),
     ("pb", np.float64, 3)]
)

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        p, i64, i32, d = C.c_void_p, C.c_int64, C.c_int, C.c_double
        L.orc_hardware_threads.restype = i32
        L.orc_pose_from_quat.argtypes = [p, i64, p]
        L.orc_hullset_create.argtypes = [p, p, p, p, p, p, i32, i32]
        L.orc_hullset_create.restype = p
        L.orc_hullset_destroy.argtypes = [p]
        L.orc_hull_walks.argtypes = [p, i32]
        L.orc_support.argtypes = [p, p, p, p, i32, p, p]
        L.orc_collide_batch.argtypes = [p, p, p, p, p, p, p, p, i64, i32, i32, p, p, i32, p]
        L.orc_collide_batch.restype = d
        L.orc_collide_contacts_batch.argtypes = [p, p, p, p, p, p, p, p, i64, i32, i32, p, p, i32, p, p]
        L.orc_collide_contacts_batch.restype = d
        L.orc_mpr_penetration_batch.argtypes = [p, p, p, p, p, p, p, p, p, i64, i32, p, d, i32, i32, p]
        L.orc_mpr_penetration_batch.restype = d
        L.orc_mesh_shape_all.argtypes = [p, p, p, p, p, i32, i32, p, p, p, p, i64]
        L.orc_mesh_shape_all.restype = i64
        L.orc_mesh_shape_batch.argtypes = [p, p, p, p, p, p, p, p, p, i64, p, i32, p, p]
        L.orc_mesh_shape_batch.restype = d
        L.orc_shape_triangle_hit.argtypes = [p, p, p, p, i32, p, p, p, p]
        L.orc_shape_triangle_hit.restype = i32
        L.orc_mesh_shape_contacts_all.argtypes = [p, p, p, p, p, i32, i32, p, p, p, p, i64]
        L.orc_mesh_shape_contacts_all.restype = i64
        L.orc_shape_triangle_contact.argtypes = [p, p, p, p, i32, p, p, p, p, p]
        L.orc_shape_triangle_contact.restype = i32
        L.orc_signed_distance_batch.argtypes = [p, p, p, p, p, p, p, p, i64, p, p, i32, p]
        L.orc_signed_distance_batch.restype = d
        L.orc_pair_algorithm.argtypes = [p, p, p, p, i32, i32, p, p, i32, i32, d, p, p]
        L.orc_pair_algorithm.restype = i32
        L.orc_capsule_capsule_distance.argtypes = [d, d, p, d, d, p]
        L.orc_capsule_capsule_distance.restype = d
        L.orc_meshset_create.argtypes = [p, p, p, p, p, p, i32, i32]
        L.orc_meshset_create.restype = p
        L.orc_meshset_destroy.argtypes = [p]
        L.orc_mesh_num_nodes.argtypes = [p, i32]
        L.orc_mesh_collide_batch.argtypes = [p, p, p, p, p, i64, i32, i32, p, p, p, p, p]
        L.orc_mesh_collide_batch.restype = d
        L.orc_mesh_collide_all.argtypes = [p, i32, i32, p, p, p, i64]
        L.orc_mesh_collide_all.restype = i64
        L.orc_mesh_collide_brute.argtypes = [p, i32, i32, p, p, p, i64]
        L.orc_mesh_collide_brute.restype = i64
        L.orc_mesh_collide_pen_batch.argtypes = [p, p, p, p, p, i64, i32, i32, p, p]
        L.orc_mesh_collide_pen_batch.restype = d
        L.orc_mesh_collide_all_pen.argtypes = [p, i32, i32, p, p, p, i64]
        L.orc_mesh_collide_all_pen.restype = i64
        L.orc_tri_tri_contacts.argtypes = [p, p, p, p, p, p, p, p]
        L.orc_tri_tri_contacts.restype = i32
        L.orc_tri_tri.argtypes = [p, p, p, p]
        L.orc_obb_disjoint.argtypes = [p, p, p, p]
        L.orc_aabbtree_create.argtypes = [p, p, p, i32]
        L.orc_aabbtree_create.restype = p
        L.orc_aabbtree_destroy.argtypes = [p]
        L.orc_aabbtree_query_batch.argtypes = [p, p, p, i64, i32, i32, p, p, p]
        L.orc_aabbtree_query_batch.restype = d
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def hardware_threads() -> int:
    return int(lib().orc_hardware_threads())


def pose12(pose7: np.ndarray) -> np.ndarray:
    """FP32 (t, q) poses -> FP64 row-major R (9) + t (3), same formula as the device."""
    pose7 = np.ascontiguousarray(pose7, np.float32).reshape(-1, 7)
    out = np.zeros((len(pose7), 12), np.float64)
    lib().orc_pose_from_quat(_p(pose7), len(pose7), _p(out))
    return out


def mat_to_pose12(R: np.ndarray, t: np.ndarray) -> np.ndarray:
    out = np.zeros(12, np.float64)
    out[:9] = np.asarray(R, np.float64).reshape(9)
    out[9:] = t
    return out


class ShapeSet:
    """Shapes as the oracle sees them; convex hulls need faces for the
    reference's vertex-adjacency support search (pass tris=None for a
    naive-scan hull)."""

    def __init__(self, types, params, verts=None, vert_off=None, vert_cnt=None, tris=None, tri_off=None,
                 tri_cnt=None, hull_of_shape=None, threads: int = 0):
        self.types = np.ascontiguousarray(types, np.int32)
        p = np.asarray(params, np.float64).reshape(len(self.types), -1)
        self.params = np.ascontiguousarray(p[:, :3])
        n = len(self.types)
        self.hull_idx = np.full(n, -1, np.int32) if hull_of_shape is None else np.ascontiguousarray(hull_of_shape, np.int32)
        self.hs = None
        if verts is not None:
            self._verts = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
            self._voff = np.ascontiguousarray(vert_off, np.int64)
            self._vcnt = np.ascontiguousarray(vert_cnt, np.int32)
            if tris is not None:
                self._tris = np.ascontiguousarray(tris, np.int32).reshape(-1, 3)
                self._toff = np.ascontiguousarray(tri_off, np.int64)
                self._tcnt = np.ascontiguousarray(tri_cnt, np.int32)
            else:
                self._tris = self._toff = self._tcnt = None
            nh = len(self._vcnt)
            self.hs = lib().orc_hullset_create(_p(self._verts), _p(self._voff), _p(self._vcnt), _p(self._tris),
                                               _p(self._toff), _p(self._tcnt), nh, threads or hardware_threads())
            if hull_of_shape is None:
                self.hull_idx = np.where(self.types == 100, np.arange(n), -1).astype(np.int32)

    def __del__(self):
        if getattr(self, "hs", None):
            lib().orc_hullset_destroy(self.hs)
            self.hs = None

    def walks(self, h: int) -> bool:
        return bool(lib().orc_hull_walks(self.hs, h))

    def support(self, i: int, d) -> np.ndarray:
        d = np.ascontiguousarray(d, np.float64)
        out = np.zeros(3)
        lib().orc_support(self.hs, _p(self.types), _p(self.params), _p(self.hull_idx), i, _p(d), _p(out))
        return out

    def collide(self, pa, pb, tfa12, tfb12, max_contacts=1, penetration=True, solver=None, guess=None,
                threads=1):
        pa = np.ascontiguousarray(pa, np.int32)
        pb = np.ascontiguousarray(pb, np.int32)
        tfa12 = np.ascontiguousarray(tfa12, np.float64).reshape(-1, 12)
        tfb12 = np.ascontiguousarray(tfb12, np.float64).reshape(-1, 12)
        sp = None if solver is None else np.ascontiguousarray(solver, np.float64)
        g = None if guess is None else np.ascontiguousarray(guess, np.float64).reshape(-1, 3)
        out = np.zeros(len(pa), COLLIDE_REC)
        secs = lib().orc_collide_batch(self.hs, _p(self.types), _p(self.params), _p(self.hull_idx), _p(pa), _p(pb),
                                       _p(tfa12), _p(tfb12), len(pa), max_contacts, 1 if penetration else 0, _p(sp),
                                       _p(g), threads, _p(out))
        return out, secs

    def collide_contacts(self, pa, pb, tfa12, tfb12, max_contacts=4, penetration=True, solver=None, threads=1):
        """As collide(), plus every kept contact: (n, 4) records of
        CONTACT_REC in CollisionResult order (entries past out['n'] zero)."""
        pa = np.ascontiguousarray(pa, np.int32)
        pb = np.ascontiguousarray(pb, np.int32)
        tfa12 = np.ascontiguousarray(tfa12, np.float64).reshape(-1, 12)
        tfb12 = np.ascontiguousarray(tfb12, np.float64).reshape(-1, 12)
        sp = None if solver is None else np.ascontiguousarray(solver, np.float64)
        out = np.zeros(len(pa), COLLIDE_REC)
        cont = np.zeros((len(pa), 4), CONTACT_REC)
        lib().orc_collide_contacts_batch(self.hs, _p(self.types), _p(self.params), _p(self.hull_idx), _p(pa),
                                         _p(pb), _p(tfa12), _p(tfb12), len(pa), max_contacts,
                                         1 if penetration else 0, _p(sp), None, threads, _p(out), _p(cont))
        return out, cont

    def directed_penetration(self, pa, pb, tfa12, tfb12, dirs, max_contacts=1, tol=1e-6, solver=None, threads=1,
                             incremental=False):
        """collisionPenetration with an MPR mode (directed, or incremental
        minimum penetration from the hint dirs): COLLIDE_REC per pair
        (epa_status = the MPR status, epa_iters its iterations; depth -1
        when it fails)."""
        pa = np.ascontiguousarray(pa, np.int32)
        pb = np.ascontiguousarray(pb, np.int32)
        tfa12 = np.ascontiguousarray(tfa12, np.float64).reshape(-1, 12)
        tfb12 = np.ascontiguousarray(tfb12, np.float64).reshape(-1, 12)
        dirs = np.ascontiguousarray(dirs, np.float64).reshape(-1, 3)
        sp = None if solver is None else np.ascontiguousarray(solver, np.float64)
        out = np.zeros(len(pa), COLLIDE_REC)
        secs = lib().orc_mpr_penetration_batch(self.hs, _p(self.types), _p(self.params), _p(self.hull_idx), _p(pa),
                                               _p(pb), _p(tfa12), _p(tfb12), _p(dirs), len(pa), max_contacts, _p(sp),
                                               tol, 1 if incremental else 0, threads, _p(out))
        return out, secs

    def incremental_penetration(self, pa, pb, tfa12, tfb12, hints, max_contacts=1, tol=1e-6, solver=None,
                                threads=1):
        return self.directed_penetration(pa, pb, tfa12, tfb12, hints, max_contacts, tol, solver, threads, True)

    def triangle_hit(self, shape, tfs12, tri9, tfm12, solver=None) -> bool:
        """Leaf test of mesh-shape collision: shape posed tfs12 vs the
        triangle tri9 (3 x 3) posed tfm12."""
        sp = None if solver is None else np.ascontiguousarray(solver, np.float64)
        return bool(lib().orc_shape_triangle_hit(self.hs, _p(self.types), _p(self.params), _p(self.hull_idx),
                                                 int(shape), _p(np.ascontiguousarray(tfs12, np.float64)),
                                                 _p(np.ascontiguousarray(tri9, np.float64)),
                                                 _p(np.ascontiguousarray(tfm12, np.float64)), _p(sp)))

    def mesh_shape(self, M: "MeshSet", mesh_ids, shape_ids, tfm12, tfs12, solver=None, threads=1):
        """Mesh-shape collision (penetration disabled): hit flags and the
        number of colliding triangles per query."""
        mesh_ids = np.ascontiguousarray(mesh_ids, np.int32)
        shape_ids = np.ascontiguousarray(shape_ids, np.int32)
        tfm12 = np.ascontiguousarray(tfm12, np.float64).reshape(-1, 12)
        tfs12 = np.ascontiguousarray(tfs12, np.float64).reshape(-1, 12)
        sp = None if solver is None else np.ascontiguousarray(solver, np.float64)
        n = len(mesh_ids)
        hit = np.zeros(n, np.int32)
        cnt = np.zeros(n, np.int64)
        lib().orc_mesh_shape_batch(M.h, self.hs, _p(self.types), _p(self.params), _p(self.hull_idx), _p(mesh_ids),
                                   _p(shape_ids), _p(tfm12), _p(tfs12), n, _p(sp), threads, _p(hit), _p(cnt))
        return hit, cnt

    def mesh_shape_all(self, M: "MeshSet", mesh, shape, tfm12, tfs12, solver=None, cap=1 << 20):
        """Every colliding triangle id (ascending) of one mesh-shape query."""
        sp = None if solver is None else np.ascontiguousarray(solver, np.float64)
        out = np.zeros(cap, np.int32)
        n = lib().orc_mesh_shape_all(M.h, self.hs, _p(self.types), _p(self.params), _p(self.hull_idx), int(mesh),
                                     int(shape), _p(np.ascontiguousarray(tfm12, np.float64)),
                                     _p(np.ascontiguousarray(tfs12, np.float64)), _p(sp), _p(out), cap)
        return out[:min(n, cap)], int(n)

    def triangle_contact(self, shape, tfs12, tri9, tfm12, solver=None):
        """Leaf test of mesh-shape collision with penetration enabled: (hit,
        pos, normal before the traversal's reversal, depth, GJK status, EPA
        status)."""
        sp = None if solver is None else np.ascontiguousarray(solver, np.float64)
        r = np.zeros(9)
        hit = lib().orc_shape_triangle_contact(self.hs, _p(self.types), _p(self.params), _p(self.hull_idx),
                                               int(shape), _p(np.ascontiguousarray(tfs12, np.float64)),
                                               _p(np.ascontiguousarray(tri9, np.float64)),
                                               _p(np.ascontiguousarray(tfm12, np.float64)), _p(sp), _p(r))
        return bool(hit), r[0:3].copy(), r[3:6].copy(), float(r[6]), int(r[7]), int(r[8])

    def mesh_shape_contacts_all(self, M: "MeshSet", mesh, shape, tfm12, tfs12, solver=None, cap=1 << 16):
        """Every contact of one penetration-enabled mesh-shape query
        (MESH_CONTACT_REC, ascending triangle id, normals reversed as the
        traversal writes them) and the total count."""
        sp = None if solver is None else np.ascontiguousarray(solver, np.float64)
        out = np.zeros(cap, MESH_CONTACT_REC)
        n = lib().orc_mesh_shape_contacts_all(M.h, self.hs, _p(self.types), _p(self.params), _p(self.hull_idx),
                                              int(mesh), int(shape), _p(np.ascontiguousarray(tfm12, np.float64)),
                                              _p(np.ascontiguousarray(tfs12, np.float64)), _p(sp), _p(out), cap)
        return out[:min(n, cap)], int(n)

    def signed_distance(self, pa, pb, tfa12, tfb12, solver=None, guess=None, threads=1):
        pa = np.ascontiguousarray(pa, np.int32)
        pb = np.ascontiguousarray(pb, np.int32)
        tfa12 = np.ascontiguousarray(tfa12, np.float64).reshape(-1, 12)
        tfb12 = np.ascontiguousarray(tfb12, np.float64).reshape(-1, 12)
        sp = None if solver is None else np.ascontiguousarray(solver, np.float64)
        g = None if guess is None else np.ascontiguousarray(guess, np.float64).reshape(-1, 3)
        out = np.zeros(len(pa), SD_REC)
        secs = lib().orc_signed_distance_batch(self.hs, _p(self.types), _p(self.params), _p(self.hull_idx), _p(pa),
                                               _p(pb), _p(tfa12), _p(tfb12), len(pa), _p(sp), _p(g), threads, _p(out))
        return out, secs

    def pair_algorithm(self, ia, ib, tfa12, tfb12, mode, max_iter, tol, guess=None):
        """mode 0: GJK boolean status, 1: MPR status, 2: GJK distance ->
        (status, distance, p0, p1)."""
        tfa12 = np.ascontiguousarray(tfa12, np.float64)
        tfb12 = np.ascontiguousarray(tfb12, np.float64)
        g = None if guess is None else np.ascontiguousarray(guess, np.float64)
        d = np.zeros(7)
        st = lib().orc_pair_algorithm(self.hs, _p(self.types), _p(self.params), _p(self.hull_idx), ia, ib,
                                      _p(tfa12), _p(tfb12), mode, max_iter, tol, _p(g), _p(d))
        return st, d[0], d[1:4], d[4:7]


def collide_contacts(S: "ShapeSet", tfa12, tfb12, max_contacts: int):
    """Every contact CollisionRequest(max_contacts) keeps on one pair
    (shapes 0 and 1 of S)."""
    out, cont = S.collide_contacts([0], [1], np.asarray(tfa12)[None], np.asarray(tfb12)[None], max_contacts)
    return cont[0, :out["n"][0]]


def capsule_capsule_distance(r1, lz1, tf1, r2, lz2, tf2) -> float:
    return float(lib().orc_capsule_capsule_distance(r1, lz1, _p(np.ascontiguousarray(tf1, np.float64)), r2, lz2,
                                                    _p(np.ascontiguousarray(tf2, np.float64))))


class MeshSet:
    def __init__(self, verts, vert_off, vert_cnt, tris, tri_off, tri_cnt, threads: int = 0):
        self._v = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
        self._vo = np.ascontiguousarray(vert_off, np.int64)
        self._vc = np.ascontiguousarray(vert_cnt, np.int32)
        self._t = np.ascontiguousarray(tris, np.int32).reshape(-1, 3)
        self._to = np.ascontiguousarray(tri_off, np.int64)
        self._tc = np.ascontiguousarray(tri_cnt, np.int32)
        self.h = lib().orc_meshset_create(_p(self._v), _p(self._vo), _p(self._vc), _p(self._t), _p(self._to),
                                          _p(self._tc), len(self._vc), threads or hardware_threads())

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_meshset_destroy(self.h)
            self.h = None

    def collide(self, pa, pb, tfa12, tfb12, max_contacts=16, threads=1):
        pa = np.ascontiguousarray(pa, np.int32)
        pb = np.ascontiguousarray(pb, np.int32)
        tfa12 = np.ascontiguousarray(tfa12, np.float64).reshape(-1, 12)
        tfb12 = np.ascontiguousarray(tfb12, np.float64).reshape(-1, 12)
        n = len(pa)
        hit = np.zeros(n, np.int32)
        nc = np.zeros(n, np.int32)
        cont = np.full((n, max(1, max_contacts), 2), -1, np.int32)
        bv = np.zeros(n, np.int64)
        lf = np.zeros(n, np.int64)
        secs = lib().orc_mesh_collide_batch(self.h, _p(pa), _p(pb), _p(tfa12), _p(tfb12), n, max_contacts, threads,
                                            _p(hit), _p(nc), _p(cont), _p(bv), _p(lf))
        return dict(hit=hit, n=nc, contacts=cont, bv_tests=bv, leaf_tests=lf), secs

    def collide_pen(self, pa, pb, tfa12, tfb12, max_contacts=16, threads=1):
        """Penetration-enabled collide: (n_contacts[n], contacts[n, max_contacts]
        of MESH_CONTACT_REC in CollisionResult order), seconds."""
        pa = np.ascontiguousarray(pa, np.int32)
        pb = np.ascontiguousarray(pb, np.int32)
        tfa12 = np.ascontiguousarray(tfa12, np.float64).reshape(-1, 12)
        tfb12 = np.ascontiguousarray(tfb12, np.float64).reshape(-1, 12)
        n = len(pa)
        nc = np.zeros(n, np.int32)
        cont = np.zeros((n, max(1, max_contacts)), MESH_CONTACT_REC)
        secs = lib().orc_mesh_collide_pen_batch(self.h, _p(pa), _p(pb), _p(tfa12), _p(tfb12), n, max_contacts,
                                                threads, _p(nc), _p(cont))
        return nc, cont, secs

    def collide_all_pen(self, ia, ib, tfa12, tfb12, cap=1 << 16):
        out = np.zeros(cap, MESH_CONTACT_REC)
        n = lib().orc_mesh_collide_all_pen(self.h, ia, ib, _p(np.ascontiguousarray(tfa12, np.float64)),
                                           _p(np.ascontiguousarray(tfb12, np.float64)), _p(out), cap)
        return out[:min(n, cap)], int(n)

    def collide_all(self, ia, ib, tfa12, tfb12, cap=1 << 20):
        out = np.zeros((cap, 2), np.int32)
        n = lib().orc_mesh_collide_all(self.h, ia, ib, _p(np.ascontiguousarray(tfa12, np.float64)),
                                       _p(np.ascontiguousarray(tfb12, np.float64)), _p(out), cap)
        return out[:min(n, cap)], int(n)

    def collide_brute(self, ia, ib, tfa12, tfb12, cap=1 << 20):
        """All colliding triangle pairs by exhaustive search (no BVH)."""
        out = np.zeros((cap, 2), np.int32)
        n = lib().orc_mesh_collide_brute(self.h, ia, ib, _p(np.ascontiguousarray(tfa12, np.float64)),
                                         _p(np.ascontiguousarray(tfb12, np.float64)), _p(out), cap)
        return out[:min(n, cap)], int(n)


def tri_tri_contacts(P9, Q9, R9, T3):
    """(hit, points[npts, 3], depth, normal) of intersect_Triangle with
    contacts (Q in its own frame, mapped into P's by R, T)."""
    a = [np.ascontiguousarray(x, np.float64) for x in (P9, Q9, R9, T3)]
    pts = np.zeros(6)
    npts = C.c_int(0)
    depth = C.c_double(0)
    nrm = np.zeros(3)
    hit = lib().orc_tri_tri_contacts(*[_p(x) for x in a], _p(pts), C.byref(npts), C.byref(depth), _p(nrm))
    return bool(hit), pts[:3 * npts.value].reshape(-1, 3), depth.value, nrm


def tri_tri(P9, Q9, R9, T3) -> bool:
    a = [np.ascontiguousarray(x, np.float64) for x in (P9, Q9, R9, T3)]
    return bool(lib().orc_tri_tri(*[_p(x) for x in a]))


def obb_disjoint(B9, T3, a3, b3) -> bool:
    a = [np.ascontiguousarray(x, np.float64) for x in (B9, T3, a3, b3)]
    return bool(lib().orc_obb_disjoint(*[_p(x) for x in a]))


class AABBTree:
    def __init__(self, mn, mx, ids):
        self._mn = np.ascontiguousarray(mn, np.float32).reshape(-1, 3)
        self._mx = np.ascontiguousarray(mx, np.float32).reshape(-1, 3)
        self._ids = np.ascontiguousarray(ids, np.int64)
        self.h = lib().orc_aabbtree_create(_p(self._mn), _p(self._mx), _p(self._ids), len(self._ids))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_aabbtree_destroy(self.h)
            self.h = None

    def query(self, qmn, qmx, cap=64, threads=1):
        qmn = np.ascontiguousarray(qmn, np.float32).reshape(-1, 3)
        qmx = np.ascontiguousarray(qmx, np.float32).reshape(-1, 3)
        n = len(qmn)
        counts = np.zeros(n, np.int32)
        ids = np.full((n, cap), -1, np.int64)
        visited = np.zeros(n, np.int64)
        secs = lib().orc_aabbtree_query_batch(self.h, _p(qmn), _p(qmx), n, cap, threads, _p(counts), _p(ids),
                                              _p(visited))
        return counts, ids, visited, secs
