"""Python binding of libmfcl.so (the C ABI in include/mfcl.h) via ctypes.

This is the host-side mirror used by the tests and the benchmark; the
product path is the HIP library itself.  Loading fails loudly when the
shared library has not been built -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libmfcl.so")

MFCL_SHAPE_BOX = 0
MFCL_SHAPE_SPHERE = 1
MFCL_SHAPE_CAPSULE = 3
MFCL_SHAPE_CYLINDER = 5
MFCL_SHAPE_CONVEX = 100
MODE_COLLIDE = 0
MODE_SIGNED_DISTANCE = 1
WS_CONVEX = 1
WS_MLP = 2
MAX_PAIR_CONTACTS = 4


class Request(C.Structure):
    _fields_ = [
        ("mode", C.c_int32),
        ("max_contacts", C.c_int32),
        ("penetration", C.c_int32),
        ("gjk_max_iterations", C.c_int32),
        ("gjk_tolerance", C.c_double),
        ("epa_max_faces", C.c_int32),
        ("epa_max_iterations", C.c_int32),
        ("epa_tolerance", C.c_double),
        ("use_guess", C.c_int32),
        ("reserved", C.c_int32),
    ]


ARM_MAX_LINKS = 16


class Arm(C.Structure):
    """mfcl_arm (include/mfcl.h)."""
    _fields_ = [
        ("n_links", C.c_int32),
        ("link_shape", C.c_uint32 * ARM_MAX_LINKS),
        ("link_len", C.c_float * ARM_MAX_LINKS),
        ("link_radius", C.c_float * ARM_MAX_LINKS),
        ("box_shape0", C.c_uint32),
        ("n_boxes", C.c_uint32),
        ("broadphase_cap", C.c_int32),
        ("gjk_max_iterations", C.c_int32),
        ("max_pairs", C.c_uint64),
        ("use_heightmap", C.c_int32),
        ("hm_box_shape0", C.c_uint32),
        ("hm_cap", C.c_int32),
        ("hm_pose7", C.c_float * 7),
    ]


def make_arm(link_shape, link_len, link_radius, box_shape0, n_boxes, broadphase_cap=32, max_pairs=0,
             heightmap=None) -> Arm:
    """heightmap: None, or dict(box_shape0=first pixel box shape id, pose7=the
    map frame's world pose, cap=candidate pixels kept per link)."""
    a = Arm()
    a.n_links = len(link_shape)
    for i in range(a.n_links):
        a.link_shape[i] = int(link_shape[i])
        a.link_len[i] = float(link_len[i])
        a.link_radius[i] = float(link_radius[i])
    a.box_shape0, a.n_boxes = int(box_shape0), int(n_boxes)
    a.broadphase_cap, a.gjk_max_iterations, a.max_pairs = int(broadphase_cap), 0, int(max_pairs)
    if heightmap is not None:
        a.use_heightmap = 1
        a.hm_box_shape0 = int(heightmap["box_shape0"])
        a.hm_cap = int(heightmap.get("cap", 4096))
        for t in range(7):
            a.hm_pose7[t] = float(heightmap["pose7"][t])
    return a


CONVEX_RESULT_DTYPE = np.dtype(
    [
        ("hit", np.int32),
        ("n_contacts", np.int32),
        ("method", np.int8),
        ("gjk_status", np.int8),
        ("epa_status", np.int8),
        ("ok", np.int8),
        ("gjk_iters", np.uint16),
        ("epa_iters", np.uint16),
        ("distance", np.float64),
        ("depth", np.float64),
        ("normal", np.float64, 3),
        ("pos", np.float64, 3),
        ("pa", np.float64, 3),
        ("pb", np.float64, 3),
    ]
)
assert CONVEX_RESULT_DTYPE.itemsize == 128

CONTACT_DTYPE = np.dtype([("normal", np.float64, 3), ("pos", np.float64, 3), ("depth", np.float64)])
assert CONTACT_DTYPE.itemsize == 56

MESH_CONTACT_DTYPE = np.dtype([("b1", np.int32), ("b2", np.int32), ("pos", np.float64, 3), ("normal", np.float64, 3),
                               ("depth", np.float64)])
assert MESH_CONTACT_DTYPE.itemsize == 64

MESH_RESULT_DTYPE = np.dtype(
    [("hit", np.int32), ("n_contacts", np.int32), ("bv_tests", np.int32), ("leaf_tests", np.int32)]
)

_lib = None


def load_library() -> C.CDLL:
    """Loads libmfcl.so; raises if it is missing (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"mfcl: {LIB_PATH} not built; run `make -C mind-fcl_amd` (or __graft_entry__.build())"
        )
    lib = C.CDLL(LIB_PATH)
    p, u64, i32, u32 = C.c_void_p, C.c_uint64, C.c_int32, C.c_uint32
    lib.mfcl_create.argtypes = [C.POINTER(p), C.c_int]
    lib.mfcl_destroy.argtypes = [p]
    lib.mfcl_last_error.argtypes = [p]
    lib.mfcl_last_error.restype = C.c_char_p
    lib.mfcl_sync.argtypes = [p]
    lib.mfcl_set_timing.argtypes = [p, C.c_int]
    lib.mfcl_last_timing.argtypes = [p, p]
    lib.mfcl_last_counters.argtypes = [p, p]
    lib.mfcl_last_counters_n.argtypes = [p, p, i32]
    lib.mfcl_request_defaults.argtypes = [C.POINTER(Request), i32]
    lib.mfcl_request_defaults.restype = None
    lib.mfcl_dev_alloc.argtypes = [p, u64, C.POINTER(p)]
    lib.mfcl_dev_free.argtypes = [p, p]
    lib.mfcl_h2d.argtypes = [p, p, p, u64]
    lib.mfcl_d2h.argtypes = [p, p, p, u64]
    lib.mfcl_upload_shapes.argtypes = [p, p, p, p, p, u32, p, u64]
    lib.mfcl_upload_convex_faces.argtypes = [p, p, u64, p, p]
    lib.mfcl_arm_query_dev.argtypes = [p, C.POINTER(Arm), p, u64, C.c_float, p, p, p, C.POINTER(u64), p]
    lib.mfcl_upload_heightmap.argtypes = [p, p, C.c_uint16, C.c_uint16, C.c_float, C.c_float]
    lib.mfcl_arm_last_heightmap_pairs.argtypes = [p, C.POINTER(u64)]
    lib.mfcl_convex_query.argtypes = [p, p, p, p, p, p, u64, C.POINTER(Request), p]
    lib.mfcl_convex_query_dev.argtypes = [p, p, p, p, p, p, u64, C.POINTER(Request), p, p]
    lib.mfcl_convex_collide_contacts.argtypes = [p, p, p, p, p, p, u64, C.POINTER(Request), p, p]
    lib.mfcl_convex_collide_contacts_dev.argtypes = [p, p, p, p, p, p, u64, C.POINTER(Request), p, p, p]
    lib.mfcl_convex_query_static_dev.argtypes = [p, p, p, p, p, p, u64, C.POINTER(Request), p, p]
    lib.mfcl_convex_query_static_dev_max_pairs.argtypes = [C.POINTER(u64)]
    lib.mfcl_graph_begin.argtypes = [p]
    lib.mfcl_graph_end.argtypes = [p, C.POINTER(p)]
    lib.mfcl_graph_launch.argtypes = [p, p]
    lib.mfcl_graph_destroy.argtypes = [p]
    lib.mfcl_reserve.argtypes = [p, i32, u64]
    lib.mfcl_reserved_pairs.argtypes = [p, C.POINTER(u64)]
    lib.mfcl_reserve_arm.argtypes = [p, C.POINTER(Arm), u64]
    lib.mfcl_convex_hulls.argtypes = [p, p, p, p, u32, p, p, p, p, p]
    lib.mfcl_convex_hulls_dev.argtypes = [p, p, p, p, u32, i32, p, p, p, p, p, p]
    lib.mfcl_upload_convex_points.argtypes = [p, p, p, p, u32, p]
    lib.mfcl_convex_directed_penetration.argtypes = [p, p, p, p, p, p, u64, p, p]
    lib.mfcl_convex_directed_penetration_dev.argtypes = [p, p, p, p, p, p, u64, p, p, p]
    lib.mfcl_convex_incremental_penetration.argtypes = [p, p, p, p, p, p, u64, p, p]
    lib.mfcl_convex_incremental_penetration_dev.argtypes = [p, p, p, p, p, p, u64, p, p, p]
    lib.mfcl_upload_meshes.argtypes = [p, p, p, p, p, p, p, u32]
    lib.mfcl_mesh_num_nodes.argtypes = [p, u32, C.POINTER(u32)]
    lib.mfcl_mesh_get_nodes.argtypes = [p, u32, p]
    lib.mfcl_mesh_shape_collide.argtypes = [p, p, p, p, p, u64, p, p, p]
    lib.mfcl_mesh_shape_collide_contacts.argtypes = [p, p, p, p, p, u64, p, p, p]
    lib.mfcl_mesh_replace_vertices.argtypes = [p, u32, p, i32]
    lib.mfcl_mesh_replace_vertices_dev.argtypes = [p, u32, p, p]
    lib.mfcl_save_meshes.argtypes = [p, C.c_char_p]
    lib.mfcl_load_meshes.argtypes = [p, C.c_char_p]
    lib.mfcl_load_obj.argtypes = [C.c_char_p, p, C.POINTER(u64), p, C.POINTER(u64)]
    lib.mfcl_build_mesh_file.argtypes = [C.c_char_p, p, p, p, p, p, p, u32]
    lib.mfcl_mesh_file_info.argtypes = [C.c_char_p, C.POINTER(u32), C.POINTER(u64), C.POINTER(u64)]
    lib.mfcl_host_last_error.argtypes = []
    lib.mfcl_host_last_error.restype = C.c_char_p
    lib.mfcl_mesh_collide.argtypes = [p, p, p, p, p, u64, i32, p, p]
    lib.mfcl_mesh_collide_dev.argtypes = [p, p, p, p, p, u64, i32, p, p, p]
    lib.mfcl_mesh_collide_contacts.argtypes = [p, p, p, p, p, u64, i32, p, p]
    lib.mfcl_mesh_collide_contacts_dev.argtypes = [p, p, p, p, p, u64, i32, p, p, p]
    lib.mfcl_aabbtree_build.argtypes = [p, p, p, p, u32]
    lib.mfcl_aabbtree_query.argtypes = [p, p, p, u64, i32, p, p]
    lib.mfcl_aabbtree_query_dev.argtypes = [p, p, p, u64, i32, p, p, p]
    lib.mfcl_tune.argtypes = [p, i32, i32]
    lib.mfcl_mlp_num_params.argtypes = []
    lib.mfcl_mlp_init.argtypes = [p, u64]
    lib.mfcl_mlp_set_params.argtypes = [p, p, i32]
    lib.mfcl_mlp_get_params.argtypes = [p, p, p, i32]
    lib.mfcl_mlp_features_dev.argtypes = [p, p, p, p, p, u64, p, p]
    lib.mfcl_mlp_guess_dev.argtypes = [p, p, p, p, p, u64, p, p, p]
    lib.mfcl_mlp_forward_dev.argtypes = [p, p, u64, p, p]
    lib.mfcl_mlp_train_step_dev.argtypes = [p, p, p, u64, p, i32, p, p]
    _lib = lib
    return lib


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def faces_from_triangles(tris, tri_off, tri_cnt):
    """Triangles (k x 3 vertex indices; shape i owns tris[tri_off[i] ..
    tri_off[i] + tri_cnt[i])) -> (faces, face_off, num_faces) in the
    reference's face encoding [3, a, b, c] per triangle."""
    tris = np.ascontiguousarray(tris, np.int32).reshape(-1, 3)
    faces = np.empty((len(tris), 4), np.int32)
    faces[:, 0] = 3
    faces[:, 1:] = tris
    return faces.reshape(-1), 4 * np.asarray(tri_off, np.uint64), np.asarray(tri_cnt, np.uint32)


OBB_NODE_DTYPE = np.dtype([("ax", "<f4", 9), ("c", "<f4", 3), ("e", "<f4", 3), ("child", "<i4")])


def _host_check(lib, rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {lib.mfcl_host_last_error().decode()}")


def load_obj(path: str):
    """Vertices (n x 3 float32) and triangles (m x 3 uint32, 0-based) of a
    Wavefront OBJ file, read as the reference's test loader reads it."""
    lib = load_library()
    nv, nt = C.c_uint64(0), C.c_uint64(0)
    _host_check(lib, lib.mfcl_load_obj(path.encode(), None, C.byref(nv), None, C.byref(nt)), "mfcl_load_obj")
    v = np.zeros((nv.value, 3), np.float32)
    t = np.zeros((nt.value, 3), np.uint32)
    _host_check(lib, lib.mfcl_load_obj(path.encode(), _ptr(v), C.byref(nv), _ptr(t), C.byref(nt)), "mfcl_load_obj")
    return v, t


def build_mesh_file(path: str, verts, vert_off, vert_cnt, tris, tri_off, tri_cnt):
    """OBB trees of a mesh set built on the host and written to a mesh-set
    file (loadable with Context.load_meshes); no GPU needed."""
    lib = load_library()
    verts = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
    tris = np.ascontiguousarray(tris, np.uint32).reshape(-1, 3)
    vo, vc = np.ascontiguousarray(vert_off, np.uint64), np.ascontiguousarray(vert_cnt, np.uint32)
    to, tc = np.ascontiguousarray(tri_off, np.uint64), np.ascontiguousarray(tri_cnt, np.uint32)
    _host_check(lib, lib.mfcl_build_mesh_file(path.encode(), _ptr(verts), _ptr(vo), _ptr(vc), _ptr(tris), _ptr(to),
                                              _ptr(tc), len(vc)), "mfcl_build_mesh_file")


def mesh_file_info(path: str):
    """(meshes, triangles, nodes) of a mesh-set file after a full check."""
    lib = load_library()
    n, nt, nn = C.c_uint32(0), C.c_uint64(0), C.c_uint64(0)
    _host_check(lib, lib.mfcl_mesh_file_info(path.encode(), C.byref(n), C.byref(nt), C.byref(nn)),
                "mfcl_mesh_file_info")
    return n.value, nt.value, nn.value


def request(mode: int = MODE_COLLIDE, **kw) -> Request:
    r = Request()
    load_library().mfcl_request_defaults(C.byref(r), mode)
    for k, v in kw.items():
        setattr(r, k, v)
    return r


class DeviceArray:
    """A device allocation owned by a Context (freed with it or explicitly)."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        ctx._check(ctx.lib.mfcl_dev_alloc(ctx.h, max(1, self.nbytes), C.byref(self.ptr)))

    @property
    def value(self):
        return self.ptr.value

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        self.ctx._check(self.ctx.lib.mfcl_h2d(self.ctx.h, self.ptr, _ptr(a), a.nbytes))

    def download(self, a: np.ndarray):
        assert a.flags.c_contiguous and a.nbytes <= self.nbytes
        self.ctx._check(self.ctx.lib.mfcl_d2h(self.ctx.h, _ptr(a), self.ptr, a.nbytes))
        return a

    def free(self):
        if self.ptr.value:
            self.ctx.lib.mfcl_dev_free(self.ctx.h, self.ptr)
            self.ptr = C.c_void_p()


class Graph:
    """A recorded query sequence (mfcl_graph_*)."""

    def __init__(self, ctx: "Context", handle: C.c_void_p):
        self.ctx = ctx
        self.h = handle

    def launch(self):
        self.ctx._check(self.ctx.lib.mfcl_graph_launch(self.h, None))

    def destroy(self):
        if self.h:
            self.ctx.lib.mfcl_graph_destroy(self.h)
            self.h = C.c_void_p()


class Context:
    def __init__(self, device: int = 0):
        self.lib = load_library()
        self._reserved = {WS_CONVEX: 0, WS_MLP: 0}
        self.h = C.c_void_p()
        rc = self.lib.mfcl_create(C.byref(self.h), device)
        if rc != 0:
            raise RuntimeError(f"mfcl_create failed ({rc}); is a GPU visible?")

    def close(self):
        if self.h:
            self.lib.mfcl_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc: int):
        if rc != 0:
            raise RuntimeError(f"mfcl error {rc}: {self.lib.mfcl_last_error(self.h).decode()}")

    def sync(self):
        self._check(self.lib.mfcl_sync(self.h))

    def set_timing(self, on: bool):
        self._check(self.lib.mfcl_set_timing(self.h, 1 if on else 0))

    def last_timing(self) -> np.ndarray:
        ms = np.zeros(6, np.float32)
        self._check(self.lib.mfcl_last_timing(self.h, _ptr(ms)))
        return ms

    def last_counters(self) -> np.ndarray:
        """[pairs, hull vertices read (GJK + EPA stages), EPA pairs, EPA
        reruns at full capacity, hull vertices read by the EPA stage, EPA
        pairs continued in a device-memory polytope of full capacity]"""
        c = np.zeros(6, np.uint64)
        self._check(self.lib.mfcl_last_counters_n(self.h, _ptr(c), 6))
        return c

    def graph_begin(self):
        """Starts recording the following "_dev" queries (mfcl_graph_begin)."""
        self._check(self.lib.mfcl_graph_begin(self.h))

    def graph_end(self) -> "Graph":
        g = C.c_void_p()
        self._check(self.lib.mfcl_graph_end(self.h, C.byref(g)))
        return Graph(self, g)

    def reserve(self, kind: int, count: int):
        """mfcl_reserve: size the workspace of the "_dev" queries of that kind
        for `count` items (the "_dev" wrappers below call it when a query
        is larger than what was reserved)."""
        if count > self._reserved[kind]:
            self._check(self.lib.mfcl_reserve(self.h, kind, int(count)))
            self._reserved[kind] = int(count)

    def reserved_pairs(self) -> int:
        n = C.c_uint64(0)
        self._check(self.lib.mfcl_reserved_pairs(self.h, C.byref(n)))
        return int(n.value)

    def reserve_arm(self, arm: "Arm", n_configs: int):
        self._check(self.lib.mfcl_reserve_arm(self.h, C.byref(arm), int(n_configs)))

    def alloc(self, nbytes: int) -> DeviceArray:
        return DeviceArray(self, nbytes)

    def to_device(self, a: np.ndarray) -> DeviceArray:
        a = np.ascontiguousarray(a)
        d = DeviceArray(self, a.nbytes)
        d.upload(a)
        return d

    # ---- shapes
    def upload_shapes(self, types, params4, vert_off=None, vert_cnt=None, verts=None, faces=None):
        """faces: optional (faces, face_off, num_faces) in the reference's
        encoding (see faces_from_triangles); hulls then use the vertex walk
        where the reference does."""
        types = np.ascontiguousarray(types, np.int32)
        params4 = np.ascontiguousarray(params4, np.float32).reshape(-1, 4)
        n = len(types)
        if verts is None:
            verts = np.zeros((1, 3), np.float32)
            vert_off = np.zeros(n, np.uint64)
            vert_cnt = np.zeros(n, np.uint32)
        verts = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
        vert_off = np.ascontiguousarray(vert_off, np.uint64)
        vert_cnt = np.ascontiguousarray(vert_cnt, np.uint32)
        self._check(
            self.lib.mfcl_upload_shapes(
                self.h, _ptr(types), _ptr(params4), _ptr(vert_off), _ptr(vert_cnt), n, _ptr(verts), len(verts)
            )
        )
        if faces is not None:
            self.upload_convex_faces(*faces)

    def upload_convex_faces(self, faces, face_off, num_faces):
        faces = np.ascontiguousarray(faces, np.int32)
        face_off = np.ascontiguousarray(face_off, np.uint64)
        num_faces = np.ascontiguousarray(num_faces, np.uint32)
        self._check(self.lib.mfcl_upload_convex_faces(self.h, _ptr(faces), len(faces), _ptr(face_off),
                                                      _ptr(num_faces)))

    def convex_query(self, pair_a, pair_b, pose_a, pose_b, req: Request, guess=None) -> np.ndarray:
        pair_a = np.ascontiguousarray(pair_a, np.uint32)
        pair_b = np.ascontiguousarray(pair_b, np.uint32)
        pose_a = np.ascontiguousarray(pose_a, np.float32).reshape(-1, 7)
        pose_b = np.ascontiguousarray(pose_b, np.float32).reshape(-1, 7)
        g = None if guess is None else np.ascontiguousarray(guess, np.float32).reshape(-1, 3)
        n = len(pair_a)
        out = np.zeros(n, CONVEX_RESULT_DTYPE)
        self._check(
            self.lib.mfcl_convex_query(
                self.h, _ptr(pair_a), _ptr(pair_b), _ptr(pose_a), _ptr(pose_b), _ptr(g), n, C.byref(req), _ptr(out)
            )
        )
        return out

    def convex_incremental_penetration(self, pair_a, pair_b, pose_a, pose_b, hints, req: Request) -> np.ndarray:
        """collisionPenetration in the incremental minimum-penetration MPR
        mode, refined from hints[i] (world unit vectors)."""
        return self.convex_directed_penetration(pair_a, pair_b, pose_a, pose_b, hints, req, incremental=True)

    def convex_directed_penetration(self, pair_a, pair_b, pose_a, pose_b, dirs, req: Request,
                                    incremental: bool = False) -> np.ndarray:
        """collisionPenetration in the directed MPR mode: per pair the
        boolean collide and, when colliding, the penetration depth along
        dirs[i] (world unit vector, shape b's escape direction)."""
        fn = self.lib.mfcl_convex_incremental_penetration if incremental else self.lib.mfcl_convex_directed_penetration
        pair_a = np.ascontiguousarray(pair_a, np.uint32)
        pair_b = np.ascontiguousarray(pair_b, np.uint32)
        pose_a = np.ascontiguousarray(pose_a, np.float32).reshape(-1, 7)
        pose_b = np.ascontiguousarray(pose_b, np.float32).reshape(-1, 7)
        dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        n = len(pair_a)
        out = np.zeros(n, CONVEX_RESULT_DTYPE)
        self._check(fn(self.h, _ptr(pair_a), _ptr(pair_b), _ptr(pose_a), _ptr(pose_b), _ptr(dirs), n, C.byref(req),
                       _ptr(out)))
        return out

    def convex_collide_contacts(self, pair_a, pair_b, pose_a, pose_b, req: Request):
        """Every kept contact per pair (collide mode with penetration):
        returns (results, contacts[n, min(max_contacts, 4)] of CONTACT_DTYPE;
        records past results['n_contacts'] are zero)."""
        pair_a = np.ascontiguousarray(pair_a, np.uint32)
        pair_b = np.ascontiguousarray(pair_b, np.uint32)
        pose_a = np.ascontiguousarray(pose_a, np.float32).reshape(-1, 7)
        pose_b = np.ascontiguousarray(pose_b, np.float32).reshape(-1, 7)
        n = len(pair_a)
        out = np.zeros(n, CONVEX_RESULT_DTYPE)
        cont = np.zeros((n, max(1, min(req.max_contacts, MAX_PAIR_CONTACTS))), CONTACT_DTYPE)
        self._check(
            self.lib.mfcl_convex_collide_contacts(
                self.h, _ptr(pair_a), _ptr(pair_b), _ptr(pose_a), _ptr(pose_b), None, n, C.byref(req), _ptr(out),
                _ptr(cont)
            )
        )
        return out, cont

    def convex_collide_contacts_dev(self, pa, pb, qa, qb, n: int, req: Request, out: DeviceArray,
                                    contacts: DeviceArray, guess: Optional[DeviceArray] = None):
        self.reserve(WS_CONVEX, n)
        self._check(
            self.lib.mfcl_convex_collide_contacts_dev(
                self.h, pa.ptr, pb.ptr, qa.ptr, qb.ptr, None if guess is None else guess.ptr, n, C.byref(req),
                out.ptr, contacts.ptr, None
            )
        )

    def convex_query_dev(self, pa: DeviceArray, pb: DeviceArray, qa: DeviceArray, qb: DeviceArray, n: int,
                         req: Request, out: DeviceArray, guess: Optional[DeviceArray] = None):
        self.reserve(WS_CONVEX, n)
        self._check(
            self.lib.mfcl_convex_query_dev(
                self.h, pa.ptr, pb.ptr, qa.ptr, qb.ptr, None if guess is None else guess.ptr, n, C.byref(req),
                out.ptr, None
            )
        )

    # ---- convex hulls
    def convex_hulls(self, points, pt_off, pt_cnt):
        """Device convex hulls of point sets: list of (status, hull vertex
        indices (ascending, within the set), triangles over the hull's
        vertex numbering)."""
        points = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        pt_off = np.ascontiguousarray(pt_off, np.uint64)
        pt_cnt = np.ascontiguousarray(pt_cnt, np.uint32)
        n = len(pt_cnt)
        total = int(max((pt_off + pt_cnt).max(), 1)) if n else 1
        hc, tc, st = np.zeros(n, np.uint32), np.zeros(n, np.uint32), np.zeros(n, np.int32)
        hi = np.zeros(total, np.uint32)
        tr = np.zeros(6 * total, np.int32)
        self._check(self.lib.mfcl_convex_hulls(self.h, _ptr(points), _ptr(pt_off), _ptr(pt_cnt), n, _ptr(hc),
                                               _ptr(hi), _ptr(tc), _ptr(tr), _ptr(st)))
        out = []
        for i in range(n):
            o = int(pt_off[i])
            out.append((int(st[i]), hi[o:o + hc[i]].copy(), tr[6 * o:6 * o + 3 * tc[i]].reshape(-1, 3).copy()))
        return out

    def upload_convex_points(self, points, pt_off, pt_cnt) -> np.ndarray:
        """Shape table of the convex hulls of point sets, built on the device
        (vertices in ascending point order, hull triangles as faces).
        Returns the hull vertex count of every shape."""
        points = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        pt_off = np.ascontiguousarray(pt_off, np.uint64)
        pt_cnt = np.ascontiguousarray(pt_cnt, np.uint32)
        hc = np.zeros(len(pt_cnt), np.uint32)
        self._check(self.lib.mfcl_upload_convex_points(self.h, _ptr(points), _ptr(pt_off), _ptr(pt_cnt),
                                                       len(pt_cnt), _ptr(hc)))
        return hc

    # ---- meshes
    def mesh_get_nodes(self, mesh: int) -> np.ndarray:
        n = C.c_uint32(0)
        self._check(self.lib.mfcl_mesh_num_nodes(self.h, mesh, C.byref(n)))
        out = np.zeros(n.value, OBB_NODE_DTYPE)
        self._check(self.lib.mfcl_mesh_get_nodes(self.h, mesh, _ptr(out)))
        return out

    def mesh_shape_collide(self, mesh_ids, shape_ids, pose_mesh, pose_shape, req: Request):
        """Mesh-shape collision (penetration disabled): (results of
        MESH_RESULT_DTYPE, colliding triangle ids n x max_contacts, -1 past
        n_contacts)."""
        mesh_ids = np.ascontiguousarray(mesh_ids, np.uint32)
        shape_ids = np.ascontiguousarray(shape_ids, np.uint32)
        pose_mesh = np.ascontiguousarray(pose_mesh, np.float32).reshape(-1, 7)
        pose_shape = np.ascontiguousarray(pose_shape, np.float32).reshape(-1, 7)
        n = len(mesh_ids)
        out = np.zeros(n, MESH_RESULT_DTYPE)
        tri = np.full((n, max(1, req.max_contacts)), -1, np.int32)
        self._check(self.lib.mfcl_mesh_shape_collide(self.h, _ptr(mesh_ids), _ptr(shape_ids), _ptr(pose_mesh),
                                                     _ptr(pose_shape), n, C.byref(req), _ptr(out), _ptr(tri)))
        return out, tri

    def mesh_shape_collide_contacts(self, mesh_ids, shape_ids, pose_mesh, pose_shape, req: Request):
        """Mesh-shape collision with penetration enabled (req.penetration
        = 1): (results of MESH_RESULT_DTYPE, contacts n x max_contacts of
        MESH_CONTACT_DTYPE, ascending b1, b1 = -1 past n_contacts)."""
        mesh_ids = np.ascontiguousarray(mesh_ids, np.uint32)
        shape_ids = np.ascontiguousarray(shape_ids, np.uint32)
        pose_mesh = np.ascontiguousarray(pose_mesh, np.float32).reshape(-1, 7)
        pose_shape = np.ascontiguousarray(pose_shape, np.float32).reshape(-1, 7)
        n = len(mesh_ids)
        out = np.zeros(n, MESH_RESULT_DTYPE)
        cont = np.zeros((n, max(1, req.max_contacts)), MESH_CONTACT_DTYPE)
        self._check(self.lib.mfcl_mesh_shape_collide_contacts(self.h, _ptr(mesh_ids), _ptr(shape_ids),
                                                              _ptr(pose_mesh), _ptr(pose_shape), n, C.byref(req),
                                                              _ptr(out), _ptr(cont)))
        return out, cont

    def mesh_replace_vertices(self, mesh: int, verts, refit: bool = True):
        verts = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
        self._check(self.lib.mfcl_mesh_replace_vertices(self.h, mesh, _ptr(verts), 1 if refit else 0))

    def mesh_replace_vertices_dev(self, mesh: int, verts: DeviceArray, stream=None):
        self._check(self.lib.mfcl_mesh_replace_vertices_dev(self.h, mesh, verts.ptr, stream))

    def save_meshes(self, path: str):
        self._check(self.lib.mfcl_save_meshes(self.h, path.encode()))

    def load_meshes(self, path: str):
        self._check(self.lib.mfcl_load_meshes(self.h, path.encode()))

    def upload_meshes(self, verts, vert_off, vert_cnt, tris, tri_off, tri_cnt):
        verts = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
        tris = np.ascontiguousarray(tris, np.uint32).reshape(-1, 3)
        vert_off = np.ascontiguousarray(vert_off, np.uint64)
        vert_cnt = np.ascontiguousarray(vert_cnt, np.uint32)
        tri_off = np.ascontiguousarray(tri_off, np.uint64)
        tri_cnt = np.ascontiguousarray(tri_cnt, np.uint32)
        self._check(
            self.lib.mfcl_upload_meshes(
                self.h, _ptr(verts), _ptr(vert_off), _ptr(vert_cnt), _ptr(tris), _ptr(tri_off), _ptr(tri_cnt),
                len(vert_cnt)
            )
        )

    def mesh_collide(self, pair_a, pair_b, pose_a, pose_b, max_contacts: int = 16):
        pair_a = np.ascontiguousarray(pair_a, np.uint32)
        pair_b = np.ascontiguousarray(pair_b, np.uint32)
        pose_a = np.ascontiguousarray(pose_a, np.float32).reshape(-1, 7)
        pose_b = np.ascontiguousarray(pose_b, np.float32).reshape(-1, 7)
        n = len(pair_a)
        out = np.zeros(n, MESH_RESULT_DTYPE)
        contacts = np.full((n, max_contacts, 2), -1, np.int32)
        self._check(
            self.lib.mfcl_mesh_collide(
                self.h, _ptr(pair_a), _ptr(pair_b), _ptr(pose_a), _ptr(pose_b), n, max_contacts, _ptr(out),
                _ptr(contacts)
            )
        )
        return out, contacts

    def mesh_collide_contacts(self, pair_a, pair_b, pose_a, pose_b, max_contacts: int = 16):
        """Penetration-enabled mesh collide: (results, contacts[n, max_contacts]
        of MESH_CONTACT_DTYPE; records past results['n_contacts'] are zero)."""
        pair_a = np.ascontiguousarray(pair_a, np.uint32)
        pair_b = np.ascontiguousarray(pair_b, np.uint32)
        pose_a = np.ascontiguousarray(pose_a, np.float32).reshape(-1, 7)
        pose_b = np.ascontiguousarray(pose_b, np.float32).reshape(-1, 7)
        n = len(pair_a)
        out = np.zeros(n, MESH_RESULT_DTYPE)
        cont = np.zeros((n, max_contacts), MESH_CONTACT_DTYPE)
        self._check(
            self.lib.mfcl_mesh_collide_contacts(
                self.h, _ptr(pair_a), _ptr(pair_b), _ptr(pose_a), _ptr(pose_b), n, max_contacts, _ptr(out),
                _ptr(cont)
            )
        )
        return out, cont

    def mesh_collide_dev(self, pa, pb, qa, qb, n, max_contacts, out, contacts):
        self._check(
            self.lib.mfcl_mesh_collide_dev(self.h, pa.ptr, pb.ptr, qa.ptr, qb.ptr, n, max_contacts, out.ptr,
                                           contacts.ptr, None)
        )

    # ---- broadphase
    def aabbtree_build(self, mn, mx, ids):
        mn = np.ascontiguousarray(mn, np.float32).reshape(-1, 3)
        mx = np.ascontiguousarray(mx, np.float32).reshape(-1, 3)
        ids = np.ascontiguousarray(ids, np.int64)
        self._check(self.lib.mfcl_aabbtree_build(self.h, _ptr(mn), _ptr(mx), _ptr(ids), len(ids)))

    def aabbtree_query(self, qmn, qmx, cap: int):
        qmn = np.ascontiguousarray(qmn, np.float32).reshape(-1, 3)
        qmx = np.ascontiguousarray(qmx, np.float32).reshape(-1, 3)
        n = len(qmn)
        counts = np.zeros(n, np.int32)
        ids = np.full((n, max(cap, 1)), -1, np.int64)
        self._check(self.lib.mfcl_aabbtree_query(self.h, _ptr(qmn), _ptr(qmx), n, cap, _ptr(counts), _ptr(ids)))
        return counts, ids


    # ---- robot-arm configurations
    def arm_query_dev(self, arm: Arm, q: DeviceArray, n_configs: int, margin: float, box_pose7: DeviceArray,
                      any_hit: DeviceArray, min_dist: DeviceArray) -> int:
        n_pairs = C.c_uint64(0)
        self._check(self.lib.mfcl_arm_query_dev(self.h, C.byref(arm), q.ptr, n_configs, margin, box_pose7.ptr,
                                                any_hit.ptr, min_dist.ptr, C.byref(n_pairs), None))
        return int(n_pairs.value)

    def upload_heightmap(self, heights_mm, res_x: float, res_y: float):
        """heights_mm: (2 half_y, 2 half_x) uint16, row y holds pixels (x, y)."""
        h = np.ascontiguousarray(heights_mm, np.uint16)
        fy, fx = h.shape
        self._check(self.lib.mfcl_upload_heightmap(self.h, h.ctypes.data, fx // 2, fy // 2, res_x, res_y))

    def arm_last_heightmap_pairs(self) -> int:
        n = C.c_uint64(0)
        self._check(self.lib.mfcl_arm_last_heightmap_pairs(self.h, C.byref(n)))
        return int(n.value)

    # ---- learned warm start
    def mlp_num_params(self) -> int:
        return int(self.lib.mfcl_mlp_num_params())

    def mlp_init(self, seed: int):
        self._check(self.lib.mfcl_mlp_init(self.h, int(seed)))

    def mlp_set_params(self, params: np.ndarray):
        params = np.ascontiguousarray(params, np.float32)
        self._check(self.lib.mfcl_mlp_set_params(self.h, _ptr(params), len(params)))

    def mlp_get_params(self, grads: bool = False):
        n = self.mlp_num_params()
        p = np.zeros(n, np.float32)
        g = np.zeros(n, np.float32) if grads else None
        self._check(self.lib.mfcl_mlp_get_params(self.h, _ptr(p), _ptr(g), n))
        return (p, g) if grads else p

    def mlp_features_dev(self, pa: DeviceArray, pb: DeviceArray, qa: DeviceArray, qb: DeviceArray, n: int,
                         feat: DeviceArray):
        self._check(self.lib.mfcl_mlp_features_dev(self.h, pa.ptr, pb.ptr, qa.ptr, qb.ptr, n, feat.ptr, None))

    def mlp_guess_dev(self, pa: DeviceArray, pb: DeviceArray, qa: DeviceArray, qb: DeviceArray, n: int,
                      guess: DeviceArray, dist: Optional[DeviceArray] = None):
        self._check(self.lib.mfcl_mlp_guess_dev(self.h, pa.ptr, pb.ptr, qa.ptr, qb.ptr, n, guess.ptr,
                                                None if dist is None else dist.ptr, None))

    def mlp_forward_dev(self, feat: DeviceArray, n: int, out4: DeviceArray):
        self.reserve(WS_MLP, n)
        self._check(self.lib.mfcl_mlp_forward_dev(self.h, feat.ptr, n, out4.ptr, None))

    def mlp_train_step_dev(self, feat: DeviceArray, target: DeviceArray, n: int, hyper=None,
                           apply_update: bool = True) -> float:
        self.reserve(WS_MLP, n)
        loss = np.zeros(1, np.float32)
        hp = None if hyper is None else np.ascontiguousarray(hyper, np.float32)
        self._check(self.lib.mfcl_mlp_train_step_dev(self.h, feat.ptr, target.ptr, n, _ptr(hp),
                                                     1 if apply_update else 0, _ptr(loss), None))
        return float(loss[0])


# Symbols the C header declares (checked by the CPU test suite).
def header_symbols(header_path: Optional[str] = None):
    import re

    if header_path is None:
        header_path = os.path.join(os.path.dirname(os.path.dirname(_HERE)), "include", "mfcl.h")
    text = open(header_path).read()
    return sorted(set(re.findall(r"\b(mfcl_[a-z0-9_]+)\s*\(", text)))
