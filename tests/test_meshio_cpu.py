"""Mesh files on the CPU: the OBJ reader (mfcl_load_obj) against the
reference's test meshes and its loader's rules (test/test_fcl_utility.h:
250-316), with no GPU needed."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _py_obj(path):
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
    return np.array(v, np.float32), np.array(f, np.uint32)


@pytest.mark.parametrize("name,n_tris", [("env.obj", 2180), ("rob.obj", 216)])
def test_reference_meshes(name, n_tris):
    import mfcl

    v, t = mfcl.load_obj(os.path.join(GOLDEN, name))
    pv, pt = _py_obj(os.path.join(GOLDEN, name))
    assert len(t) == n_tris
    assert np.array_equal(v, pv)
    assert np.array_equal(t, pt)


def test_loader_rules(tmp_path):
    import mfcl

    p = tmp_path / "quad.obj"
    p.write_text("# comment\n"
                 "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 0 0 1\n"
                 "f 1 2 3 4\n"          # before any vn/vt line: 2 copies of (1,2,3)
                 "vn 0 0 1\n"
                 "f 1/1/1 2/2/1 3/3/1 4/4/1\n"  # after: the fan (1,2,3), (1,3,4)
                 "f 5\t2 3\r\n")
    v, t = mfcl.load_obj(str(p))
    assert v.shape == (5, 3) and v[4, 2] == 1.0
    assert t.tolist() == [[0, 1, 2], [0, 1, 2], [0, 1, 2], [0, 2, 3], [4, 1, 2]]


@pytest.mark.parametrize("body", ["v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n",
                                  "v 0 0\n", "f 1 2\n"])
def test_loader_errors(tmp_path, body):
    import mfcl

    p = tmp_path / "bad.obj"
    p.write_text(body)
    with pytest.raises(RuntimeError):
        mfcl.load_obj(str(p))
    with pytest.raises(RuntimeError):
        mfcl.load_obj(str(tmp_path / "missing.obj"))


def test_mesh_set_file_host(tmp_path):
    """A mesh-set file built on the host reads back with the right counts;
    every damaged variant (flipped bytes anywhere, truncation, another
    magic) is rejected."""
    import mfcl
    from mfcl import workloads as W

    pool = W.mesh_pool(4, 5, 200, 800)
    path = str(tmp_path / "set.mfclbvh")
    mfcl.build_mesh_file(path, pool["verts"], pool["vert_off"], pool["vert_cnt"], pool["tris"], pool["tri_off"],
                         pool["tri_cnt"])
    nt = int(pool["tri_cnt"].sum())
    assert mfcl.mesh_file_info(path) == (4, nt, 2 * nt - 4)
    raw = open(path, "rb").read()
    bad = tmp_path / "bad.mfclbvh"
    for pos in list(range(0, 160, 3)) + list(range(160, len(raw), max(1, len(raw) // 97))):
        b = bytearray(raw)
        b[pos] ^= 0x10
        bad.write_bytes(bytes(b))
        with pytest.raises(RuntimeError):
            mfcl.mesh_file_info(str(bad))
    for cut in (0, 7, 40, 150, len(raw) // 2, len(raw) - 1):
        bad.write_bytes(raw[:cut])
        with pytest.raises(RuntimeError):
            mfcl.mesh_file_info(str(bad))
    bad.write_bytes(raw[len(raw) // 2:])
    with pytest.raises(RuntimeError, match="not a mesh-set file"):
        mfcl.mesh_file_info(str(bad))
    b = bytearray(raw)
    b[len(b) // 2] ^= 1
    bad.write_bytes(bytes(b))
    with pytest.raises(RuntimeError, match="checksum"):
        mfcl.mesh_file_info(str(bad))
