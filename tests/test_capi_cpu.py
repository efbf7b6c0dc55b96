"""CPU: the C-ABI library loads, exports every entry point the header
declares, and its structs have the layout the Python binding assumes.

No compute entry point is called here (there is no GPU in this process);
only mfcl_request_defaults, a pure host function, is exercised.
"""
import ctypes as C
import os
import subprocess
import textwrap

import numpy as np
import pytest

import mfcl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mfcl.h")


@pytest.fixture(scope="module")
def lib():
    return mfcl.load_library()


def test_header_symbols_exported(lib):
    syms = mfcl.header_symbols(HEADER)
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_exports_are_extern_c():
    """Every exported mfcl_* symbol is unmangled and declared in the header."""
    out = subprocess.run(["nm", "-D", "--defined-only", mfcl.LIB_PATH], capture_output=True, text=True, check=True)
    exported = sorted({ln.split()[-1] for ln in out.stdout.splitlines() if ln.split()[-1].startswith("mfcl_")})
    assert set(exported) == set(mfcl.header_symbols(HEADER))


def _c_layout(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(textwrap.dedent("""
        #include <stddef.h>
        #include <stdio.h>
        #include "mfcl.h"
        #define OFF(T, f) printf(#T "." #f " %zu\\n", offsetof(T, f))
        int main(void) {
          printf("sizeof.request %zu\\n", sizeof(mfcl_request));
          printf("sizeof.convex_result %zu\\n", sizeof(mfcl_convex_result));
          printf("sizeof.mesh_result %zu\\n", sizeof(mfcl_mesh_result));
          printf("sizeof.arm %zu\\n", sizeof(mfcl_arm));
          OFF(mfcl_arm, n_links); OFF(mfcl_arm, link_shape); OFF(mfcl_arm, link_len); OFF(mfcl_arm, link_radius);
          OFF(mfcl_arm, box_shape0); OFF(mfcl_arm, n_boxes); OFF(mfcl_arm, broadphase_cap);
          OFF(mfcl_arm, gjk_max_iterations); OFF(mfcl_arm, max_pairs); OFF(mfcl_arm, use_heightmap);
          OFF(mfcl_arm, hm_box_shape0); OFF(mfcl_arm, hm_cap); OFF(mfcl_arm, hm_pose7);
          OFF(mfcl_request, mode); OFF(mfcl_request, max_contacts); OFF(mfcl_request, penetration);
          OFF(mfcl_request, gjk_max_iterations); OFF(mfcl_request, gjk_tolerance);
          OFF(mfcl_request, epa_max_faces); OFF(mfcl_request, epa_max_iterations);
          OFF(mfcl_request, epa_tolerance); OFF(mfcl_request, use_guess);
          OFF(mfcl_convex_result, hit); OFF(mfcl_convex_result, n_contacts); OFF(mfcl_convex_result, method);
          OFF(mfcl_convex_result, gjk_status); OFF(mfcl_convex_result, epa_status); OFF(mfcl_convex_result, ok);
          OFF(mfcl_convex_result, gjk_iters); OFF(mfcl_convex_result, epa_iters);
          OFF(mfcl_convex_result, distance); OFF(mfcl_convex_result, depth); OFF(mfcl_convex_result, normal);
          OFF(mfcl_convex_result, pos); OFF(mfcl_convex_result, pa); OFF(mfcl_convex_result, pb);
          OFF(mfcl_mesh_result, hit); OFF(mfcl_mesh_result, n_contacts); OFF(mfcl_mesh_result, bv_tests);
          OFF(mfcl_mesh_result, leaf_tests);
          return 0;
        }
    """))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)],
                   check=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    return dict((ln.split()[0], int(ln.split()[1])) for ln in res.splitlines())


def test_struct_layout_matches_binding(tmp_path):
    lay = _c_layout(tmp_path)
    assert lay["sizeof.request"] == C.sizeof(mfcl.Request)
    for name, _ in mfcl.Request._fields_:
        key = f"mfcl_request.{name}"
        if key in lay:
            assert lay[key] == getattr(mfcl.Request, name).offset, name
    assert lay["sizeof.convex_result"] == mfcl.CONVEX_RESULT_DTYPE.itemsize == 128
    for name in mfcl.CONVEX_RESULT_DTYPE.names:
        key = f"mfcl_convex_result.{name}"
        if key in lay:
            assert lay[key] == mfcl.CONVEX_RESULT_DTYPE.fields[name][1], name
    assert lay["sizeof.arm"] == C.sizeof(mfcl.Arm)
    for name, _ in mfcl.Arm._fields_:
        assert lay[f"mfcl_arm.{name}"] == getattr(mfcl.Arm, name).offset, name
    assert lay["sizeof.mesh_result"] == mfcl.MESH_RESULT_DTYPE.itemsize
    for name in mfcl.MESH_RESULT_DTYPE.names:
        key = f"mfcl_mesh_result.{name}"
        if key in lay:
            assert lay[key] == mfcl.MESH_RESULT_DTYPE.fields[name][1], name


def test_request_defaults_match_reference_solver(lib):
    """CollisionRequest() defaults (collision_request-inl.h:48-60): one
    contact, penetration disabled, tolerances 1e-6; fcl::collide's GJKSolver
    runs 128 GJK iterations and EPA with 256 faces / 255 iterations
    (gjk_solver-inl.h:89, 1124-1127)."""
    r = mfcl.request(mfcl.MODE_COLLIDE)
    assert (r.gjk_max_iterations, r.gjk_tolerance) == (128, 1e-6)
    assert (r.epa_max_faces, r.epa_max_iterations, r.epa_tolerance) == (256, 255, 1e-6)
    assert r.max_contacts == 1 and r.penetration == 0
    s = mfcl.request(mfcl.MODE_SIGNED_DISTANCE)
    assert s.mode == mfcl.MODE_SIGNED_DISTANCE and s.penetration == 1


def test_product_has_no_oracle_dependency():
    """The product package and library never reference the oracle."""
    pkg = os.path.join(ROOT, "mind-fcl_amd")
    for dirpath, _, files in os.walk(pkg):
        if "_build" in dirpath:
            continue
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "pyoracle" not in text and "liborc" not in text and "oracle/" not in text, f
    out = subprocess.run(["nm", "-D", mfcl.LIB_PATH], capture_output=True, text=True, check=True).stdout
    assert "orc_" not in out


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(mfcl, "_lib", None)
    monkeypatch.setattr(mfcl, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(RuntimeError):
        mfcl.load_library()


def test_dtype_views_roundtrip():
    a = np.zeros(3, mfcl.CONVEX_RESULT_DTYPE)
    a["distance"] = [1.0, -2.0, 3.5]
    b = np.frombuffer(a.tobytes(), mfcl.CONVEX_RESULT_DTYPE)
    assert np.array_equal(b["distance"], a["distance"])
