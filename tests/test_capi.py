"""CPU tests of the C ABI library: it loads, exports every function include/thea_gpu.h declares,
and reports argument errors through status codes and thg_last_error() without a GPU."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "thea_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(thg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_expected_entry_points():
    names = declared_functions()
    for must in ("thg_closest_point", "thg_k_closest", "thg_ray_first_hit", "thg_ray_any_hit", "thg_sample_area",
                 "thg_bvh_build_mesh", "thg_bvh_build_points", "thg_mesh_create", "thg_batch_closest"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from thea_amd import _capi
    lib = _capi.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", _capi.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (thg_\w+)", out))
    assert set(declared_functions()) <= exported
    # every exported thg_ symbol is declared in the header (no undocumented entry points)
    assert exported <= set(declared_functions()), exported - set(declared_functions())


def test_python_binding_covers_header():
    from thea_amd import _capi
    assert set(declared_functions()) == set(_capi.SIGNATURES)
    m = re.search(r"#define THG_NOALLOC_MAX_QUERIES (\d+)", open(HEADER).read())
    assert m and int(m.group(1)) == _capi.THG_NOALLOC_MAX_QUERIES


def test_noalloc_entry_point_checks_arguments_first():
    """The no-allocation entry point rejects a NULL handle and a negative count before any device work."""
    from thea_amd import _capi
    lib = _capi.lib()
    assert lib.thg_closest_point_noalloc(None, None, 0, -1.0, None, None, None, None, None) == _capi.THG_E_INVALID
    assert b"NULL" in lib.thg_last_error()


def test_null_handles_are_invalid_arguments():
    from thea_amd import _capi
    lib = _capi.lib()
    out = C.c_void_p()
    assert lib.thg_mesh_create(None, None, 0, None, 0, 1, C.byref(out)) == _capi.THG_E_INVALID
    assert b"NULL" in lib.thg_last_error()
    assert lib.thg_closest_point(None, None, 0, -1.0, None, None, None, None, 1, None) == _capi.THG_E_INVALID
    assert lib.thg_k_closest(None, None, 0, 4, -1.0, 0, None, None, None, 1, None) == _capi.THG_E_INVALID
    assert lib.thg_ray_first_hit(None, None, None, 0, -1.0, None, None, None, 1, None) == _capi.THG_E_INVALID
    assert lib.thg_sample_area(None, 0, 0, 0, None, None, None, None, 1, None) == _capi.THG_E_INVALID
    assert lib.thg_bvh_info(None, None) == _capi.THG_E_INVALID
    assert lib.thg_version().startswith(b"thea_gpu")


def test_null_handles_of_the_feature_entry_points():
    """The ray features, separation sampling and APPROXIMATE sampling reject NULL handles and bad
    sizes before touching a device."""
    from thea_amd import _capi
    lib = _capi.lib()
    got = C.c_int64()
    E = _capi.THG_E_INVALID
    assert lib.thg_shape_diameter(None, None, None, 0, 1.0, 1, 1, None, None, None) == E
    assert lib.thg_visibility(None, None, 0, 100, 0, -1.0, 1, None, None) == E
    assert lib.thg_furthest_point_sampling(None, None, 0, 0, None, 1, None, None, C.byref(got)) == E
    assert b"NULL" in lib.thg_last_error()
    assert lib.thg_sample_separation(None, 10, -1.0, 0, 0, None, None, None, None, C.byref(got), 1, None) == E
    assert lib.thg_sample_area_approx(None, 10, 0, 0, 0, None, None, None, C.byref(got), 1, None) == E
    assert lib.thg_proximity_graph(None, None, None, 1, None, None, None, None, None) == E
    assert lib.thg_icp_align(None, None, 0, 1, None, None, 1, None, None, None, None) == E
    assert lib.thg_bvh_bounds(None, None, None) == E


def test_file_entry_points_report_errors(tmp_path):
    from thea_amd import _capi
    lib = _capi.lib()
    m = _capi.MeshFile()
    assert lib.thg_mesh_file_read(str(tmp_path / "missing.obj").encode(), C.byref(m)) == _capi.THG_E_INVALID
    assert b"could not open" in lib.thg_last_error()
    lib.thg_mesh_file_free(C.byref(m))  # zeroed by the failed read: a no-op
    assert lib.thg_pts_write(None, 0, None, None, None) == _capi.THG_E_INVALID
    n = C.c_int64()
    p = C.POINTER(C.c_float)()
    assert lib.thg_pts_read(str(tmp_path / "missing.pts").encode(), C.byref(n), C.byref(p)) == _capi.THG_E_INVALID


def test_no_device_reports_error():
    from thea_amd import _capi
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    lib = _capi.lib()
    h = C.c_void_p()
    rc = lib.thg_ctx_create(0, C.byref(h))
    assert rc == _capi.THG_E_DEVICE
    assert len(lib.thg_last_error()) > 0


def test_cpp_shim_compiles(tmp_path):
    # include/thea_gpu_bvh.hpp: Thea-named C++ host API over the C ABI
    src = os.path.join(ROOT, "tests", "shim", "shim_example.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-c", src, "-o", str(tmp_path / "shim.o")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
