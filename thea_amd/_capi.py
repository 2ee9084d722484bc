"""ctypes binding of libthea_gpu.so (the C ABI declared in include/thea_gpu.h).

The library is loaded from this package directory only; if it is missing or fails to load, every
call raises -- there is no CPU fallback.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libthea_gpu.so")

THG_OK = 0
THG_E_INVALID = -1
THG_E_OOM = -2
THG_E_DEVICE = -3
THG_E_BUDGET = -4
THG_E_UNSUPPORTED = -5
THG_E_CANCELLED = -6
THG_MEM_DEVICE = 0
THG_MEM_HOST = 1
THG_NOALLOC_MAX_QUERIES = 4194304  # include/thea_gpu.h

_ERR_NAMES = {
    THG_E_INVALID: "THG_E_INVALID", THG_E_OOM: "THG_E_OOM", THG_E_DEVICE: "THG_E_DEVICE",
    THG_E_BUDGET: "THG_E_BUDGET", THG_E_UNSUPPORTED: "THG_E_UNSUPPORTED", THG_E_CANCELLED: "THG_E_CANCELLED",
}


class TheaGpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


class BuildOpts(C.Structure):
    _fields_ = [("leaf_size", C.c_int32), ("reserved", C.c_int32 * 7)]


class ServerOpts(C.Structure):
    _fields_ = [("memory_budget_bytes", C.c_int64), ("num_streams", C.c_int32), ("max_batch_requests", C.c_int32),
                ("max_batch_queries", C.c_int64), ("coalesce", C.c_int32), ("reserved", C.c_int32 * 7)]


class Request(C.Structure):
    _fields_ = [("type", C.c_int32), ("mesh", C.c_int32), ("n", C.c_int64), ("k", C.c_int32),
                ("exclude_self", C.c_int32), ("bound", C.c_double), ("q", C.c_void_p), ("dirs", C.c_void_p),
                ("idx", C.c_void_p), ("dist", C.c_void_p), ("cp", C.c_void_p)]


class IcpOpts(C.Structure):
    _fields_ = [("fractional_error_threshold", C.c_double), ("min_iterations", C.c_int32),
                ("max_iterations", C.c_int32), ("has_up", C.c_int32), ("kabsch", C.c_int32),
                ("up", C.c_double * 3), ("translation_weights", C.c_void_p), ("rotation_weights", C.c_void_p),
                ("reserved", C.c_int64 * 4)]


class GraphOpts(C.Structure):
    _fields_ = [("max_degree", C.c_int32), ("separation_trials", C.c_int32), ("seed", C.c_uint64),
                ("num_original", C.c_int64), ("surface", C.c_void_p), ("reserved", C.c_int64 * 2)]


class MeshFile(C.Structure):
    _fields_ = [("num_vertices", C.c_int64), ("num_faces", C.c_int64), ("num_triangles", C.c_int64),
                ("vertices", C.POINTER(C.c_float)), ("triangles", C.POINTER(C.c_int32)),
                ("triangle_faces", C.POINTER(C.c_int64))]


THG_REQ_CLOSEST, THG_REQ_KNN, THG_REQ_RAY = 0, 1, 2
THG_PENDING = 1


vp, i32, i64, u64, f32, f64 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64, C.c_float, C.c_double

# name -> (restype, argtypes)
SIGNATURES = {
    "thg_last_error": (C.c_char_p, []),
    "thg_version": (C.c_char_p, []),
    "thg_ctx_create": (C.c_int, [i32, C.POINTER(vp)]),
    "thg_ctx_destroy": (C.c_int, [vp]),
    "thg_mesh_create": (C.c_int, [vp, vp, i64, vp, i64, i32, C.POINTER(vp)]),
    "thg_mesh_destroy": (C.c_int, [vp]),
    "thg_mesh_num_triangles": (i64, [vp]),
    "thg_mesh_total_area": (C.c_int, [vp, C.POINTER(f64)]),
    "thg_bvh_build_mesh": (C.c_int, [vp, C.POINTER(BuildOpts), vp, C.POINTER(vp)]),
    "thg_bvh_build_points": (C.c_int, [vp, vp, i64, i32, C.POINTER(BuildOpts), vp, C.POINTER(vp)]),
    "thg_bvh_destroy": (C.c_int, [vp]),
    "thg_bvh_info": (C.c_int, [vp, C.POINTER(i64)]),
    "thg_bvh_export": (C.c_int, [vp, vp, vp]),
    "thg_bvh_export_ray_tree": (C.c_int, [vp, vp, vp]),
    "thg_bvh_bounds": (C.c_int, [vp, vp, vp]),
    "thg_closest_point": (C.c_int, [vp, vp, i64, f64, vp, vp, vp, vp, i32, vp]),
    "thg_closest_point_noalloc": (C.c_int, [vp, vp, i64, f64, vp, vp, vp, vp, vp]),
    "thg_bvh_build_points_f64": (C.c_int, [vp, vp, i64, i32, C.POINTER(BuildOpts), vp, C.POINTER(vp)]),
    "thg_closest_point_f64": (C.c_int, [vp, vp, i64, f64, vp, vp, vp, vp, i32, vp]),
    "thg_k_closest": (C.c_int, [vp, vp, i64, i32, f64, i32, vp, vp, vp, i32, vp]),
    "thg_ray_first_hit": (C.c_int, [vp, vp, vp, i64, f32, vp, vp, vp, i32, vp]),
    "thg_ray_any_hit": (C.c_int, [vp, vp, vp, i64, f32, vp, i32, vp]),
    "thg_sample_area": (C.c_int, [vp, i64, u64, u64, vp, vp, vp, C.POINTER(i64), i32, vp]),
    "thg_sample_area_approx": (C.c_int, [vp, i64, u64, u64, i64, vp, vp, vp, C.POINTER(i64), i32, vp]),
    "thg_closest_stats": (C.c_int, [vp, vp, i64, vp, C.POINTER(i64)]),
    "thg_knn_wave_stats": (C.c_int, [vp, vp, i64, i32, vp, vp]),
    "thg_ray_stats": (C.c_int, [vp, vp, vp, i64, vp, vp, vp]),
    "thg_ctx_set_timing": (C.c_int, [vp, i32]),
    "thg_ctx_kernel_time": (C.c_int, [vp, C.c_char_p, C.POINTER(f64), C.POINTER(i64), i32]),
    "thg_proximity_graph": (C.c_int, [vp, vp, C.POINTER(GraphOpts), i32, vp, vp, vp, vp, vp]),
    "thg_mesh_file_read": (C.c_int, [C.c_char_p, C.POINTER(MeshFile)]),
    "thg_mesh_file_free": (None, [C.POINTER(MeshFile)]),
    "thg_pts_write": (C.c_int, [C.c_char_p, i64, vp, vp, vp]),
    "thg_pts_read": (C.c_int, [C.c_char_p, C.POINTER(i64), C.POINTER(C.POINTER(C.c_float))]),
    "thg_free": (None, [vp]),
    "thg_visibility": (C.c_int, [vp, vp, i64, i32, u64, f32, i32, vp, vp]),
    "thg_furthest_point_sampling": (C.c_int, [vp, vp, i64, i64, C.POINTER(GraphOpts), i32, vp, vp, C.POINTER(i64)]),
    "thg_sample_separation": (C.c_int, [vp, i64, f32, u64, u64, C.POINTER(GraphOpts), vp, vp, vp, C.POINTER(i64), i32,
                                        vp]),
    "thg_shape_diameter": (C.c_int, [vp, vp, vp, i64, C.c_float, i32, i32, vp, vp, vp]),
    "thg_icp_align": (C.c_int, [vp, vp, i64, i32, vp, C.POINTER(IcpOpts), i32, vp, vp, vp, vp]),
    "thg_server_create": (C.c_int, [vp, C.POINTER(ServerOpts), C.POINTER(vp)]),
    "thg_server_destroy": (C.c_int, [vp]),
    "thg_server_add_mesh": (C.c_int, [vp, vp, i64, vp, i64, i32, i32, C.POINTER(i32)]),
    "thg_server_submit": (C.c_int, [vp, C.POINTER(Request), C.POINTER(i64)]),
    "thg_server_wait": (C.c_int, [vp, i64, i64, C.POINTER(f64)]),
    "thg_server_cancel": (C.c_int, [vp, i64]),
    "thg_server_stats": (C.c_int, [vp, C.POINTER(f64), i32]),
    "thg_batch_plan_create": (C.c_int, [vp, i32, vp, vp, vp, vp, C.POINTER(vp)]),
    "thg_batch_plan_run": (C.c_int, [vp, vp, vp, vp, vp, vp, vp, vp]),
    "thg_batch_plan_destroy": (C.c_int, [vp]),
    "thg_autotune_leaf": (C.c_int, [vp, i32, vp, vp, vp, vp, vp, vp, vp]),
    "thg_autotune_plan": (C.c_int, [vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "thg_batch_plan_set_launch": (C.c_int, [vp, vp, i32, i32]),
    "thg_batch_closest": (C.c_int, [vp, i32, vp, vp, vp, vp, vp, vp, i32, vp, vp, vp, vp]),
}

_lib = None


def lib():
    """Load libthea_gpu.so (raises OSError if it is absent or cannot be loaded)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(f"{LIB_PATH} is missing: build it with thea_amd/build.py (no CPU fallback exists)")
    # PyTorch-ROCm ships its own HIP runtime with the same SONAME (libamdhip64.so.7). Loading torch
    # first makes this library bind to that runtime, so one process never holds two HIP runtimes
    # (a second one cannot see the GPU the first has opened).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(code):
    if code != THG_OK:
        msg = lib().thg_last_error()
        raise TheaGpuError(code, msg.decode() if msg else "")
    return code
