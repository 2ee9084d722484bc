"""TEST INFRASTRUCTURE ONLY -- ctypes wrapper of the oracle (liboracle.so).

The oracle is a CPU restatement of Thea's BVH-query and mesh-sampling path (see thea_oracle.cpp).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module, and
only as the checker; the product path (thea_amd) never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")


def build():
    """Compile liboracle.so from thea_oracle.cpp (make, in this directory)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    vp, i64, i32, f32, f64, u32, u64 = C.c_void_p, C.c_int64, C.c_int, C.c_float, C.c_double, C.c_uint32, C.c_uint64
    sig = {
        "or_tri_props": (None, [f32p, i32p, i64, f32p]),
        "or_tri_closest": (f32, [f32p, f32p, f32p]),
        "or_tri_ray": (f32, [f32p, f32p, f32p, f32]),
        "or_kd_build_mesh": (vp, [f32p, i32p, i64, i32]),
        "or_kd_build_points": (vp, [f32p, i64, i32]),
        "or_kd_free": (None, [vp]),
        "or_kd_num_nodes": (i64, [vp]),
        "or_kd_nodes": (None, [vp, f32p, i64p]),
        "or_kd_order": (None, [vp, i64p]),
        "or_kd_closest": (None, [vp, f32p, i64, f64, i32, i32, i32, i32p, vp, vp, vp]),
        "or_kd_knn": (None, [vp, f32p, i64, i32, f64, i32, i32, i32, i32p, f32p, vp]),
        "or_kd_ray": (None, [vp, f32p, f32p, i64, f32, i32, f32, i32, i32p, f32p, vp]),
        "or_brute_closest": (None, [f32p, i32p, i64, f32p, i64, f64, i32, i32p, vp, vp]),
        "or_brute_closest_f64": (None, [f32p, i32p, i64, f32p, i64, i32, i32p, f64p]),
        "or_brute_closest_points_f64": (None, [f64p, i64, f64p, i64, f64, i32, i32p, vp]),
        "or_extract_original": (None, [i32p, f32p, i32p, i64, i32, i64, i32, i32p, f32p, i32p]),
        "or_dist_f64": (None, [f32p, i32p, f32p, i32p, i64, f64p]),
        "or_d2_to": (None, [f32p, i32p, f32p, i32p, i64, f32p]),
        "or_brute_knn": (None, [f32p, i64, f32p, i64, i32, f64, i32, i32, i32p, f32p]),
        "or_brute_ray": (None, [f32p, i32p, i64, f32p, f32p, i64, f32, i32, i32p, f32p]),
        "or_ray_f64": (None, [f32p, i32p, f32p, f32p, i32p, i64, f64p]),
        "or_philox": (None, [u32p, u32p, u32p]),
        "or_u01": (f32, [u32]),
        "or_thea_mt": (None, [u32, i64, u32p]),
        "or_thea_mt_u01": (None, [u32, i64, f32p]),
        "or_sample_area": (i64, [f32p, i32p, i64, i64, u64, u64, f32p, vp, vp, vp]),
        "or_geodesic_sphere": (i64, [i32, f32p, i64, i32p, i64, vp]),
        "or_fast_sqrt": (None, [f32p, i64, f32p]),
        "or_sdf_rays": (None, [f32p, f32p, i64, f32, f32p, f32p]),
        "or_sdf_reduce": (None, [f32p, i64, f64, f64p]),
        "or_proximity_graph": (i32, [f32p, vp, i64, i32, i64p, i64, i32p, f32p, i32p, f32p, vp]),
        "or_sample_area_approx": (i64, [f32p, i32p, i64, i64, u64, u64, i64, f32p, vp, vp]),
        "or_fps": (i64, [i32p, f32p, i64, i32, i64, i64p]),
        "or_vis_rays": (None, [f32p, i64, i32, f32, u64, f32p, f32p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def default_threads():
    return max(1, min(16, os.cpu_count() or 1))


def tri_props(verts, tris):
    verts, tris = _f32(verts), _i32(tris)
    out = np.zeros((len(tris), 13), np.float32)
    lib().or_tri_props(verts, tris, len(tris), out)
    return out


def tri_closest(v9, q):
    cp = np.zeros(3, np.float32)
    d2 = lib().or_tri_closest(_f32(v9).reshape(9), _f32(q).reshape(3), cp)
    return np.float32(d2), cp


def tri_ray(v9, o, d, max_time=-1.0):
    return np.float32(lib().or_tri_ray(_f32(v9).reshape(9), _f32(o).reshape(3), _f32(d).reshape(3), max_time))


class KdTree:
    """Thea's kd-tree (median split on the longest axis, <= leaf elements per leaf, x1.05 padding)."""

    def __init__(self, verts=None, tris=None, points=None, leaf=5):
        if points is not None:
            self.points = _f32(points).reshape(-1, 3)
            self.h = lib().or_kd_build_points(self.points, len(self.points), leaf)
        else:
            self.verts, self.tris = _f32(verts).reshape(-1, 3), _i32(tris).reshape(-1, 3)
            self.h = lib().or_kd_build_mesh(self.verts, self.tris, len(self.tris), leaf)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.or_kd_free(self.h)
            self.h = None

    def nodes(self):
        n = lib().or_kd_num_nodes(self.h)
        boxes = np.zeros((n, 6), np.float32)
        links = np.zeros((n, 4), np.int64)
        lib().or_kd_nodes(self.h, boxes, links)
        return boxes, links

    def closest(self, q, dist_bound=-1.0, mode=0, exclude_self=False, nthreads=None, want_cp=True):
        """mode 0: canonical (lowest index on exact ties); mode 1: Thea traversal order."""
        q = _f32(q).reshape(-1, 3)
        n = len(q)
        idx = np.zeros(n, np.int32)
        cp = np.zeros((n, 3), np.float32) if want_cp else None
        d2 = np.zeros(n, np.float32)
        stats = np.zeros(3, np.int64)
        lib().or_kd_closest(self.h, q, n, float(dist_bound), mode, int(exclude_self), nthreads or default_threads(),
                            idx, _ptr(cp), _ptr(d2), _ptr(stats))
        return idx, cp, d2, stats

    def knn(self, q, k, dist_bound=-1.0, mode=0, exclude_self=False, nthreads=None):
        q = _f32(q).reshape(-1, 3)
        n = len(q)
        idx = np.zeros((n, k), np.int32)
        d2 = np.zeros((n, k), np.float32)
        stats = np.zeros(3, np.int64)
        lib().or_kd_knn(self.h, q, n, k, float(dist_bound), mode, int(exclude_self), nthreads or default_threads(),
                        idx, d2, _ptr(stats))
        return idx, d2, stats

    def ray(self, orig, dirs, max_time=-1.0, mode=0, slack=None, nthreads=None):
        o, d = _f32(orig).reshape(-1, 3), _f32(dirs).reshape(-1, 3)
        n = len(o)
        if slack is None:
            slack = float(np.abs(self.verts).max()) * 2.0 ** -16 if len(self.verts) else 0.0
        idx = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        stats = np.zeros(3, np.int64)
        lib().or_kd_ray(self.h, o, d, n, float(max_time), mode, float(slack), nthreads or default_threads(), idx, t,
                        _ptr(stats))
        return idx, t, stats


def brute_closest(verts, tris, q, dist_bound=-1.0, nthreads=None):
    verts, tris, q = _f32(verts), _i32(tris), _f32(q).reshape(-1, 3)
    n = len(q)
    idx = np.zeros(n, np.int32)
    cp = np.zeros((n, 3), np.float32)
    d2 = np.zeros(n, np.float32)
    lib().or_brute_closest(verts, tris, len(tris), q, n, float(dist_bound), nthreads or default_threads(), idx,
                           _ptr(cp), _ptr(d2))
    return idx, cp, d2


def brute_closest_f64(verts, tris, q, nthreads=None):
    verts, tris, q = _f32(verts), _i32(tris), _f32(q).reshape(-1, 3)
    n = len(q)
    idx = np.zeros(n, np.int32)
    dist = np.zeros(n, np.float64)
    lib().or_brute_closest_f64(verts, tris, len(tris), q, n, nthreads or default_threads(), idx, dist)
    return idx, dist


def brute_closest_points_f64(points, q, dist_bound=-1.0, nthreads=None):
    """Nearest point in double precision (lowest index on exact ties): (idx int32, d2 float64)."""
    pts = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    q = np.ascontiguousarray(q, np.float64).reshape(-1, 3)
    idx = np.zeros(len(q), np.int32)
    d2 = np.zeros(len(q), np.float64)
    lib().or_brute_closest_points_f64(pts, len(pts), q, len(q), float(dist_bound), nthreads or default_threads(), idx,
                                      _ptr(d2))
    return idx, d2


def dist_f64(verts, tris, q, which):
    verts, tris, q, which = _f32(verts), _i32(tris), _f32(q).reshape(-1, 3), _i32(which)
    out = np.zeros(len(q), np.float64)
    lib().or_dist_f64(verts, tris, q, which, len(q), out)
    return out


def d2_to(verts, tris, q, which):
    """Thea's float squared distance from each query to the triangle `which` names (-1 for -1)."""
    verts, tris, q, which = _f32(verts), _i32(tris), _f32(q).reshape(-1, 3), _i32(which)
    out = np.zeros(len(q), np.float32)
    lib().or_d2_to(verts, tris, q, which, len(q), out)
    return out


def brute_knn(points, q, k, dist_bound=-1.0, exclude_self=False, nthreads=None):
    points, q = _f32(points).reshape(-1, 3), _f32(q).reshape(-1, 3)
    n = len(q)
    idx = np.zeros((n, k), np.int32)
    d2 = np.zeros((n, k), np.float32)
    lib().or_brute_knn(points, len(points), q, n, k, float(dist_bound), int(exclude_self),
                       nthreads or default_threads(), idx, d2)
    return idx, d2


def brute_ray(verts, tris, orig, dirs, max_time=-1.0, nthreads=None):
    verts, tris = _f32(verts), _i32(tris)
    o, d = _f32(orig).reshape(-1, 3), _f32(dirs).reshape(-1, 3)
    n = len(o)
    idx = np.zeros(n, np.int32)
    t = np.zeros(n, np.float32)
    lib().or_brute_ray(verts, tris, len(tris), o, d, n, float(max_time), nthreads or default_threads(), idx, t)
    return idx, t


def ray_f64(verts, tris, orig, dirs, which):
    verts, tris = _f32(verts), _i32(tris)
    o, d, which = _f32(orig).reshape(-1, 3), _f32(dirs).reshape(-1, 3), _i32(which)
    out = np.zeros(len(o), np.float64)
    lib().or_ray_f64(verts, tris, o, d, which, len(o), out)
    return out


def philox(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().or_philox(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def u01(bits):
    return np.float32(lib().or_u01(int(bits)))


def thea_mt(seed, n):
    out = np.zeros(n, np.uint32)
    lib().or_thea_mt(seed, n, out)
    return out


def thea_mt_u01(seed, n):
    out = np.zeros(n, np.float32)
    lib().or_thea_mt_u01(seed, n, out)
    return out


def sample_area(verts, tris, n, seed, stream=0, want_cdf=False):
    verts, tris = _f32(verts), _i32(tris)
    pos = np.zeros((max(n, 0), 3), np.float32)
    nrm = np.zeros((max(n, 0), 3), np.float32)
    tri = np.zeros(max(n, 0), np.int32)
    cdf = np.zeros(len(tris), np.float64) if want_cdf else None
    m = lib().or_sample_area(verts, tris, len(tris), n, seed, stream, pos, _ptr(nrm), _ptr(tri), _ptr(cdf))
    res = (pos[:m], nrm[:m], tri[:m])
    return res + (cdf,) if want_cdf else res


def sample_area_approx(verts, tris, n, seed, stream=0):
    """APPROXIMATE-mode area sampling restated (see or_sample_area_approx): (pos, nrm, tri)."""
    verts, tris = _f32(verts), _i32(tris)
    cap = int(n + 10 * np.sqrt(max(n, 1)) + len(tris) + 16)
    pos = np.zeros((cap, 3), np.float32)
    nrm = np.zeros((cap, 3), np.float32)
    tri = np.zeros(cap, np.int32)
    m = lib().or_sample_area_approx(verts, tris, len(tris), n, seed, stream, cap, pos, _ptr(nrm), _ptr(tri))
    assert m <= cap
    return pos[:m], nrm[:m], tri[:m]


def geodesic_sphere(depth):
    nv = 10 * 4 ** depth + 2
    nt = 20 * 4 ** depth
    v = np.zeros((nv, 3), np.float32)
    t = np.zeros((nt, 3), np.int32)
    ntr = np.zeros(1, np.int64)
    got = lib().or_geodesic_sphere(depth, v, nv, t, nt, _ptr(ntr))
    return v[:got], t[: int(ntr[0])]


def fast_sqrt(x):
    """Thea's Math::fastSqrt for float32 (Quake III reciprocal square root, one Newton step)."""
    x = _f32(x).ravel()
    out = np.zeros_like(x)
    lib().or_fast_sqrt(x, len(x), out)
    return out


def graph_picks(n, trials, seed):
    """The separation-estimate picks of the proximity graph: pick t = (w * n) >> 32 for the first word
    w of Philox4x32-10 at counter (t, 0, 0x47524150, 0), key (seed lo, seed hi)."""
    t = min(trials, n)
    key = [seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF]
    return np.array([(int(philox([i & 0xFFFFFFFF, i >> 32, 0x47524150, 0], key)[0]) * n) >> 32 for i in range(t)],
                    np.int64)


def proximity_graph(points, normals=None, max_degree=8, picks=None, surface=None):
    """PointSet3::updateGraph restated (see thea_oracle.cpp): (idx, sep, degree, (initial, final avg)).
    surface: a mesh KdTree for the reachability test (used with normals)."""
    pts = _f32(points).reshape(-1, 3)
    n = len(pts)
    nrm = None if normals is None else _f32(normals).reshape(-1, 3)
    picks = np.ascontiguousarray(picks, np.int64)
    idx = np.zeros((n, max_degree), np.int32)
    sep = np.zeros((n, max_degree), np.float32)
    deg = np.zeros(n, np.int32)
    avg = np.zeros(2, np.float32)
    rc = lib().or_proximity_graph(pts, None if nrm is None else nrm.ctypes.data_as(C.c_void_p), n, max_degree, picks,
                                  len(picks), idx, sep, deg, avg, None if surface is None else surface.h)
    if rc != 0:
        raise ValueError("or_proximity_graph failed")
    return idx, sep, deg, (float(avg[0]), float(avg[1]))


def extract_original(nbr_idx, nbr_sep, degree, num_original, nthreads=None):
    """PointSet3::extractOriginalAdjacencies restated (see or_extract_original): the graph of all n
    samples (n x md neighbour indices and separations, degrees) -> for the first num_original samples,
    (idx (S, md), sep (S, md) float32, degree (S,))."""
    nbr_idx = _i32(nbr_idx)
    n, md = nbr_idx.shape
    S = int(num_original)
    idx = np.zeros((S, md), np.int32)
    sep = np.zeros((S, md), np.float32)
    deg = np.zeros(S, np.int32)
    lib().or_extract_original(nbr_idx, _f32(nbr_sep), _i32(degree), n, md, S, nthreads or default_threads(), idx, sep,
                              deg)
    return idx, sep, deg


def vis_rays(positions, num_rays, scale, seed):
    """Visibility::compute's rays (see or_vis_rays): origins and directions (n * num_rays, 3)."""
    pos = _f32(positions).reshape(-1, 3)
    m = len(pos) * num_rays
    org = np.zeros((m, 3), np.float32)
    dr = np.zeros((m, 3), np.float32)
    lib().or_vis_rays(pos, len(pos), num_rays, np.float32(scale), seed, org, dr)
    return org, dr


def visibility(kd, positions, num_rays, scale, seed):
    """Visibility::compute restated: the fraction of each point's rays with no hit (kd-tree)."""
    org, dr = vis_rays(positions, num_rays, scale, seed)
    idx, _, _ = kd.ray(org, dr, mode=0)
    escaped = (idx < 0).reshape(-1, num_rays).sum(1)
    return escaped.astype(np.float64) / num_rays


def fps(idx, sep, k):
    """FurthestPointSampling::subsample over a proximity graph (idx, sep: (n, max_degree)) restated
    with a per-pick multi-source Dijkstra (see or_fps). Returns the picked indices (int64)."""
    idx = np.ascontiguousarray(idx, np.int32)
    sep = _f32(sep)
    n, md = idx.shape
    sel = np.zeros(max(k, 1), np.int64)
    got = lib().or_fps(idx, sep, n, md, k, sel)
    return sel[:got]


def shape_diameter(kd, tri_normals, positions, normals, scale, only_interior=True):
    """ShapeDiameter::compute restated: rays from C++ (or_sdf_rays), first hits from the kd-tree
    (canonical), the interior test on the hit face normal, and the reduction (or_sdf_reduce).
    Returns (sdf (n,), ray values (n, 30))."""
    pos = _f32(positions).reshape(-1, 3)
    nrm = _f32(normals).reshape(-1, 3)
    n = len(pos)
    org = np.zeros((30 * n, 3), np.float32)
    dr = np.zeros((30 * n, 3), np.float32)
    lib().or_sdf_rays(pos, nrm, n, np.float32(scale), org, dr)
    idx, t, _ = kd.ray(org, dr, mode=0)
    val = np.where(idx >= 0, t, np.float32(-1)).astype(np.float32)
    if only_interior:
        hn = tri_normals[np.maximum(idx, 0)]
        dot = hn[:, 0] * dr[:, 0] + (hn[:, 1] * dr[:, 1] + hn[:, 2] * dr[:, 2])
        val = np.where((idx >= 0) & (dot >= 0), val, np.float32(-1)).astype(np.float32)
    out = np.zeros(n, np.float64)
    lib().or_sdf_reduce(np.ascontiguousarray(val), n, float(np.float32(scale)), out)
    return out, val.reshape(n, 30)

