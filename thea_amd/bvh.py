"""Python mirror of Thea's BVH query and mesh sampling interface over libthea_gpu's C ABI.

Method names follow Thea (Algorithms/BvhN.hpp, Algorithms/MeshBvh.hpp, Algorithms/MeshSampler.hpp)
and the north-star spellings (closestElement, closestPoint, kClosestElements, rayIntersection,
sampleUniformRandom). Every method takes a batch of queries. Arrays may be numpy arrays (results
come back as numpy, the call is synchronous) or torch tensors on the GPU (results are torch tensors
on the same device, computed asynchronously on torch's current stream).

Return conventions follow Thea: a query with no result has index -1 and distance/time -1
(BvhN.hpp:1062, :1315). Exact ties resolve to the lowest element index.
"""
import ctypes as C

import numpy as np

from . import _capi as capi


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _torch():
    import torch  # noqa: WPS433 (imported only when torch tensors are used)
    return torch


class _Args:
    """Marshals one call's arrays: host (numpy) or device (torch) pointers, plus result arrays."""

    def __init__(self, device_mode, torch_device=None):
        self.device_mode = device_mode
        self.torch_device = torch_device
        self.keep = []

    @property
    def mem(self):
        return capi.THG_MEM_DEVICE if self.device_mode else capi.THG_MEM_HOST

    @property
    def stream(self):
        if not self.device_mode:
            return None
        torch = _torch()
        return C.c_void_p(torch.cuda.current_stream(self.torch_device).cuda_stream)

    def inp(self, a, dtype, cols=None):
        if self.device_mode:
            torch = _torch()
            tdt = {np.float32: torch.float32, np.float64: torch.float64, np.int32: torch.int32}[dtype]
            if not _is_torch(a):
                a = torch.as_tensor(np.asarray(a), device=self.torch_device)
            t = a.to(device=self.torch_device, dtype=tdt).contiguous()
            if cols:
                t = t.reshape(-1, cols)
            self.keep.append(t)
            return t, C.c_void_p(t.data_ptr()), t.shape[0]
        arr = np.ascontiguousarray(np.asarray(a), dtype=dtype)
        if cols:
            arr = arr.reshape(-1, cols)
        self.keep.append(arr)
        return arr, arr.ctypes.data_as(C.c_void_p), arr.shape[0]

    def out(self, shape, dtype, want=True):
        if not want:
            return None, None
        if self.device_mode:
            torch = _torch()
            tdt = {np.float32: torch.float32, np.float64: torch.float64, np.int32: torch.int32,
                   np.uint8: torch.uint8}[dtype]
            t = torch.empty(shape, dtype=tdt, device=self.torch_device)
            return t, C.c_void_p(t.data_ptr()) if t.numel() else None
        arr = np.empty(shape, dtype=dtype)
        return arr, arr.ctypes.data_as(C.c_void_p) if arr.size else None


class Context:
    """A device context (one per GPU)."""

    def __init__(self, device=0):
        self.device = device
        h = C.c_void_p()
        capi.check(capi.lib().thg_ctx_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            capi.lib().thg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_timing(self, enable=True):
        """Record HIP events around every query/sampling kernel launched through this context."""
        capi.check(capi.lib().thg_ctx_set_timing(self.h, 1 if enable else 0))

    def kernel_time(self, kernel, reset=True):
        """(summed milliseconds, launch count) of `kernel` since the last reset."""
        ms, cnt = C.c_double(), C.c_int64()
        capi.check(capi.lib().thg_ctx_kernel_time(self.h, kernel.encode(), C.byref(ms), C.byref(cnt),
                                                  1 if reset else 0))
        return ms.value, cnt.value

    def torch_device(self):
        return _torch().device("cuda", self.device)


_default_ctx = {}


def default_context(device=0):
    if device not in _default_ctx:
        _default_ctx[device] = Context(device)
    return _default_ctx[device]


class Mesh:
    """Triangle mesh resident in device memory (MeshBvh::add / MeshTriangles, MeshBvh.hpp:55-67)."""

    def __init__(self, vertices, triangles, ctx=None):
        self.ctx = ctx or default_context()
        dev = _is_torch(vertices) and vertices.is_cuda
        a = _Args(dev, self.ctx.torch_device() if dev else None)
        _, vp, nv = a.inp(vertices, np.float32, 3)
        _, tp, nt = a.inp(triangles, np.int32, 3)
        h = C.c_void_p()
        capi.check(capi.lib().thg_mesh_create(self.ctx.h, vp, nv, tp, nt, a.mem, C.byref(h)))
        if dev:
            _torch().cuda.synchronize(self.ctx.device)
        self.h = h
        self.num_vertices, self.num_triangles = int(nv), int(nt)

    def close(self):
        if getattr(self, "h", None):
            capi.lib().thg_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def total_area(self):
        a = C.c_double()
        capi.check(capi.lib().thg_mesh_total_area(self.h, C.byref(a)))
        return a.value


class _BvhBase:
    points = False

    def _finish_build(self, h):
        self.h = h
        info = (C.c_int64 * 6)()
        capi.check(capi.lib().thg_bvh_info(self.h, info))
        self.info = {"nodes": info[0], "leaves": info[1], "max_depth": info[2], "elements": info[3],
                     "device_bytes": info[4], "leaf_size": info[5]}

    def export(self):
        """(floats, ints, elem_ids): the nodes as float32 (nodes, 16) and int32 (nodes, 16) views
        (child 0 box in floats 0-5, child 1 box in 6-11; refs and counts in ints 12-15) and the caller
        indices of the elements in leaf order."""
        raw = np.zeros((self.info["nodes"], 16), np.int32)
        ids = np.zeros(self.info["elements"], np.int32)
        capi.check(capi.lib().thg_bvh_export(self.h, raw.ctypes.data_as(C.c_void_p), ids.ctypes.data_as(C.c_void_p)))
        return raw.view(np.float32), raw, ids

    def exportRayTree(self):
        """The mesh BVH's 4-wide quantized ray tree as an int32 array (nodes, 16) (layout in
        thg_bvh_export_ray_tree); empty for point BVHs."""
        cnt = C.c_int64()
        capi.check(capi.lib().thg_bvh_export_ray_tree(self.h, None, C.byref(cnt)))
        raw = np.zeros((cnt.value, 16), np.int32)
        if cnt.value:
            capi.check(capi.lib().thg_bvh_export_ray_tree(self.h, raw.ctypes.data_as(C.c_void_p), C.byref(cnt)))
        return raw

    def getBounds(self):
        """(lo, hi) float32: BvhN::getBounds (BvhN.hpp:991), the union of the element boxes scaled by
        1.05 about its centre in float arithmetic (AxisAlignedBoxN.hpp:274,316-324; BvhN.hpp:1437)."""
        lo, hi = np.zeros(3, np.float32), np.zeros(3, np.float32)
        capi.check(capi.lib().thg_bvh_bounds(self.h, lo.ctypes.data_as(C.c_void_p), hi.ctypes.data_as(C.c_void_p)))
        if self.info["elements"] == 0:
            return lo, hi
        c = np.float32(0.5) * (hi + lo)
        f = np.float32(1.05)
        return f * (lo - c) + c, f * (hi - c) + c

    def close(self):
        if getattr(self, "h", None):
            capi.lib().thg_bvh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _args(self, x):
        dev = _is_torch(x) and x.is_cuda
        return _Args(dev, self.ctx.torch_device() if dev else None)

    # -- closest element (BvhN::closestElement / closestPair, BvhN.hpp:1050-1114) -----------------
    def closest(self, queries, dist_bound=-1.0, want_point=True, want_dist=True, want_d2=False):
        """Returns (idx, dist, closest_point, d2) for a batch; unrequested outputs are None."""
        a = self._args(queries)
        _, qp, n = a.inp(queries, np.float32, 3)
        idx, ip = a.out((n,), np.int32)
        dist, dp = a.out((n,), np.float32, want_dist)
        cp, cpp = a.out((n, 3), np.float32, want_point)
        d2, d2p = a.out((n,), np.float32, want_d2)
        capi.check(capi.lib().thg_closest_point(self.h, qp, n, float(dist_bound), ip, cpp, dp, d2p, a.mem, a.stream))
        return idx, dist, cp, d2

    def closestElement(self, queries, dist_bound=-1.0):
        """Index of the nearest element per query (-1 if none within dist_bound)."""
        return self.closest(queries, dist_bound, want_point=False, want_dist=False)[0]

    def closestElementWithDistance(self, queries, dist_bound=-1.0):
        idx, dist, _, _ = self.closest(queries, dist_bound, want_point=False)
        return idx, dist

    def closestPoint(self, queries, dist_bound=-1.0):
        """(index, distance, closest point) per query: closestElement(..., &dist, &closest_point)."""
        idx, dist, cp, _ = self.closest(queries, dist_bound)
        return idx, dist, cp

    # -- k nearest (BvhN::kClosestPairs, BvhN.hpp:1139-1146) ---------------------------------------
    def kClosestElements(self, queries, k, dist_bound=-1.0, exclude_self=False, want_d2=False):
        """(idx[n,k], dist[n,k]) sorted by (distance, index); missing entries are -1."""
        a = self._args(queries)
        _, qp, n = a.inp(queries, np.float32, 3)
        idx, ip = a.out((n, k), np.int32)
        dist, dp = a.out((n, k), np.float32)
        d2, d2p = a.out((n, k), np.float32, want_d2)
        capi.check(capi.lib().thg_k_closest(self.h, qp, n, int(k), float(dist_bound), int(bool(exclude_self)), ip, dp,
                                            d2p, a.mem, a.stream))
        return (idx, dist, d2) if want_d2 else (idx, dist)

    kClosestPairs = kClosestElements


class MeshBvh(_BvhBase):
    """BVH over a triangle mesh (Thea's MeshBvh<Mesh>, Algorithms/MeshBvh.hpp:31-165)."""

    def __init__(self, mesh, leaf_size=4):
        self.mesh = mesh
        self.ctx = mesh.ctx
        opts = capi.BuildOpts(leaf_size=leaf_size)
        h = C.c_void_p()
        capi.check(capi.lib().thg_bvh_build_mesh(mesh.h, C.byref(opts), None, C.byref(h)))
        self._finish_build(h)

    def shapeDiameter(self, positions, normals=None, scale=None, only_hit_interior_surfaces=True,
                      want_ray_times=False):
        """ShapeDiameter::compute (SurfaceFeatures/Local/ShapeDiameter.hpp:162-256) at each position:
        SDF in [0, 1] (float64), -1 where no ray counts. normals=None takes the normal of the nearest
        face. scale: the normalisation length; None or <= 0 takes the diagonal of getBounds(), the
        default of Thea's constructor over a precomputed BVH (ShapeDiameter.hpp:110-120; its mesh
        constructors default to the bounding-sphere diameter instead). want_ray_times adds the 30
        per-point ray values (-1: not counted). Results are numpy arrays for host inputs and device
        tensors for device inputs."""
        if scale is None or scale <= 0:
            lo, hi = self.getBounds()
            e = hi - lo
            scale = float(np.sqrt(e[0] * e[0] + (e[1] * e[1] + e[2] * e[2])))
            if not scale > 0:
                raise ValueError("the BVH has an empty box; pass a positive scale")
        dev = _is_torch(positions) and positions.is_cuda
        a = _Args(dev, self.ctx.torch_device() if dev else None)
        _, pp, n = a.inp(positions, np.float32, 3)
        npp = None
        if normals is not None:
            _, npp, nn = a.inp(normals, np.float32, 3)
            if nn != n:
                raise ValueError("one normal per position")
        sdf, sp = a.out((n,), np.float64)
        times, tp = a.out((n, 30), np.float32, want_ray_times)
        capi.check(capi.lib().thg_shape_diameter(self.h, pp, npp, n, float(scale), 1 if only_hit_interior_surfaces else 0,
                                                 a.mem, a.stream, sp, tp))
        return (sdf, times) if want_ray_times else sdf

    def visibility(self, positions, num_rays=100, seed=0, scale=None):
        """Visibility::compute (SurfaceFeatures/Local/Visibility.hpp:99-125) at each position: the
        fraction (float64) of num_rays sphere-uniform rays that escape the mesh. Directions come from
        Philox with `seed` (Thea uses its time-seeded global generator). scale None: the diagonal of
        getBounds(), as Thea. numpy in, numpy out; device tensors in, device tensor out."""
        dev = _is_torch(positions) and positions.is_cuda
        a = _Args(dev, self.ctx.torch_device() if dev else None)
        _, pp, n = a.inp(positions, np.float32, 3)
        vis, vp_ = a.out((n,), np.float64)
        capi.check(capi.lib().thg_visibility(self.h, pp, n, int(num_rays), C.c_uint64(seed),
                                             -1.0 if scale is None else float(scale), a.mem, a.stream, vp_))
        return vis

    # -- rays (BvhN::rayStructureIntersection / rayIntersectionTime / rayIntersects) --------------
    def rayStructureIntersection(self, origins, directions, max_time=-1.0, want_normal=True):
        """(idx, t, normal) of the first hit per ray; misses give idx -1, t -1."""
        a = self._args(origins)
        _, op, n = a.inp(origins, np.float32, 3)
        _, dp_, n2 = a.inp(directions, np.float32, 3)
        if n != n2:
            raise ValueError("origins and directions differ in length")
        idx, ip = a.out((n,), np.int32)
        t, tp = a.out((n,), np.float32)
        nrm, np_ = a.out((n, 3), np.float32, want_normal)
        capi.check(capi.lib().thg_ray_first_hit(self.h, op, dp_, n, float(max_time), ip, tp, np_, a.mem, a.stream))
        return idx, t, nrm

    rayIntersection = rayStructureIntersection

    def rayIntersectionTime(self, origins, directions, max_time=-1.0):
        return self.rayStructureIntersection(origins, directions, max_time, want_normal=False)[1]

    def rayIntersects(self, origins, directions, max_time=-1.0):
        a = self._args(origins)
        _, op, n = a.inp(origins, np.float32, 3)
        _, dp_, _ = a.inp(directions, np.float32, 3)
        hit, hp = a.out((n,), np.uint8)
        capi.check(capi.lib().thg_ray_any_hit(self.h, op, dp_, n, float(max_time), hp, a.mem, a.stream))
        return hit


class PointBvh(_BvhBase):
    """BVH over points (Thea's BvhN<Vector3, 3> over a point set)."""

    points = True

    def __init__(self, points, leaf_size=4, ctx=None):
        self.ctx = ctx or default_context()
        a = self._args(points)
        _, pp, n = a.inp(points, np.float32, 3)
        opts = capi.BuildOpts(leaf_size=leaf_size)
        h = C.c_void_p()
        capi.check(capi.lib().thg_bvh_build_points(self.ctx.h, pp, n, a.mem, C.byref(opts), a.stream, C.byref(h)))
        self._finish_build(h)

    def proximityGraph(self, normals=None, max_degree=8, separation_trials=100, seed=0, average=True,
                       num_original=None, surface=None):
        """PointSet3::updateGraph (PointSet3.hpp:324-386) over these points as the samples:
        (neighbour indices (S, max_degree), separations (S, max_degree), degrees (S,),
        (initial separation estimate, average edge length)). With num_original = S < n the first S
        points are the samples and the rest their oversampling (PointSet3::addOversampling): the
        graph is built over all points and each sample keeps its max_degree nearest other samples
        by shortest-path distance (extractOriginalAdjacencies, PointSet3.cpp:158-178). Outputs are
        numpy arrays, or torch tensors when the normals are a torch device tensor. average=False
        skips the averages (the final one is a sequential float sum on the host, as in Thea) and
        returns None for them. surface: a MeshBvh for the reachability test of
        updateGraph(surface) (PointSet3.hpp:633-641; needs normals)."""
        n = self.info["elements"]
        dev = _is_torch(normals) and normals.is_cuda
        a = _Args(dev, self.ctx.torch_device() if dev else None)
        if normals is not None:
            _, np_, nn = a.inp(normals, np.float32, 3)
            if nn != n:
                raise ValueError("one normal per point")
        else:
            np_ = None
        S = n if num_original is None else int(num_original)
        idx, ip = a.out((S, max_degree), np.int32)
        sep, sp = a.out((S, max_degree), np.float32)
        deg, dp = a.out((S,), np.int32)
        avg = (C.c_float * 2)()
        o = capi.GraphOpts(max_degree=max_degree, separation_trials=separation_trials, seed=seed,
                           num_original=0 if num_original is None else S,
                           surface=None if surface is None else surface.h)
        capi.check(capi.lib().thg_proximity_graph(self.h, np_, C.byref(o), a.mem, a.stream, ip, sp, dp,
                                                  avg if average else None))
        return idx, sep, deg, ((avg[0], avg[1]) if average else None)


class PointBvhF64(PointBvh):
    """BVH over double-precision points (Thea's BvhN<Vector3d, 3, double> over a point set, the
    target structure of ShapeAlign's ICP, Tools/ShapeAlign/ShapeAlign.cpp:339). closest() searches
    in double; the float methods inherited from PointBvh search the points rounded to float."""

    def __init__(self, points, leaf_size=4, ctx=None):
        self.ctx = ctx or default_context()
        a = self._args(points)
        _, pp, n = a.inp(points, np.float64, 3)
        opts = capi.BuildOpts(leaf_size=leaf_size)
        h = C.c_void_p()
        capi.check(capi.lib().thg_bvh_build_points_f64(self.ctx.h, pp, n, a.mem, C.byref(opts), a.stream,
                                                       C.byref(h)))
        self._finish_build(h)

    def closest(self, queries, dist_bound=-1.0, want_point=True, want_dist=True, want_d2=False):
        """(idx, dist, closest_point, d2) in double (BvhN<Vector3d>::closestElement, BvhN.hpp:1050-1063);
        unrequested outputs are None."""
        a = self._args(queries)
        _, qp, n = a.inp(queries, np.float64, 3)
        idx, ip = a.out((n,), np.int32)
        dist, dp = a.out((n,), np.float64, want_dist)
        cp, cpp = a.out((n, 3), np.float64, want_point)
        d2, d2p = a.out((n,), np.float64, want_d2)
        capi.check(capi.lib().thg_closest_point_f64(self.h, qp, n, float(dist_bound), ip, cpp, dp, d2p, a.mem,
                                                    a.stream))
        return idx, dist, cp, d2


class MeshSampler:
    """Area-weighted sampling (Thea's MeshSampler::sampleEvenlyByArea, CountMode::EXACT)."""

    def __init__(self, mesh):
        self.mesh = mesh

    def sampleEvenlyByArea(self, n, seed=0, stream_id=0, device=False, count_mode="exact"):
        """(positions[n,3], face_normals[n,3], triangle_indices[n]); fewer rows for an empty mesh.
        count_mode "approximate" (Thea's CountMode::APPROXIMATE) returns about n samples, in
        triangle order."""
        a = _Args(device, self.mesh.ctx.torch_device() if device else None)
        got = C.c_int64()
        if count_mode == "approximate":
            capi.check(capi.lib().thg_sample_area_approx(self.mesh.h, int(n), C.c_uint64(seed), C.c_uint64(stream_id),
                                                         0, None, None, None, C.byref(got), a.mem, None))
            m = got.value
            pos, pp = a.out((m, 3), np.float32)
            nrm, np_ = a.out((m, 3), np.float32)
            tri, tp = a.out((m,), np.int32)
            capi.check(capi.lib().thg_sample_area_approx(self.mesh.h, int(n), C.c_uint64(seed), C.c_uint64(stream_id),
                                                         m, pp, np_, tp, C.byref(got), a.mem, a.stream))
            return pos, nrm, tri
        if count_mode != "exact":
            raise ValueError("count_mode is 'exact' or 'approximate'")
        pos, pp = a.out((n, 3), np.float32)
        nrm, np_ = a.out((n, 3), np.float32)
        tri, tp = a.out((n,), np.int32)
        capi.check(capi.lib().thg_sample_area(self.mesh.h, int(n), C.c_uint64(seed), C.c_uint64(stream_id), pp, np_,
                                              tp, C.byref(got), a.mem, a.stream))
        m = got.value
        return pos[:m], nrm[:m], tri[:m]

    sampleUniformRandom = sampleEvenlyByArea

    def sampleEvenlyBySeparation(self, n, seed=0, stream_id=0, oversampling_factor=-1.0, max_degree=8,
                                 separation_trials=100, graph_seed=0, device=False):
        """MeshSampler::sampleEvenlyBySeparation (MeshSampler.hpp:270-329): (positions[n,3],
        face_normals[n,3], triangle_indices[n]) of n furthest-point picks among ceil(factor * n)
        area samples (sampleEvenlyByArea with seed/stream_id; factor < 0: 3). Empty arrays when the
        mesh has no area or fewer than n picks were possible, as Thea returns 0."""
        a = _Args(device, self.mesh.ctx.torch_device() if device else None)
        pos, pp = a.out((n, 3), np.float32)
        nrm, np_ = a.out((n, 3), np.float32)
        tri, tp = a.out((n,), np.int32)
        got = C.c_int64()
        o = capi.GraphOpts(max_degree=max_degree, separation_trials=separation_trials, seed=graph_seed)
        capi.check(capi.lib().thg_sample_separation(self.mesh.h, int(n), float(oversampling_factor), C.c_uint64(seed),
                                                    C.c_uint64(stream_id), C.byref(o), pp, np_, tp, C.byref(got),
                                                    a.mem, a.stream))
        m = got.value
        return pos[:m], nrm[:m], tri[:m]


def furthestPointSampling(points, k, max_degree=8, separation_trials=100, graph_seed=0, ctx=None):
    """FurthestPointSampling::subsample with geodesic distances (FurthestPointSampling.cpp:47-125):
    indices (int64) of k points picked greedily along the points' proximity graph, starting with
    point 0. Fewer than k when every remaining point is at distance 0. Device tensors in, device
    tensor out."""
    ctx = ctx or default_context()
    dev = _is_torch(points) and points.is_cuda
    a = _Args(dev, ctx.torch_device() if dev else None)
    _, pp, n = a.inp(points, np.float32, 3)
    if dev:
        sel = _torch().empty((k,), dtype=_torch().int64, device=ctx.torch_device())
        sp = C.c_void_p(sel.data_ptr()) if k else None
    else:
        sel = np.empty(k, np.int64)
        sp = sel.ctypes.data_as(C.c_void_p) if k else None
    got = C.c_int64()
    o = capi.GraphOpts(max_degree=max_degree, separation_trials=separation_trials, seed=graph_seed)
    capi.check(capi.lib().thg_furthest_point_sampling(ctx.h, pp, n, int(k), C.byref(o), a.mem, a.stream, sp,
                                                      C.byref(got)))
    return sel[:got.value]


def version():
    return capi.lib().thg_version().decode()
