"""Python binding of the batched many-mesh closest-point plan (thg_batch_plan_*, thg_autotune_leaf,
thg_autotune_plan)."""
import ctypes as C

import numpy as np

from . import _capi as capi
from .bvh import default_context


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def concat(meshes, queries):
    """Concatenate [(verts, tris)] and [queries] into the flat arrays the plan takes."""
    nv = _i64([len(v) for v, _ in meshes])
    nt = _i64([len(t) for _, t in meshes])
    nq = _i64([len(q) for q in queries])
    verts = np.concatenate([np.asarray(v, np.float32).reshape(-1, 3) for v, _ in meshes]) if len(meshes) else \
        np.zeros((0, 3), np.float32)
    tris = np.concatenate([np.asarray(t, np.int32).reshape(-1, 3) for _, t in meshes])
    qs = np.concatenate([np.asarray(q, np.float32).reshape(-1, 3) for q in queries])
    return nv, nt, nq, verts, tris, qs


class BatchPlan:
    """A reusable plan for one set of mesh/query sizes; run() takes torch device tensors."""

    def __init__(self, nv, nt, nq, leaf_sizes=None, ctx=None, child_orders=None, launch=None):
        """launch: (waves per workgroup, 8-resident-wave build) of the query launch, as autotune_plan
        returns it; child_orders: per mesh (0 majority, 1 nearest, 2 build order)."""
        self.ctx = ctx or default_context()
        self.nv, self.nt, self.nq = _i64(nv), _i64(nt), _i64(nq)
        self.nm = len(self.nt)
        ls = None if leaf_sizes is None else np.ascontiguousarray(leaf_sizes, np.int32)
        h = C.c_void_p()
        capi.check(capi.lib().thg_batch_plan_create(
            self.ctx.h, self.nm, self.nv.ctypes.data_as(C.c_void_p), self.nt.ctypes.data_as(C.c_void_p),
            self.nq.ctypes.data_as(C.c_void_p), None if ls is None else ls.ctypes.data_as(C.c_void_p), C.byref(h)))
        self.h = h
        if child_orders is not None or launch is not None:
            self.set_launch(child_orders, *(launch if launch is not None else (4, 1)))

    def set_launch(self, child_orders=None, block_waves=4, spill_build=1):
        co = None if child_orders is None else np.ascontiguousarray(child_orders, np.int32)
        capi.check(capi.lib().thg_batch_plan_set_launch(self.h, None if co is None else co.ctypes.data_as(C.c_void_p),
                                                        int(block_waves), int(spill_build)))

    def close(self):
        if getattr(self, "h", None):
            capi.lib().thg_batch_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, verts, tris, queries, timed=False):
        """(idx, dist) torch tensors (local triangle index per query); with timed=True an eager run
        that also returns (build_ms, query_ms)."""
        import torch
        n = int(self.nq.sum())
        idx = torch.empty(n, dtype=torch.int32, device=queries.device)
        dist = torch.empty(n, dtype=torch.float32, device=queries.device)
        tm = (C.c_double * 2)()
        stream = C.c_void_p(torch.cuda.current_stream(queries.device).cuda_stream)
        capi.check(capi.lib().thg_batch_plan_run(self.h, C.c_void_p(verts.data_ptr()), C.c_void_p(tris.data_ptr()),
                                                 C.c_void_p(queries.data_ptr()), C.c_void_p(idx.data_ptr()),
                                                 C.c_void_p(dist.data_ptr()), stream, tm if timed else None))
        return (idx, dist, (tm[0], tm[1])) if timed else (idx, dist)


def autotune_leaf(nv, nt, nq, verts, tris, queries, ctx=None):
    """Per-mesh leaf sizes chosen by timing the closest-point kernel per size class."""
    ctx = ctx or default_context()
    nv, nt, nq = _i64(nv), _i64(nt), _i64(nq)
    out = np.zeros(len(nt), np.int32)
    capi.check(capi.lib().thg_autotune_leaf(ctx.h, len(nt), nv.ctypes.data_as(C.c_void_p), nt.ctypes.data_as(C.c_void_p),
                                            nq.ctypes.data_as(C.c_void_p), C.c_void_p(verts.data_ptr()),
                                            C.c_void_p(tris.data_ptr()), C.c_void_p(queries.data_ptr()),
                                            out.ctypes.data_as(C.c_void_p)))
    return out


def autotune_plan(nv, nt, nq, verts, tris, queries, ctx=None):
    """(leaf sizes, child orders, (waves per workgroup, 8-resident-wave build)) chosen by timing the
    plan's closest-point kernel per size class and over the sampled meshes (thg_autotune_plan)."""
    ctx = ctx or default_context()
    nv, nt, nq = _i64(nv), _i64(nt), _i64(nq)
    leaf = np.zeros(len(nt), np.int32)
    order = np.zeros(len(nt), np.int32)
    launch = np.zeros(2, np.int32)
    capi.check(capi.lib().thg_autotune_plan(ctx.h, len(nt), nv.ctypes.data_as(C.c_void_p), nt.ctypes.data_as(C.c_void_p),
                                            nq.ctypes.data_as(C.c_void_p), C.c_void_p(verts.data_ptr()),
                                            C.c_void_p(tris.data_ptr()), C.c_void_p(queries.data_ptr()),
                                            leaf.ctypes.data_as(C.c_void_p), order.ctypes.data_as(C.c_void_p),
                                            launch.ctypes.data_as(C.c_void_p)))
    return leaf, order, (int(launch[0]), int(launch[1]))
