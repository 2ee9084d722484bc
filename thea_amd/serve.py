"""Python binding of the request server (thg_server_* in include/thea_gpu.h).

Meshes stay resident on the GPU under a memory budget; closest-point, k-NN and ray requests are
submitted with torch device tensors and served with continuous batching (see serve.hip).
"""
import ctypes as C

import numpy as np

from . import _capi as capi
from .bvh import default_context


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


class Server:
    def __init__(self, ctx=None, memory_budget_bytes=0, num_streams=4, max_batch_requests=256,
                 max_batch_queries=16 << 20, coalesce=True):
        self.ctx = ctx or default_context()
        o = capi.ServerOpts(memory_budget_bytes=memory_budget_bytes, num_streams=num_streams,
                            max_batch_requests=max_batch_requests, max_batch_queries=max_batch_queries,
                            coalesce=1 if coalesce else -1)
        h = C.c_void_p()
        capi.check(capi.lib().thg_server_create(self.ctx.h, C.byref(o), C.byref(h)))
        self.h = h
        self._keep = {}

    def close(self):
        if getattr(self, "h", None):
            capi.lib().thg_server_destroy(self.h)
            self.h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_mesh(self, vertices, triangles, leaf_size=4):
        v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
        t = np.ascontiguousarray(triangles, np.int32).reshape(-1, 3)
        mid = C.c_int32()
        capi.check(capi.lib().thg_server_add_mesh(self.h, v.ctypes.data_as(C.c_void_p), len(v),
                                                  t.ctypes.data_as(C.c_void_p), len(t), capi.THG_MEM_HOST,
                                                  leaf_size, C.byref(mid)))
        return mid.value

    def submit(self, kind, mesh, q, dirs=None, k=0, bound=-1.0, exclude_self=False, outputs=None):
        """Queue a request on torch device tensors; returns (request id, outputs dict).
        kind: "closest" | "knn" | "ray". Outputs are allocated here unless given."""
        import torch
        n = q.shape[0]
        dev = q.device
        if outputs is None:
            if kind == "closest":
                outputs = {"idx": torch.empty(n, dtype=torch.int32, device=dev),
                           "dist": torch.empty(n, dtype=torch.float32, device=dev),
                           "cp": torch.empty((n, 3), dtype=torch.float32, device=dev)}
            elif kind == "knn":
                outputs = {"idx": torch.empty((n, k), dtype=torch.int32, device=dev),
                           "dist": torch.empty((n, k), dtype=torch.float32, device=dev), "cp": None}
            else:
                outputs = {"idx": torch.empty(n, dtype=torch.int32, device=dev),
                           "dist": torch.empty(n, dtype=torch.float32, device=dev),
                           "cp": torch.empty((n, 3), dtype=torch.float32, device=dev)}
        typ = {"closest": capi.THG_REQ_CLOSEST, "knn": capi.THG_REQ_KNN, "ray": capi.THG_REQ_RAY}[kind]
        r = capi.Request(type=typ, mesh=mesh, n=n, k=k, exclude_self=int(bool(exclude_self)), bound=float(bound),
                         q=q.data_ptr(), dirs=dirs.data_ptr() if dirs is not None else None,
                         idx=outputs["idx"].data_ptr(),
                         dist=outputs["dist"].data_ptr() if outputs.get("dist") is not None else None,
                         cp=outputs["cp"].data_ptr() if outputs.get("cp") is not None else None)
        rid = C.c_int64()
        capi.check(capi.lib().thg_server_submit(self.h, C.byref(r), C.byref(rid)))
        self._keep[rid.value] = (q, dirs, outputs)
        return rid.value, outputs

    def wait(self, rid, timeout_us=-1):
        """Status (THG_OK, an error code, or THG_PENDING) and latency in microseconds."""
        lat = C.c_double()
        rc = capi.lib().thg_server_wait(self.h, rid, int(timeout_us), C.byref(lat))
        if rc != capi.THG_PENDING:
            self._keep.pop(rid, None)
        return rc, lat.value

    def cancel(self, rid):
        return capi.lib().thg_server_cancel(self.h, rid)

    def stats(self, reset=False):
        out = (C.c_double * 10)()
        capi.check(capi.lib().thg_server_stats(self.h, out, 1 if reset else 0))
        keys = ["completed", "launches", "requests_per_launch", "p50_us", "p99_us", "resident_bytes",
                "peak_inflight_bytes", "rejected", "queries", "mean_us"]
        return dict(zip(keys, list(out)))
