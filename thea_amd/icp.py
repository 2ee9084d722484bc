"""Iterative closest point with Thea's interface (Algorithms/Icp3.hpp): Icp3(...).align(from, to_bvh)
returns the 3x4 transform (row-major [L | t]) and the alignment error; alignBatch runs several
alignments of the same points from different initial transforms in one call (thg_icp_align)."""
import ctypes as C

import numpy as np

from . import _capi as capi


class Icp3:
    """Icp3<double> (Icp3.hpp:34-66): negative arguments select Thea's defaults (threshold 1e-4,
    at least 3 and at most 10 iterations). kabsch=True replaces Thea's fit (rot = U V^T of the
    cross-covariance, the transpose of the least-squares rotation) with the least-squares rotation."""

    def __init__(self, fractional_error_threshold=-1.0, min_iterations=-1, max_iterations=-1, verbose=False, *,
                 kabsch=False):
        self.fractional_error_threshold = float(fractional_error_threshold)
        self.min_iterations = int(min_iterations)
        self.max_iterations = int(max_iterations)
        self.kabsch = bool(kabsch)
        self.up = None

    def setUpVector(self, up):
        up = np.asarray(up, np.float64).reshape(3)
        self.up = up / np.linalg.norm(up)

    def hasUpVector(self):
        return self.up is not None

    def getUpVector(self):
        return self.up

    def clearUpVector(self):
        self.up = None

    def _opts(self, wt, wr):
        o = capi.IcpOpts(fractional_error_threshold=self.fractional_error_threshold,
                         min_iterations=self.min_iterations, max_iterations=self.max_iterations,
                         has_up=1 if self.up is not None else 0, kabsch=1 if self.kabsch else 0)
        if self.up is not None:
            for a in range(3):
                o.up[a] = float(self.up[a])
        o.translation_weights = None if wt is None else wt.ctypes.data
        o.rotation_weights = None if wr is None else wr.ctypes.data
        return o

    def alignBatch(self, from_points, to, initial=None, translation_weights=None, rotation_weights=None):
        """(transforms (K, 3, 4), errors (K,), iterations (K,)) for the source points moved by each of
        the K initial (3, 4) transforms first (identity when None, K = 1); `to` is a MeshBvh or PointBvh
        (positions rounded to float for the search), or a PointBvhF64 (searched in double, as
        ShapeAlign's BvhN<Vector3d, 3, double> target, Tools/ShapeAlign/ShapeAlign.cpp:339)."""
        f = np.ascontiguousarray(from_points, np.float64).reshape(-1, 3)
        init = None if initial is None else np.ascontiguousarray(initial, np.float64).reshape(-1, 12)
        K = 1 if init is None else len(init)
        wt = None if translation_weights is None else np.ascontiguousarray(translation_weights, np.float64)
        wr = None if rotation_weights is None else np.ascontiguousarray(rotation_weights, np.float64)
        for w in (wt, wr):
            if w is not None and len(w) != len(f):
                raise ValueError("one weight per source point")
        tr = np.zeros((K, 12), np.float64)
        err = np.zeros(K, np.float64)
        its = np.zeros(K, np.int32)
        o = self._opts(wt, wr)
        capi.check(capi.lib().thg_icp_align(to.h, f.ctypes.data_as(C.c_void_p), len(f), K,
                                            None if init is None else init.ctypes.data_as(C.c_void_p), C.byref(o),
                                            capi.THG_MEM_HOST, None, tr.ctypes.data_as(C.c_void_p),
                                            err.ctypes.data_as(C.c_void_p), its.ctypes.data_as(C.c_void_p)))
        return tr.reshape(K, 3, 4), err, its

    def align(self, from_points, to, translation_weights=None, rotation_weights=None):
        """(transform (3, 4), error): Icp3::align(n, from, to, &error)."""
        tr, err, _ = self.alignBatch(from_points, to, None, translation_weights, rotation_weights)
        return tr[0], float(err[0])
