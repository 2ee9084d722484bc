"""thea_amd: MI355X-native BVH queries and mesh sampling with Thea's interface.

The compute path is libthea_gpu.so (HIP kernels for gfx950 behind the C ABI in
include/thea_gpu.h). This package is the Python host side: ctypes bindings (_capi), Thea-named
classes (bvh, icp), the request server (serve), batched plans (batch), and synthetic workload
generators (workloads).
"""
from .bvh import (Context, Mesh, MeshBvh, MeshSampler, PointBvh, PointBvhF64, default_context, furthestPointSampling,  # noqa: F401
                  version)
from .icp import Icp3  # noqa: F401
from ._capi import TheaGpuError  # noqa: F401
