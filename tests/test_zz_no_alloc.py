"""thg_closest_point_noalloc (the entry point that allocates nothing) against thg_closest_point.

Bar: indices, closest points, distances and squared distances bit-identical to the allocating entry
point on the same inputs. Calls run back to back on one stream, each on a batch smaller than or
different from the one before, so a result that depended on what an earlier call left in the
shared sort memory would differ. Covers the sorted wave path at 8 and at 1 resident waves per SIMD
(batches above and below 256 CUs x 24 waves x 64 queries), a bounded search, an unsorted sparse
batch, a point BVH, and the size limit.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(L, fn, bvh, qd, n, bound, stream):
    import torch
    from thea_amd import _capi as capi
    dev = qd.device
    idx = torch.full((n,), -7, dtype=torch.int32, device=dev)
    cp = torch.full((n, 3), 7.0, dtype=torch.float32, device=dev)
    dist = torch.full((n,), 7.0, dtype=torch.float32, device=dev)
    d2 = torch.full((n,), 7.0, dtype=torch.float32, device=dev)
    P = lambda x: C.c_void_p(x.data_ptr())  # noqa: E731
    if fn == "noalloc":
        capi.check(L.thg_closest_point_noalloc(bvh.h, P(qd), n, bound, P(idx), P(cp), P(dist), P(d2), stream))
    else:
        capi.check(L.thg_closest_point(bvh.h, P(qd), n, bound, P(idx), P(cp), P(dist), P(d2), capi.THG_MEM_DEVICE,
                                       stream))
    return idx, cp, dist, d2


def _same(a, b, label):
    import torch
    torch.cuda.synchronize()
    for name, x, y in zip(("idx", "cp", "dist", "d2"), a, b):
        xa, ya = x.cpu().numpy(), y.cpu().numpy()
        assert np.array_equal(xa.view(np.uint32), ya.view(np.uint32)), f"{label}: {name} differs"
    if "bound -1" in label:
        assert int((a[0] < 0).sum()) == 0, f"{label}: unbounded query without a result"


def test_noalloc_equals_closest_point():
    import torch
    import thea_amd
    from thea_amd import _capi as capi
    from thea_amd import workloads as W
    L = capi.lib()
    thea_amd.default_context(0)
    v, t, q, _, _ = W.torus_c3(n_queries=1_700_000, n_rays=0)
    b = thea_amd.MeshBvh(thea_amd.Mesh(v, t), leaf_size=8)
    qd = torch.from_numpy(q).cuda()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    # (first query, count, bound): 8-wave sorted batch, then smaller and different ones
    cases = [(0, 1_000_000, -1.0), (1_000_000, 150_000, -1.0), (1_150_000, 300_000, 0.05), (1_450_000, 2_000, -1.0),
             (1_452_000, 4_096, -1.0)]
    for first, n, bound in cases:
        label = f"C3 torus leaf 8, queries [{first}, {first + n}), bound {bound}"
        print(f"case {label}", flush=True)
        sub = qd[first:first + n]
        got = _call(L, "noalloc", b, sub, n, bound, stream)
        want = _call(L, "alloc", b, sub, n, bound, stream)
        _same(got, want, label)


def test_noalloc_point_bvh_and_empty_batch():
    import torch
    import thea_amd
    from thea_amd import _capi as capi
    L = capi.lib()
    thea_amd.default_context(0)
    g = np.random.default_rng(5)
    pts = g.normal(size=(200_000, 3)).astype(np.float32)
    q = g.uniform(-2, 2, size=(120_000, 3)).astype(np.float32)
    b = thea_amd.PointBvh(pts, leaf_size=4)
    qd = torch.from_numpy(q).cuda()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for n in (120_000, 50_000):
        print(f"case point BVH 200k points x {n} queries", flush=True)
        got = _call(L, "noalloc", b, qd[:n], n, -1.0, stream)
        want = _call(L, "alloc", b, qd[:n], n, -1.0, stream)
        _same(got, want, f"points n={n}")
    assert L.thg_closest_point_noalloc(b.h, None, 0, -1.0, None, None, None, None, stream) == capi.THG_OK


def test_noalloc_rejects_batches_above_the_limit():
    """The limit is checked on the host before any device work."""
    import thea_amd
    from thea_amd import _capi as capi
    L = capi.lib()
    thea_amd.default_context(0)
    g = np.random.default_rng(6)
    b = thea_amd.PointBvh(g.normal(size=(1000, 3)).astype(np.float32))
    over = capi.THG_NOALLOC_MAX_QUERIES + 1
    rc = L.thg_closest_point_noalloc(b.h, C.c_void_p(16), over, -1.0, C.c_void_p(16), None, None, None, None)
    assert rc == capi.THG_E_INVALID
    assert b"THG_NOALLOC_MAX_QUERIES" in L.thg_last_error()
