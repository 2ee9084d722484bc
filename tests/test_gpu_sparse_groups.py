"""GPU parity of the sparse-batch kernels that share one query among G lanes (G = 2..64).

A batch with fewer queries than about one per 8 leaves takes the per-query kernels; below 2^17
queries each query gets G = 2^lg lanes (n * 2G > 2^18 fixes lg), and the lanes split the tree. The
cases use one 500k-triangle closed surface at leaf size 1 (about 500k leaves) with queries uniform
in its inflated box, so most queries lie inside it, far from the nearest triangle, and meet a large
part of the tree. A 100k batch takes the dense (wave-cooperative) kernel; subsets of the same
queries of 100 ... 50,000 take G = 64 ... 2. Bar: every row of every subset bit-identical to the
dense result (indices, squared distances, closest points), and the dense result and the smallest
subset bit-identical to the oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# subset size -> log2 lanes per query chosen for it
SUBSETS = {100: 6, 3000: 5, 6000: 4, 12000: 3, 24000: 2, 50000: 1}


@pytest.fixture(scope="module")
def sphere():
    import thea_amd
    from thea_amd import workloads as W
    g = np.random.default_rng(41)
    v, t = W.uv_sphere(354, 708)
    v = (v * (1 + 0.02 * g.normal(size=len(v)))[:, None]).astype(np.float32)
    t = t.astype(np.int32)
    lo, hi = W.inflated_box(v, 1.2)
    q = g.uniform(lo, hi, size=(100_000, 3)).astype(np.float32)
    b = thea_amd.MeshBvh(thea_amd.Mesh(v, t), leaf_size=1)
    perm = g.permutation(len(q))
    print(f"case sphere: {len(t)} tris, bvh {b.info}", flush=True)
    assert len(q) >= b.info["leaves"] / 8 > max(SUBSETS), "sizes no longer select the intended kernels"
    return v, t, q, b, perm


def _same(a, b_, label):
    assert np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b_).view(np.uint32)), label


def test_closest_groups_equal_dense(sphere, oracle):
    v, t, q, b, perm = sphere
    idx, dist, cp, d2 = b.closest(q, want_d2=True)
    for n, lg in SUBSETS.items():
        sel = perm[:n]
        print(f"case closest subset n={n} (G={1 << lg})", flush=True)
        i2, dist2, cp2, d22 = b.closest(q[sel], want_d2=True)
        assert np.array_equal(i2, idx[sel]), f"n={n}: {np.count_nonzero(i2 != idx[sel])} index mismatches"
        _same(d22, d2[sel], f"n={n}: d2")
        _same(cp2, cp[sel], f"n={n}: closest points")
        _same(dist2, dist[sel], f"n={n}: dist")
    sel = perm[:100]
    kd = oracle.KdTree(v, t)
    ridx, rcp, rd2, _ = kd.closest(q[sel], mode=0)
    assert np.array_equal(idx[sel], ridx)
    _same(d2[sel], rd2, "oracle d2")
    _same(cp[sel], rcp, "oracle closest points")


def test_closest_groups_with_bound(sphere):
    v, t, q, b, perm = sphere
    bound = 0.05  # most queries inside the sphere find nothing within it
    idx, _, cp, d2 = b.closest(q, dist_bound=bound, want_d2=True)
    assert 0 < np.count_nonzero(idx >= 0) < len(q)
    for n in (100, 6000, 50000):
        sel = perm[:n]
        print(f"case closest bound subset n={n}", flush=True)
        i2, _, cp2, d22 = b.closest(q[sel], dist_bound=bound, want_d2=True)
        assert np.array_equal(i2, idx[sel])
        hit = i2 >= 0
        _same(d22[hit], d2[sel][hit], f"n={n}: d2")
        _same(cp2[hit], cp[sel][hit], f"n={n}: cp")


@pytest.mark.parametrize("k", [3, 8, 16])
def test_knn_groups_equal_dense(sphere, oracle, k):
    v, t, q, b, perm = sphere
    idx, dist, d2 = b.kClosestElements(q, k, want_d2=True)
    for n, lg in SUBSETS.items():
        sel = perm[:n]
        print(f"case knn k={k} subset n={n} (G={1 << lg})", flush=True)
        i2, dist2, d22 = b.kClosestElements(q[sel], k, want_d2=True)
        assert np.array_equal(i2, idx[sel]), f"n={n}: {np.count_nonzero(i2 != idx[sel])} index mismatches"
        _same(d22, d2[sel], f"n={n}: d2")
        _same(dist2, dist[sel], f"n={n}: dist")
    sel = perm[:100]
    ridx, rd2, _ = oracle.KdTree(v, t).knn(q[sel], k, mode=0)
    assert np.array_equal(idx[sel], ridx)
    _same(d2[sel], rd2, "oracle d2")


def test_points_groups_equal_dense(sphere, oracle):
    import thea_amd
    v, _, _, _, _ = sphere
    g = np.random.default_rng(43)
    pts = v[g.permutation(len(v))[:240_000]]
    pts = np.concatenate([pts, pts[:500]])  # exact duplicates: ties on equal distance
    pb = thea_amd.PointBvh(pts, leaf_size=1)
    lo, hi = pts.min(0) * 1.2, pts.max(0) * 1.2
    q = g.uniform(lo, hi, size=(60_000, 3)).astype(np.float32)
    q[:300] = pts[:300]
    assert len(q) >= pb.info["leaves"] / 8 > 6000
    idx, _, cp, d2 = pb.closest(q, want_d2=True)
    kidx, _, kd2 = pb.kClosestElements(q, 6, want_d2=True)
    for n in (100, 3000, 6000):
        sel = np.arange(n)
        print(f"case points subset n={n}", flush=True)
        i2, _, cp2, d22 = pb.closest(q[sel], want_d2=True)
        assert np.array_equal(i2, idx[sel])
        _same(d22, d2[sel], "points d2")
        _same(cp2, cp[sel], "points cp")
        k2, _, kd22 = pb.kClosestElements(q[sel], 6, want_d2=True)
        assert np.array_equal(k2, kidx[sel])
        _same(kd22, kd2[sel], "points knn d2")
    ridx, rd2, _ = oracle.KdTree(points=pts).knn(q[:100], 6, mode=0)
    assert np.array_equal(kidx[:100], ridx)
    _same(kd2[:100], rd2, "oracle points knn d2")
