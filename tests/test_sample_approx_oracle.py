"""Pins the oracle's APPROXIMATE-mode sampler (MeshSampler::sampleEvenlyByArea, CountMode::
APPROXIMATE): each triangle receives floor(frac) or ceil(frac) samples with frac = n * area / total,
the extra sample appears with probability equal to frac's fractional part, the expected total is n,
samples come in triangle order and lie in their triangles."""
import numpy as np


def test_counts_and_points(oracle):
    from thea_amd import workloads as W
    v, t = W.torus(64, 32, seed=2)
    n = 20000
    pos, nrm, tri = oracle.sample_area_approx(v, t, n, seed=3, stream=1)
    # triangle areas with Thea's formula come from the exact-mode CDF
    cdf = oracle.sample_area(v, t, 1, 0, 0, want_cdf=True)[3]
    a = np.diff(np.concatenate([[0.0], cdf]))
    frac = n * a / cdf[-1]
    cnt = np.bincount(tri, minlength=len(t))
    lo, hi = np.floor(frac), np.ceil(frac)
    assert np.all((cnt == lo) | (cnt == hi))
    assert np.all(np.diff(tri) >= 0)  # triangle order
    # the extra samples: as many as the fractional parts add up to, within 5 standard deviations
    r = frac - lo
    extra = (cnt - lo).sum()
    assert abs(extra - r.sum()) < 5 * np.sqrt((r * (1 - r)).sum())
    assert abs(len(pos) - n) < 5 * np.sqrt((r * (1 - r)).sum()) + 1
    # points lie in their triangles: barycentric coordinates in [0, 1]
    p0, p1, p2 = v[t[tri, 0]], v[t[tri, 1]], v[t[tri, 2]]
    e1, e2, d = (p1 - p0).astype(np.float64), (p2 - p0).astype(np.float64), (pos - p0).astype(np.float64)
    g = np.stack([np.einsum("ij,ij->i", e1, e1), np.einsum("ij,ij->i", e1, e2), np.einsum("ij,ij->i", e2, e2)], 1)
    rhs = np.stack([np.einsum("ij,ij->i", d, e1), np.einsum("ij,ij->i", d, e2)], 1)
    det = g[:, 0] * g[:, 2] - g[:, 1] ** 2
    s = (g[:, 2] * rhs[:, 0] - g[:, 1] * rhs[:, 1]) / det
    u = (g[:, 0] * rhs[:, 1] - g[:, 1] * rhs[:, 0]) / det
    assert s.min() > -1e-4 and u.min() > -1e-4 and (s + u).max() < 1 + 1e-4


def test_integral_fractions_draw_nothing(oracle):
    # two triangles of equal area and n even: frac is an integer for both, no draw is made
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [5, 0, 0], [6, 0, 0], [5, 1, 0]], np.float32)
    t = np.array([[0, 1, 2], [3, 4, 5]], np.int32)
    for seed in range(5):
        pos, _, tri = oracle.sample_area_approx(v, t, 10, seed=seed)
        assert list(np.bincount(tri)) == [5, 5]
