"""GPU parity of the double-precision point search (thg_closest_point_f64, Thea's
BvhN<Vector3d, 3, double>::closestElement over points, the target structure of ShapeAlign's ICP)
against the oracle's double brute force, and of ICP on such a target against the double-precision
ICP restatement.

Bar: point indices and double squared distances bit-exact (lowest index on exact ties); closest
points are the target's double coordinates; distances sqrt(d^2) in double. ICP: identical iteration
counts, transforms within 1e-9 and errors within 1e-9 relative (numpy sums in another order).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _check(b, pts, q, bound=-1.0, label=""):
    from oracle import oracle as O
    idx, dist, cp, d2 = b.closest(q, dist_bound=bound, want_d2=True)
    ridx, rd2 = O.brute_closest_points_f64(pts, q, bound)
    bad = np.nonzero(idx != ridx)[0]
    assert len(bad) == 0, f"{label}: {len(bad)} index mismatches, first {bad[:5]}"
    hit = ridx >= 0
    assert np.array_equal(d2[hit], rd2[hit]), f"{label}: d2 differs"
    assert np.array_equal(cp[hit], pts[ridx[hit]]), f"{label}: closest points differ"
    assert np.array_equal(dist[hit], np.sqrt(rd2[hit]))
    assert np.all(idx[~hit] == -1) and np.all(dist[~hit] == -1)
    return idx


def test_closest64_random_and_bounded():
    import thea_amd
    g = np.random.default_rng(31)
    pts = g.normal(size=(60_000, 3))
    q = g.uniform(-3, 3, size=(150_000, 3))
    print("case f64 points 60k x 150k (sorted wave walk) and 3k (unsorted)", flush=True)
    for leaf in (1, 4, 16):
        b = thea_amd.PointBvhF64(pts, leaf_size=leaf)
        _check(b, pts, q, label=f"leaf {leaf}")
        _check(b, pts, q[:3000], label=f"leaf {leaf} small")
        _check(b, pts, q[:20000], bound=0.05, label=f"leaf {leaf} bounded")


def test_closest64_double_detail_and_ties():
    """Points that coincide in float but not in double: a float search cannot tell them apart, the
    double search must. Exact duplicates resolve to the lowest index."""
    import thea_amd
    g = np.random.default_rng(32)
    base = g.uniform(-1, 1, size=(5000, 3)).astype(np.float32).astype(np.float64)  # float-representable
    eps = np.spacing(np.abs(base).astype(np.float32)).astype(np.float64) * 0.25  # a quarter ulp per coordinate
    pts = np.concatenate([base, base + eps, base - eps, base[:100]])  # sub-float offsets and duplicates
    q = np.concatenate([base + eps * 0.9, base - eps * 0.6, base[:100],
                        1000.0 + g.uniform(0, 1e-3, size=(500, 3))])  # far queries
    print("case f64 points below float resolution", flush=True)
    b = thea_amd.PointBvhF64(pts, leaf_size=4)
    idx = _check(b, pts, q, label="detail")
    # the queries nearest the +eps copies find them (index 5000..9999), which the float search cannot
    assert np.array_equal(idx[:5000], np.arange(5000, 10000))
    assert np.array_equal(idx[5000:10000], np.arange(10000, 15000))
    assert np.array_equal(idx[10000:10100], np.arange(100))  # exact duplicates: lowest index
    # in float the three copies are one point (and the query rounds onto it): the lowest index wins
    fidx = thea_amd.PointBvh(pts.astype(np.float32)).closestElement(q[:5000].astype(np.float32))
    assert np.array_equal(fidx, np.arange(5000))


def test_closest64_device_and_empty():
    import torch
    import thea_amd
    g = np.random.default_rng(33)
    pts = g.normal(size=(20_000, 3))
    q = g.normal(size=(50_000, 3))
    b = thea_amd.PointBvhF64(torch.from_numpy(pts).cuda())
    hidx, hdist, hcp, _ = b.closest(q)
    didx, ddist, dcp, _ = b.closest(torch.from_numpy(q).cuda())
    torch.cuda.synchronize()
    assert np.array_equal(hidx, didx.cpu().numpy()) and np.array_equal(hdist, ddist.cpu().numpy())
    assert np.array_equal(hcp, dcp.cpu().numpy())
    e = thea_amd.PointBvhF64(np.zeros((0, 3)))
    idx, dist, _, _ = e.closest(q[:10])
    assert np.all(idx == -1) and np.all(dist == -1)


@pytest.mark.parametrize("kabsch", [False, True])
def test_icp_on_double_target_matches_oracle(oracle, kabsch):
    """ShapeAlign's configuration: Icp3<double> over a BvhN<Vector3d> of the target points."""
    import thea_amd
    from oracle import icp as I
    print(f"case icp f64 kabsch={kabsch}", flush=True)
    g = np.random.default_rng(34)
    u = g.normal(size=(30000, 3))
    to = u / np.linalg.norm(u, axis=1)[:, None] * np.array([1.0, 0.6, 0.3])  # double ellipsoid samples
    c, s_ = np.cos(0.08), np.sin(0.08)
    R = np.array([[c, -s_, 0], [s_, c, 0], [0, 0, 1.0]])
    src = to[::7] @ R.T + np.array([0.02, -0.01, 0.03])
    bvh = thea_amd.PointBvhF64(to)
    icp = thea_amd.Icp3(kabsch=kabsch)
    tr, err, its = icp.alignBatch(src, bvh)
    ref = I.icp_align(lambda x: to[oracle.brute_closest_points_f64(to, x)[0]], src, kabsch=kabsch, f64=True)
    rtr, rerr, rits = ref
    assert its[0] == rits, f"iterations {its[0]} vs {rits}"
    assert np.allclose(tr[0], rtr, rtol=0, atol=1e-9), np.abs(tr[0] - rtr).max()
    assert abs(err[0] - rerr) <= 1e-9 * max(abs(rerr), 1e-30)
