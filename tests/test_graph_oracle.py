"""Pins the oracle's restatement of Thea's proximity graph (PointSet3::updateGraph) and fastSqrt.

fastSqrt: the published error bound of the Quake III reciprocal square root with one Newton step
(max relative error 0.175228%, Lomont 2003) over a sweep of floats, and its known value at 1.0
(0.998307...). Graph: the grid-based restatement against a numpy brute-force statement of the same
definition (ball of radius sqrt(max_degree) * RMS separation, doubling up to 3 times until min_degree
neighbours; normals not opposed; max_degree smallest (fastSqrt(d^2), index))."""
import numpy as np

from oracle import oracle as O


def test_fast_sqrt_bound_and_known_value():
    g = np.random.default_rng(1)
    x = np.concatenate([np.logspace(-30, 30, 20001), g.uniform(0, 4, 20000), [1.0, 2.0, 4.0, 0.25]]).astype(np.float32)
    fs = O.fast_sqrt(x).astype(np.float64)
    ref = np.sqrt(x.astype(np.float64))
    rel = np.abs(fs / ref - 1)
    assert rel.max() < 0.00176 and rel.max() > 0.0015  # one Newton step: the bound is reached
    assert abs(1.0 / O.fast_sqrt(np.array([1.0], np.float32))[0] - 0.998307) < 1e-6


def _brute_graph(p, nrm, md, picks):
    p = p.astype(np.float32)
    n = len(p)
    acc = np.float32(0)
    for a in picks:
        d = ((p - p[a]) ** 2).astype(np.float32)
        d2 = d[:, 0] + (d[:, 1] + d[:, 2])
        d2[a] = np.inf
        acc = np.float32(acc + d2.min())
    avg = np.float32(np.sqrt(np.float32(acc / np.float32(len(picks)))))
    r0 = np.float32(np.sqrt(np.float32(md)) * avg)
    mind = min(md, max(4, int(0.25 * md)))
    idx = -np.ones((n, md), np.int32)
    sep = -np.ones((n, md), np.float32)
    deg = np.zeros(n, np.int32)
    for i in range(n):
        diff = (p - p[i]).astype(np.float32)
        sq = diff * diff
        d2 = sq[:, 0] + (sq[:, 1] + sq[:, 2])
        ok = np.arange(n) != i
        if nrm is not None:
            pr = nrm * nrm[i]
            ok &= (pr[:, 0] + (pr[:, 1] + pr[:, 2])) >= np.float32(-0.5)
        radius = r0
        for _ in range(3):
            cand = np.nonzero(ok & (d2 <= np.float32(radius * radius)))[0]
            keys = O.fast_sqrt(d2[cand])
            order = np.lexsort((cand, keys))[:md]
            if len(cand) >= mind:
                break
            radius = np.float32(radius * np.float32(2))
        k = len(order)
        idx[i, :k], sep[i, :k], deg[i] = cand[order], keys[order], k
    return idx, sep, deg, avg


def test_graph_oracle_matches_brute_force():
    g = np.random.default_rng(2)
    u = g.normal(size=(3000, 3))
    p = (u / np.linalg.norm(u, axis=1)[:, None] * np.array([1.0, 0.7, 0.4])).astype(np.float32)
    p[2000:2050] = p[0]  # duplicates: zero separations, ties by index
    nrm = (u / np.linalg.norm(u, axis=1)[:, None]).astype(np.float32)
    picks = O.graph_picks(len(p), 100, 5)
    for md, normals in ((8, None), (5, nrm), (16, nrm)):
        idx, sep, deg, avg = O.proximity_graph(p, normals, md, picks)
        bi, bs, bd, bavg = _brute_graph(p, normals, md, picks)
        assert np.array_equal(idx, bi) and np.array_equal(sep.view(np.int32), bs.view(np.int32))
        assert np.array_equal(deg, bd) and np.float32(avg[0]) == bavg
        total = np.float32(0)
        for i in range(len(p)):
            for j in range(bd[i]):
                total = np.float32(total + bs[i, j])
        assert np.float32(avg[1]) == np.float32(total / np.float32(bd.sum()))


def _scipy_extract(idx, sep, deg, S):
    """extractOriginalAdjacencies by an independent route: all-source shortest paths from
    scipy.sparse.csgraph.dijkstra (double sums of the float weights) over the directed graph, then
    the max_degree smallest (distance, index) other original samples, reported as
    (float(distance), index) in ascending order."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import dijkstra
    n, md = idx.shape
    rows = np.repeat(np.arange(n), deg)
    cols = np.concatenate([idx[i, :deg[i]] for i in range(n)]).astype(np.int64)
    w = np.concatenate([sep[i, :deg[i]] for i in range(n)]).astype(np.float64)
    # zero weights must survive the sparse matrix: shift by a tiny tag and subtract it back is not
    # exact, so the test graphs have positive weights only
    assert np.all(w > 0)
    G = csr_matrix((w, (rows, cols)), shape=(n, n))
    D = dijkstra(G, directed=True, indices=np.arange(S))
    oi = -np.ones((S, md), np.int32)
    os_ = -np.ones((S, md), np.float32)
    od = np.zeros(S, np.int32)
    for s in range(S):
        d = D[s, :S].copy()
        d[s] = np.inf
        order = np.lexsort((np.arange(S), d))
        keep = [j for j in order[:md] if np.isfinite(d[j])]
        pairs = sorted((np.float32(d[j]), j) for j in keep)
        od[s] = len(pairs)
        for k, (fd, j) in enumerate(pairs):
            oi[s, k], os_[s, k] = j, fd
    return oi, os_, od


def test_extract_original_against_scipy_dijkstra():
    """The oracle's per-sample Dijkstra with early stop equals the k smallest all-pairs shortest
    paths among the original samples, on a random proximity graph of 3,000 points, 300 originals."""
    g = np.random.default_rng(21)
    u = g.normal(size=(3000, 3))
    p = (u / np.linalg.norm(u, axis=1)[:, None]).astype(np.float32)
    idx, sep, deg, _ = O.proximity_graph(p, None, 8, O.graph_picks(len(p), 100, 3))
    S = 300
    got = O.extract_original(idx, sep, deg, S)
    ref = _scipy_extract(idx, sep, deg, S)
    assert np.array_equal(got[2], ref[2])
    assert np.array_equal(got[0], ref[0])
    assert np.array_equal(got[1].view(np.int32), ref[1].view(np.int32))


def test_extract_original_path_graph_closed_form():
    """A directed path 0 -> 1 -> ... -> n-1 with unit weights and originals every 3rd vertex: sample
    3a reaches originals 3(a+1), 3(a+2), ... at distances 3, 6, ...; the last ones see none."""
    n, md = 30, 4
    idx = -np.ones((n, md), np.int32)
    sep = -np.ones((n, md), np.float32)
    deg = np.zeros(n, np.int32)
    for i in range(n - 1):
        idx[i, 0], sep[i, 0], deg[i] = i + 1, 1.0, 1
    # originals are 0..9 by index: relabel so that original a sits at path position 3a
    perm = np.empty(n, np.int64)  # new index of path position q
    orig_pos = np.arange(0, n, 3)
    others = np.setdiff1d(np.arange(n), orig_pos)
    perm[orig_pos] = np.arange(len(orig_pos))
    perm[others] = len(orig_pos) + np.arange(len(others))
    ridx = -np.ones_like(idx)
    rsep = -np.ones_like(sep)
    rdeg = np.zeros_like(deg)
    for q in range(n):
        v = perm[q]
        rdeg[v] = deg[q]
        if deg[q]:
            ridx[v, 0], rsep[v, 0] = perm[idx[q, 0]], sep[q, 0]
    S = len(orig_pos)
    oi, os_, od = O.extract_original(ridx, rsep, rdeg, S)
    for a in range(S):
        want = [b for b in range(a + 1, S)][:md]
        assert od[a] == len(want)
        assert list(oi[a, :len(want)]) == want
        assert list(os_[a, :len(want)]) == [3.0 * (b - a) for b in want]
