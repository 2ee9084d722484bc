"""Pins the oracle's furthest-point sampling (or_fps, a restatement of Thea's FurthestPointSampling
with a per-pick Dijkstra): closed forms on path graphs (connected and in two components) and, on
proximity graphs of sampled surfaces, an independent selection loop built on
scipy.sparse.csgraph.dijkstra."""
import numpy as np
import pytest


def path_graph(n, offset=0, total=None):
    total = total or n
    idx = -np.ones((total, 2), np.int32)
    sep = -np.ones((total, 2), np.float32)
    for i in range(n):
        nb = [offset + j for j in (i - 1, i + 1) if 0 <= j < n]
        idx[offset + i, :len(nb)] = nb
        sep[offset + i, :len(nb)] = 1.0
    return idx, sep


def test_path_graph_closed_form(oracle):
    idx, sep = path_graph(9)
    # 0 first; then the far end; the middle; then ties at equal distance resolve to the higher index
    assert list(oracle.fps(idx, sep, 9)) == [0, 8, 4, 6, 2, 7, 5, 3, 1]
    assert list(oracle.fps(idx, sep, 3)) == [0, 8, 4]


def test_two_components_unreachable_first(oracle):
    a_idx, a_sep = path_graph(5, 0, 10)
    b_idx, b_sep = path_graph(5, 5, 10)
    idx = np.where(a_idx >= 0, a_idx, b_idx)
    sep = np.where(a_idx >= 0, a_sep, b_sep)
    # after 0, the other component is unreachable: its last vertex (9) is next
    assert list(oracle.fps(idx, sep, 10)) == [0, 9, 5, 4, 7, 2, 8, 6, 3, 1]


def test_stops_when_everything_is_picked(oracle):
    idx, sep = path_graph(4)
    assert list(oracle.fps(idx, sep, 4)) == [0, 3, 2, 1]
    # every vertex at distance 0 from the picks: no fifth pick is possible on a 4-vertex graph
    dup_idx = np.array([[1, -1], [0, -1]], np.int32)
    dup_sep = np.array([[0.0, -1], [0.0, -1]], np.float32)  # coincident points
    assert list(oracle.fps(dup_idx, dup_sep, 2)) == [0]


def scipy_fps(idx, sep, k):
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import dijkstra
    n, md = idx.shape
    rows = np.repeat(np.arange(n), md)
    keep = idx.reshape(-1) >= 0
    g = csr_matrix((sep.reshape(-1)[keep].astype(np.float64), (rows[keep], idx.reshape(-1)[keep])), shape=(n, n))
    picks = [0]
    while len(picks) < k:
        d = dijkstra(g, directed=True, indices=picks, min_only=True)
        inf = np.nonzero(np.isinf(d))[0]
        if len(inf):
            nxt = int(inf.max())
        else:
            m = d.max()
            nxt = int(np.nonzero(d == m)[0].max())
        if d[nxt] == 0:
            break
        picks.append(nxt)
    return np.array(picks, np.int64)


@pytest.mark.parametrize("components", [1, 3])
def test_matches_scipy_dijkstra(oracle, components):
    g = np.random.default_rng(components)
    pts = []
    for c in range(components):
        u = g.normal(size=(1500, 3))
        pts.append(u / np.linalg.norm(u, axis=1)[:, None] * [1.0, 0.6, 0.8] + [4.0 * c, 0, 0])
    p = np.concatenate(pts).astype(np.float32)
    p = p[g.permutation(len(p))]
    idx, sep, _, _ = oracle.proximity_graph(p, None, 8, oracle.graph_picks(len(p), 100, 3))
    ref = scipy_fps(idx, sep, 400)
    got = oracle.fps(idx, sep, 400)
    assert np.array_equal(got, ref)
