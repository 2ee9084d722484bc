"""GPU tests of thg_proximity_graph (Thea's PointSet3::updateGraph) against the oracle: neighbour
indices, separations (Thea's fastSqrt) and degrees bit for bit, and both separation averages."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cmp(got, ref, label):
    idx, sep, deg, avg = got
    ridx, rsep, rdeg, ravg = ref
    bad = np.nonzero((idx != ridx).any(1))[0]
    assert len(bad) == 0, f"{label}: {len(bad)} samples differ, first {bad[:3]}: {idx[bad[:1]]} vs {ridx[bad[:1]]}"
    assert np.array_equal(sep.view(np.int32), rsep.view(np.int32)), label
    assert np.array_equal(deg, rdeg), label
    assert np.float32(avg[0]) == np.float32(ravg[0]) and np.float32(avg[1]) == np.float32(ravg[1]), (label, avg, ravg)


@pytest.mark.parametrize("md,with_normals,leaf", [(8, False, 4), (8, True, 8), (4, True, 1), (16, False, 12),
                                                  (32, True, 4)])
def test_graph_matches_oracle(oracle, md, with_normals, leaf):
    import thea_amd
    print(f"case graph md={md} normals={with_normals} leaf={leaf}", flush=True)
    g = np.random.default_rng(md + leaf)
    u = g.normal(size=(60000, 3))
    nrm = (u / np.linalg.norm(u, axis=1)[:, None]).astype(np.float32)
    p = (nrm * np.array([1.0, 0.7, 0.4], np.float32)).astype(np.float32)
    p[50000:50100] = p[7]  # duplicates
    b = thea_amd.PointBvh(p, leaf_size=leaf)
    got = b.proximityGraph(normals=nrm if with_normals else None, max_degree=md, seed=9)
    ref = oracle.proximity_graph(p, nrm if with_normals else None, md, oracle.graph_picks(len(p), 100, 9))
    _cmp(got, ref, f"md={md}")
    assert got[2].min() >= 4 or md < 4


def test_graph_sparse_regions_expand_radius(oracle):
    """Clusters far apart: samples in small clusters find fewer than min_degree neighbours at the
    first radius and take the doubling steps."""
    import thea_amd
    print("case graph clusters", flush=True)
    g = np.random.default_rng(3)
    dense = g.normal(size=(20000, 3)).astype(np.float32) * 0.1
    sparse = (g.normal(size=(300, 3)) * 0.3 + 5).astype(np.float32)
    p = np.concatenate([dense, sparse])
    b = thea_amd.PointBvh(p)
    got = b.proximityGraph(max_degree=8, seed=1)
    ref = oracle.proximity_graph(p, None, 8, oracle.graph_picks(len(p), 100, 1))
    _cmp(got, ref, "clusters")


def test_graph_c2_samples(oracle):
    """The 16-neighbour graph of C2's 1M area samples (SampleGraph on the k-NN workload's points)."""
    import thea_amd
    from thea_amd import workloads as W
    print("case graph C2 samples", flush=True)
    v, t = W.ellipsoids_c2()
    pos, nrm, _ = thea_amd.MeshSampler(thea_amd.Mesh(v, t)).sampleEvenlyByArea(1_000_000, seed=2,
                                                                               stream_id=W.STREAM_SAMPLES)
    b = thea_amd.PointBvh(pos, leaf_size=12)
    got = b.proximityGraph(normals=nrm, max_degree=16, seed=4)
    ref = oracle.proximity_graph(pos, nrm, 16, oracle.graph_picks(len(pos), 100, 4))
    _cmp(got, ref, "C2")
