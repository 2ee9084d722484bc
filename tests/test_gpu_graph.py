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


def _oversampled_ref(oracle, p, nrm, md, S, seed):
    gi, gs, gd, gavg = oracle.proximity_graph(p, nrm, md, oracle.graph_picks(len(p), 100, seed))
    ridx, rsep, rdeg = oracle.extract_original(gi, gs, gd, S)
    tot = np.float32(0)
    for i in range(S):  # updateAverageSeparation over the original samples, in order
        for j in range(rdeg[i]):
            tot = np.float32(tot + rsep[i, j])
    e = int(rdeg.sum())
    return ridx, rsep, rdeg, (gavg[0], float(np.float32(tot / np.float32(e))) if e else 0.0)


@pytest.mark.parametrize("S,md,with_normals", [(2000, 8, True), (5000, 8, False), (800, 16, True)])
def test_graph_oversampled_matches_oracle(oracle, S, md, with_normals):
    """SampleGraph below min_samples: S samples of a mesh plus dense area samples up to 30,000 points
    (SampleGraph.cpp:296-334); each sample keeps its max_degree nearest other samples by shortest
    path through the dense graph (PointSet3.cpp:158-178)."""
    import thea_amd
    from thea_amd import workloads as W
    print(f"case oversampled graph S={S} md={md} normals={with_normals}", flush=True)
    v, t = W.torus(96, 48, R=1.0, r=0.3, amp=0.03, seed=7)
    sampler = thea_amd.MeshSampler(thea_amd.Mesh(v, t))
    sp, sn, _ = sampler.sampleEvenlyByArea(S, seed=5, stream_id=1)
    dp, dn, _ = sampler.sampleEvenlyByArea(30000 - S, seed=5, stream_id=2)
    p = np.concatenate([sp, dp])
    nrm = np.concatenate([sn, dn]) if with_normals else None
    b = thea_amd.PointBvh(p, leaf_size=8)
    got = b.proximityGraph(normals=nrm, max_degree=md, seed=11, num_original=S)
    ref = _oversampled_ref(oracle, p, nrm, md, S, 11)
    assert got[0].shape == (S, md)
    _cmp(got, ref, f"oversampled S={S}")
    assert got[2].min() >= 1


def test_graph_oversampled_sparse_originals_use_the_large_tables(oracle):
    """40 samples among 40,000 points: every search settles thousands of vertices, more than the
    per-wave LDS tables hold, and is rerun with tables for the whole graph."""
    import thea_amd
    print("case oversampled graph, 40 originals", flush=True)
    g = np.random.default_rng(8)
    u = g.normal(size=(40000, 3))
    p = (u / np.linalg.norm(u, axis=1)[:, None]).astype(np.float32)
    S = 40
    b = thea_amd.PointBvh(p)
    got = b.proximityGraph(max_degree=8, seed=2, num_original=S)
    ref = _oversampled_ref(oracle, p, None, 8, S, 2)
    _cmp(got, ref, "sparse originals")


@pytest.mark.parametrize("S", [None, 1500])
def test_graph_reachability_matches_oracle(oracle, S):
    """updateGraph(surface) (SampleGraph -r): the 64 ellipsoids of the k-NN workload, whose surfaces
    cut each other, so candidates across another ellipsoid's surface are blocked; with and without
    oversampling. The graph must differ from the one without the surface."""
    import thea_amd
    from thea_amd import workloads as W
    print(f"case graph reachability S={S}", flush=True)
    v, t = W.ellipsoids_c2()
    mesh = thea_amd.Mesh(v, t)
    p, nrm, _ = thea_amd.MeshSampler(mesh).sampleEvenlyByArea(20000, seed=3, stream_id=1)
    mb = thea_amd.MeshBvh(mesh, leaf_size=1)
    b = thea_amd.PointBvh(p, leaf_size=8)
    got = b.proximityGraph(normals=nrm, max_degree=8, seed=6, num_original=S, surface=mb)
    kd = oracle.KdTree(v, t)
    gi, gs, gd, gavg = oracle.proximity_graph(p, nrm, 8, oracle.graph_picks(len(p), 100, 6), surface=kd)
    if S is None:
        ref = (gi, gs, gd, gavg)
    else:
        ridx, rsep, rdeg = oracle.extract_original(gi, gs, gd, S)
        tot = np.float32(0)
        for i in range(S):
            for j in range(rdeg[i]):
                tot = np.float32(tot + rsep[i, j])
        ref = (ridx, rsep, rdeg, (gavg[0], float(np.float32(tot / np.float32(rdeg.sum())))))
    _cmp(got, ref, f"reachability S={S}")
    plain = b.proximityGraph(normals=nrm, max_degree=8, seed=6, num_original=S)
    assert not np.array_equal(plain[0], got[0])
