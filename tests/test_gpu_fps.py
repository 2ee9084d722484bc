"""GPU tests of separation-based sampling (FurthestPointSampling::subsample and
MeshSampler::sampleEvenlyBySeparation) against the oracle: the same picks in the same order, and the
same sample positions, normals and triangles."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def surface_points(seed, n, components=1):
    g = np.random.default_rng(seed)
    pts = []
    for c in range(components):
        u = g.normal(size=(n // components, 3))
        pts.append(u / np.linalg.norm(u, axis=1)[:, None] * [1.0, 0.6, 0.8] + [4.0 * c, 0, 0])
    p = np.concatenate(pts).astype(np.float32)
    return p[g.permutation(len(p))]


@pytest.mark.parametrize("n,k,md,components", [(3000, 1000, 8, 1), (6000, 2000, 16, 3), (2000, 2000, 8, 1),
                                               (500, 1, 8, 1), (2, 2, 8, 1)])
def test_fps_matches_oracle(oracle, n, k, md, components):
    import thea_amd
    print(f"case fps n={n} k={k} md={md} components={components}", flush=True)
    p = surface_points(n + k, n, components)
    got = thea_amd.furthestPointSampling(p, k, max_degree=md, graph_seed=5)
    idx, sep, _, _ = oracle.proximity_graph(p, None, md, oracle.graph_picks(len(p), 100, 5))
    ref = oracle.fps(idx, sep, k)
    assert got.dtype == np.int64
    assert np.array_equal(got, ref), (len(got), len(ref), np.nonzero(got[:len(ref)] != ref[:len(got)])[0][:5])
    if k == n:
        assert sorted(got) == list(range(n))


def test_fps_device_arrays_and_empty():
    import torch
    import thea_amd
    print("case fps device", flush=True)
    p = surface_points(11, 4000)
    host = thea_amd.furthestPointSampling(p, 700)
    dev = thea_amd.furthestPointSampling(torch.as_tensor(p, device="cuda"), 700)
    assert dev.is_cuda and np.array_equal(dev.cpu().numpy(), host)
    assert len(thea_amd.furthestPointSampling(p, 0)) == 0
    with pytest.raises(thea_amd.TheaGpuError):
        thea_amd.furthestPointSampling(p[:10], 11)


@pytest.mark.parametrize("factor", [-1.0, 2.5])
def test_sample_separation_matches_oracle(oracle, factor):
    import thea_amd
    from thea_amd import workloads as W
    print(f"case separation factor={factor}", flush=True)
    v, t = W.torus(96, 48, R=1.0, r=0.3, amp=0.04, seed=2)
    mesh = thea_amd.Mesh(v, t)
    n = 1500
    pos, nrm, tri = thea_amd.MeshSampler(mesh).sampleEvenlyBySeparation(n, seed=4, stream_id=1,
                                                                       oversampling_factor=factor, graph_seed=6)
    over = int(np.ceil(np.float32(3.0 if factor < 0 else factor) * np.float32(n)))
    opos, onrm, otri = oracle.sample_area(v, t, over, 4, 1)[:3]
    idx, sep, _, _ = oracle.proximity_graph(opos, None, 8, oracle.graph_picks(over, 100, 6))
    sel = oracle.fps(idx, sep, n)
    assert len(sel) == n and len(pos) == n
    assert np.array_equal(pos, opos[sel]) and np.array_equal(nrm, onrm[sel]) and np.array_equal(tri, otri[sel])
    # the picks are spread out: the smallest pairwise distance is far above that of random samples
    from scipy.spatial import cKDTree
    d_sel = cKDTree(pos).query(pos, k=2)[0][:, 1].min()
    d_rand = cKDTree(opos[:n]).query(opos[:n], k=2)[0][:, 1].min()
    assert d_sel > 3 * d_rand


def test_sample_separation_device():
    import torch
    import thea_amd
    from thea_amd import workloads as W
    print("case separation device", flush=True)
    v, t = W.torus(64, 32, seed=3)
    mesh = thea_amd.Mesh(v, t)
    s = thea_amd.MeshSampler(mesh)
    hp, hn, ht = s.sampleEvenlyBySeparation(800, seed=1)
    dp, dn, dt = s.sampleEvenlyBySeparation(800, seed=1, device=True)
    assert dp.is_cuda
    assert np.array_equal(dp.cpu().numpy(), hp) and np.array_equal(dn.cpu().numpy(), hn)
    assert np.array_equal(dt.cpu().numpy(), ht)
    assert len(s.sampleEvenlyBySeparation(0)[0]) == 0
