"""GPU tests of thg_icp_align (Thea's Icp3::align) against the numpy restatement in oracle/icp.py
with the oracle's kd-tree as the nearest-point search: same iteration counts, transforms within
1e-9, errors within 1e-9 relative; a batch of K starting transforms gives each problem exactly the
result of aligning it alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ellipsoid(g, n, axes=(1.0, 0.6, 0.3)):
    u = g.normal(size=(n, 3))
    return (u / np.linalg.norm(u, axis=1)[:, None] * np.array(axes)).astype(np.float32)


def _axis_rot(axis, angle):
    ax = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * K @ K


def _xf(R, t):
    return np.hstack([R, np.asarray(t, np.float64).reshape(3, 1)])


def _check(tr, err, its, ref, label):
    rtr, rerr, rits = ref
    assert its == rits, f"{label}: iterations {its} vs {rits}"
    assert np.allclose(tr, rtr, rtol=0, atol=1e-9), f"{label}: transform differs by {np.abs(tr - rtr).max()}"
    assert abs(err - rerr) <= 1e-9 * max(abs(rerr), 1e-30), f"{label}: error {err} vs {rerr}"


@pytest.mark.parametrize("kabsch", [False, True])
def test_icp_points_matches_oracle(oracle, kabsch):
    import thea_amd
    from oracle import icp as I
    print(f"case icp points kabsch={kabsch}", flush=True)
    g = np.random.default_rng(11)
    to = _ellipsoid(g, 30000)
    src = to[::7].astype(np.float64) @ _axis_rot([1, 2, 3], 0.08).T + np.array([0.02, -0.01, 0.03])
    bvh = thea_amd.PointBvh(to)
    icp = thea_amd.Icp3(kabsch=kabsch)
    tr, err, its = icp.alignBatch(src, bvh)
    kd = oracle.KdTree(points=to)
    ref = I.icp_align(lambda x: kd.closest(x, mode=0)[1], src, kabsch=kabsch)
    _check(tr[0], err[0], its[0], ref, "points")
    if kabsch:
        assert err[0] < 0.2 * I.measure_error(I.identity(), src, kd.closest(src.astype(np.float32), mode=0)[1])


def test_icp_batch_equals_single_and_oracle(oracle):
    import thea_amd
    from oracle import icp as I
    print("case icp batch", flush=True)
    g = np.random.default_rng(12)
    to = _ellipsoid(g, 20000)
    src = to[::5].astype(np.float64) + 0.01
    inits = np.stack([_xf(_axis_rot(g.normal(size=3), a), g.normal(size=3) * 0.05)
                      for a in (0.0, 0.05, 0.3, 1.0, 2.5, 3.1)])
    bvh = thea_amd.PointBvh(to)
    icp = thea_amd.Icp3(kabsch=True)
    tr, err, its = icp.alignBatch(src, bvh, initial=inits)
    kd = oracle.KdTree(points=to)
    for k in range(len(inits)):
        t1, e1, i1 = icp.alignBatch(src, bvh, initial=inits[k:k + 1])
        assert np.array_equal(t1[0], tr[k]) and e1[0] == err[k] and i1[0] == its[k], f"problem {k}"
        ref = I.icp_align(lambda x: kd.closest(x, mode=0)[1], I.xf_apply(inits[k], src), kabsch=True)
        _check(tr[k], err[k], its[k], ref, f"problem {k}")


def test_icp_mesh_weights_up(oracle):
    import thea_amd
    from oracle import icp as I
    from thea_amd import workloads as W
    print("case icp mesh", flush=True)
    v, t = W.torus(96, 48, R=1.0, r=0.3, amp=0.03, seed=7)
    g = np.random.default_rng(13)
    sampler = thea_amd.MeshSampler(thea_amd.Mesh(v, t))
    pts, _, _ = sampler.sampleEvenlyByArea(8000, seed=3)
    src = pts.astype(np.float64) @ _axis_rot([0, 0, 1], 0.1).T + np.array([0.05, 0.0, 0.0])
    wt, wr = g.uniform(0.5, 1.5, len(src)), g.uniform(0.5, 1.5, len(src))
    bvh = thea_amd.MeshBvh(thea_amd.Mesh(v, t))
    icp = thea_amd.Icp3(max_iterations=15, kabsch=True)
    icp.setUpVector([0, 0, 2.0])
    tr, err, its = icp.alignBatch(src, bvh, translation_weights=wt, rotation_weights=wr)
    kd = oracle.KdTree(v, t)
    ref = I.icp_align(lambda x: kd.closest(x, mode=0)[1], src, max_it=15, wt=wt, wr=wr, up=[0, 0, 1.0], kabsch=True)
    _check(tr[0], err[0], its[0], ref, "mesh")
    assert np.allclose(tr[0][2, :3], [0, 0, 1], atol=1e-12)


def test_icp_edge_cases():
    import thea_amd
    from thea_amd import _capi as capi
    print("case icp edge cases", flush=True)
    g = np.random.default_rng(14)
    to = _ellipsoid(g, 1000)
    bvh = thea_amd.PointBvh(to)
    icp = thea_amd.Icp3()
    tr, err = icp.align(np.zeros((0, 3)), bvh)
    assert np.array_equal(tr, np.hstack([np.eye(3), np.zeros((3, 1))])) and err == 0
    tr, err, its = icp.alignBatch(to.astype(np.float64), bvh)
    assert its[0] == 0 and err[0] == 0 and np.array_equal(tr[0], np.hstack([np.eye(3), np.zeros((3, 1))]))
    empty = thea_amd.PointBvh(np.zeros((0, 3), np.float32))
    with pytest.raises(capi.TheaGpuError) as e:
        icp.align(to.astype(np.float64), empty)
    assert e.value.code == capi.THG_E_INVALID
