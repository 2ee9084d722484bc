"""GPU tests of thg_visibility (Visibility::compute) against the oracle: the same rays, the same
hits, the same fractions."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("num_rays,explicit_scale", [(100, False), (37, True)])
def test_visibility_matches_oracle(oracle, num_rays, explicit_scale):
    import thea_amd
    from thea_amd import workloads as W
    print(f"case visibility rays={num_rays} scale={explicit_scale}", flush=True)
    v, t = W.torus(192, 96, R=1.0, r=0.3, amp=0.05, seed=4)
    mesh = thea_amd.Mesh(v, t)
    pos, _, _ = thea_amd.MeshSampler(mesh).sampleEvenlyByArea(5000, seed=5)
    b = thea_amd.MeshBvh(mesh)
    lo, hi = b.getBounds()
    e = hi - lo
    default_scale = np.float32(np.sqrt(e[0] * e[0] + (e[1] * e[1] + e[2] * e[2])))
    scale = 2.5 if explicit_scale else None
    vis = b.visibility(pos, num_rays=num_rays, seed=11, scale=scale)
    ref = oracle.visibility(oracle.KdTree(v, t), pos, num_rays, default_scale if scale is None else scale, 11)
    assert np.array_equal(vis, ref)
    # a surface point of a closed mesh sees at most half the sphere
    assert vis.max() <= 0.75 and 0.2 < vis.mean() < 0.55


def test_visibility_device_and_chunks(oracle):
    import torch
    import thea_amd
    from thea_amd import workloads as W
    print("case visibility device", flush=True)
    v, t = W.torus(64, 32, seed=1)
    mesh = thea_amd.Mesh(v, t)
    pos, _, _ = thea_amd.MeshSampler(mesh).sampleEvenlyByArea(2000, seed=2)
    b = thea_amd.MeshBvh(mesh)
    host = b.visibility(pos, num_rays=64, seed=3)
    dev = b.visibility(torch.as_tensor(pos, device="cuda"), num_rays=64, seed=3)
    assert dev.is_cuda and np.array_equal(dev.cpu().numpy(), host)
    # 1,700 points x 10,000 rays: more than one 2^24-ray chunk
    many = b.visibility(pos[:1700], num_rays=10000, seed=4, scale=2.0)
    assert np.array_equal(many, oracle.visibility(oracle.KdTree(v, t), pos[:1700], 10000, 2.0, 4))
    assert len(b.visibility(np.zeros((0, 3), np.float32))) == 0
