"""GPU tests of thg_shape_diameter (ShapeDiameter::compute) against the oracle: every ray's value
and every SDF bit for bit, with given normals and with the nearest face's normal."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("given_normals,only_interior", [(True, True), (False, True), (True, False)])
def test_sdf_matches_oracle(oracle, given_normals, only_interior):
    import thea_amd
    from thea_amd import workloads as W
    print(f"case sdf normals={given_normals} interior={only_interior}", flush=True)
    v, t = W.torus(256, 128, R=1.0, r=0.3, amp=0.04, seed=5)
    mesh = thea_amd.Mesh(v, t)
    pos, nrm, tri = thea_amd.MeshSampler(mesh).sampleEvenlyByArea(20000, seed=8)
    b = thea_amd.MeshBvh(mesh)
    scale = 2.6
    sdf, times = b.shapeDiameter(pos, nrm if given_normals else None, scale=scale,
                                 only_hit_interior_surfaces=only_interior, want_ray_times=True)
    kd = oracle.KdTree(v, t)
    tn = oracle.tri_props(v, t)[:, :3]
    if given_normals:
        ref_n = nrm
    else:
        ridx, _, _, _ = kd.closest(pos, mode=0, want_cp=False)
        ref_n = tn[ridx]
    ref, rtimes = oracle.shape_diameter(kd, tn, pos, ref_n, scale, only_interior)
    assert np.array_equal(times.view(np.int32), rtimes.view(np.int32))
    assert np.array_equal(sdf, ref)
    assert (sdf >= 0).mean() > 0.99 and sdf.max() <= 1
    # the tube's diameter is 2 r = 0.6 (+- the displacement): SDF ~ 0.6 / 2.6 inside
    assert 0.15 < np.median(sdf[sdf >= 0]) < 0.3


def test_sdf_device_arrays_match_host():
    """Device-resident inputs give device results identical to the host-array call; an empty batch
    is accepted."""
    import torch
    import thea_amd
    from thea_amd import workloads as W
    print("case sdf device arrays", flush=True)
    v, t = W.torus(128, 64, R=1.0, r=0.35, amp=0.03, seed=6)
    mesh = thea_amd.Mesh(v, t)
    pos, nrm, _ = thea_amd.MeshSampler(mesh).sampleEvenlyByArea(5000, seed=3)
    b = thea_amd.MeshBvh(mesh)
    host, htimes = b.shapeDiameter(pos, nrm, scale=2.7, want_ray_times=True)
    dpos, dnrm = torch.as_tensor(pos, device="cuda"), torch.as_tensor(nrm, device="cuda")
    dev, dtimes = b.shapeDiameter(dpos, dnrm, scale=2.7, want_ray_times=True)
    assert dev.is_cuda and dev.dtype == torch.float64
    assert np.array_equal(dev.cpu().numpy(), host)
    assert np.array_equal(dtimes.cpu().numpy().view(np.int32), htimes.view(np.int32))
    assert np.array_equal(b.shapeDiameter(dpos, None, scale=2.7).cpu().numpy(), b.shapeDiameter(pos, None, scale=2.7))
    assert b.shapeDiameter(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32), scale=1.0).shape == (0,)
