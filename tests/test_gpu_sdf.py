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
