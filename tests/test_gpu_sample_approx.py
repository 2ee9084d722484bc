"""GPU tests of APPROXIMATE-mode area sampling (thg_sample_area_approx) against the oracle: the same
count, triangles, positions and normals bit for bit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,seed", [(100000, 1), (777, 9), (3, 4)])
def test_approx_matches_oracle(oracle, n, seed):
    import thea_amd
    from thea_amd import workloads as W
    print(f"case approx n={n} seed={seed}", flush=True)
    v, t = W.ellipsoids_c2(n_ellipsoids=8)
    mesh = thea_amd.Mesh(v, t)
    pos, nrm, tri = thea_amd.MeshSampler(mesh).sampleEvenlyByArea(n, seed=seed, stream_id=2, count_mode="approximate")
    rpos, rnrm, rtri = oracle.sample_area_approx(v, t, n, seed, 2)
    assert len(pos) == len(rpos)
    assert np.array_equal(tri, rtri) and np.array_equal(pos, rpos) and np.array_equal(nrm, rnrm)


def test_approx_device_and_capacity():
    import torch
    import thea_amd
    from thea_amd import _capi as capi
    from thea_amd import workloads as W
    print("case approx device", flush=True)
    v, t = W.torus(64, 32, seed=1)
    mesh = thea_amd.Mesh(v, t)
    s = thea_amd.MeshSampler(mesh)
    hp, hn, ht = s.sampleEvenlyByArea(5000, seed=2, count_mode="approximate")
    dp, dn, dt = s.sampleEvenlyByArea(5000, seed=2, count_mode="approximate", device=True)
    assert dp.is_cuda and np.array_equal(dp.cpu().numpy(), hp) and np.array_equal(dt.cpu().numpy(), ht)
    assert len(s.sampleEvenlyByArea(0, count_mode="approximate")[0]) == 0
    out = np.zeros((10, 3), np.float32)
    got = C.c_int64()
    rc = capi.lib().thg_sample_area_approx(mesh.h, 5000, C.c_uint64(2), C.c_uint64(0), 10,
                                           out.ctypes.data_as(C.c_void_p), None, None, C.byref(got),
                                           capi.THG_MEM_HOST, None)
    assert rc == capi.THG_E_INVALID
