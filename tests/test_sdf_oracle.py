"""Pins the oracle's restatement of ShapeDiameter (ray generation and reduction) with what the
mathematics fixes: the ray directions form an orthonormal frame around the inward normal (each ray
makes the angle of its cone direction with it), origins sit 0.001 scale inside, and the reduction of
equal hit times is that time over the scale, clamped to [0, 1]; no hits give -1."""
import numpy as np

from oracle import oracle as O


def test_rays_follow_the_inward_frame():
    g = np.random.default_rng(1)
    nrm = g.normal(size=(500, 3)).astype(np.float32)
    nrm[:5] = [[1, 0, 0], [0, 1, 0], [0, 0, -1], [1, 1, 0], [0, 0, 1]]
    pos = g.normal(size=(500, 3)).astype(np.float32)
    org = np.zeros((15000, 3), np.float32)
    dr = np.zeros((15000, 3), np.float32)
    O.lib().or_sdf_rays(pos, nrm, 500, np.float32(2.0), org, dr)
    inw = -nrm / np.linalg.norm(nrm, axis=1)[:, None]
    d = dr.reshape(500, 30, 3).astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=2), 1, atol=1e-5)
    cz = np.einsum("nrk,nk->nr", d, inw)  # cos angle to the inward normal = the cone direction's z
    dirs = d.reshape(500, 30, 3)
    assert cz.min() > 0.5
    # the same cone z for every point: the frame is a rotation
    assert np.allclose(cz, cz[0][None, :], atol=1e-5)
    assert np.allclose(org.reshape(500, 30, 3), (pos + 0.002 * inw)[:, None, :], atol=1e-6)
    # pairwise angles between rays are the cone's (rigid frame, right-handed: no reflection)
    G = np.einsum("nrk,nsk->nrs", dirs, dirs)
    assert np.allclose(G, G[0][None], atol=1e-5)
    tri = np.einsum("k,k->", np.cross(dirs[0, 0], dirs[0, 1]), dirs[0, 2])
    assert np.allclose([np.einsum("k,k->", np.cross(dirs[i, 0], dirs[i, 1]), dirs[i, 2]) for i in range(500)], tri,
                       atol=1e-5)


def test_reduction_closed_forms():
    vals = np.full((3, 30), -1, np.float32)
    vals[0, :] = 0.5                      # all rays hit at 0.5: sdf = 0.5 / scale
    vals[1, [3, 7, 11]] = [0.2, 0.2, 0.2]  # three rays at the same time
    out = np.zeros(3)
    O.lib().or_sdf_reduce(vals.ravel(), 3, 2.0, out)
    assert out[0] == np.float64(np.float32(0.5)) / 2.0
    assert out[1] == np.float64(np.float32(0.2)) / 2.0
    assert out[2] == -1.0
    vals[0, :] = 100.0
    O.lib().or_sdf_reduce(vals.ravel(), 3, 2.0, out)
    assert out[0] == 1.0  # clamped
