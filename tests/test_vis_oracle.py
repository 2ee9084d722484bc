"""Pins the oracle's Visibility restatement: Knop's sphere directions are unit vectors with the
uniform distribution (z and azimuth uniform), and the visibility of points inside, on and outside a
sphere mesh matches the closed forms (0, 1/2, and the complement of the sphere's cap fraction)."""
import numpy as np
import pytest


def test_directions_uniform_on_sphere(oracle):
    from scipy import stats
    pos = np.zeros((1, 3), np.float32)
    org, d = oracle.vis_rays(pos, 200000, 2.0, 7)
    assert np.abs(np.linalg.norm(d.astype(np.float64), axis=1) - 1).max() < 1e-6
    assert np.allclose(org, 0.002 * d, atol=1e-7)
    assert stats.kstest(d[:, 2], "uniform", args=(-1, 2)).pvalue > 1e-3
    az = np.arctan2(d[:, 1], d[:, 0])
    assert stats.kstest(az, "uniform", args=(-np.pi, 2 * np.pi)).pvalue > 1e-3
    # different points and seeds draw different directions
    _, d2 = oracle.vis_rays(np.zeros((2, 3), np.float32), 5, 2.0, 7)
    _, d3 = oracle.vis_rays(np.zeros((1, 3), np.float32), 5, 2.0, 8)
    assert not np.array_equal(d2[:5], d2[5:]) and not np.array_equal(d2[:5], d3)


@pytest.mark.parametrize("where,expect,tol", [("inside", 0.0, 0.0), ("surface", 0.5, 0.02), ("outside", None, 0.02)])
def test_sphere_closed_forms(oracle, where, expect, tol):
    v, t = oracle.geodesic_sphere(5)
    kd = oracle.KdTree(v, t)
    R = 20000
    if where == "inside":
        p = np.array([[0.1, -0.2, 0.05]], np.float32)
    elif where == "surface":
        p = v[:1].astype(np.float32)
    else:
        D = 3.0
        p = np.array([[D, 0, 0]], np.float32)
        # rays from outside hit the unit sphere iff they point into its cap of half-angle asin(1 / D)
        expect = 1 - (1 - np.sqrt(1 - 1 / D ** 2)) / 2
    vis = oracle.visibility(kd, p, R, 2.0, 3)
    assert abs(vis[0] - expect) <= tol + 1e-12, (where, vis[0], expect)
