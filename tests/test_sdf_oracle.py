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


# cone z components (the weights) of rays 5, 9, 17 and 26 of ShapeDiameter's table
# (Algorithms/SurfaceFeatures/Local/ShapeDiameter.hpp:197-228)
CONE_Z = {5: 0.689520, 9: 0.732200, 17: 0.890588, 26: 0.570771, 0: 0.903309}


def _reduce(vals_by_ray, scale):
    vals = np.full(30, -1, np.float32)
    for r, x in vals_by_ray.items():
        vals[r] = x
    out = np.zeros(1)
    O.lib().or_sdf_reduce(vals, 1, float(scale), out)
    return out[0]


def test_reduction_window_and_weights_three_hits():
    """Three hits (rays 5, 9, 17 at 0.5, 0.125, 0.25). With three values std::nth_element is an
    insertion sort, so the values become [0.125, 0.25, 0.5] while the weights stay in ray order
    (ShapeDiameter.hpp:230-252). Median 0.25, variance 7/288: 0.125 is kept ((0.125)^2 <= var),
    0.5 rejected. Each plausible slip changes the answer: centring on the mean rejects 0.125,
    comparing |v - median| with the variance rejects it too, comparing the square with the standard
    deviation keeps 0.5, and pairing each value with its own ray's weight or dropping the weights
    changes the weighted mean."""
    w = {r: float(np.float32(z)) for r, z in CONE_Z.items()}
    got = _reduce({5: 0.5, 9: 0.125, 17: 0.25}, 2.0)
    want = (w[5] * 0.125 + w[9] * 0.25) / ((w[5] + w[9]) * 2.0)
    assert got == want
    assert got != (w[9] * 0.125 + w[17] * 0.25) / ((w[9] + w[17]) * 2.0)  # own-ray weights
    assert got != (0.125 + 0.25) / 2 / 2.0                                   # unweighted


def test_reduction_five_hits_one_outlier():
    """Five hits, one far outlier (rays 0, 5, 9, 17, 26 at 1, 3, 2, 20, 4): median 3, mean 6,
    variance 50, so 1, 2, 3, 4 are kept and 20 is dropped. The values are paired with the ray-order
    weights in the order std::nth_element leaves them, taken from the host library itself."""
    import os
    import subprocess
    import tempfile
    rays = [0, 5, 9, 17, 26]
    vals = [1.0, 3.0, 2.0, 20.0, 4.0]
    src = ("#include <algorithm>\n#include <cstdio>\nint main(){double v[5]={1,3,2,20,4};"
           "std::nth_element(v,v+2,v+5);for(double x:v)std::printf(\"%.17g \",x);return 0;}\n")
    with tempfile.TemporaryDirectory() as d:
        c, exe = os.path.join(d, "nth.cpp"), os.path.join(d, "nth")
        with open(c, "w") as f:
            f.write(src)
        subprocess.run(["g++", "-O2", c, "-o", exe], check=True)
        perm = [float(x) for x in subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()]
    assert sorted(perm) == sorted(vals) and perm[2] == 3.0
    w = [float(np.float32(CONE_Z[r])) for r in rays]
    sv = sw = 0.0
    for k in range(5):
        if (perm[k] - 3.0) ** 2 <= 50.0:
            sv += w[k] * perm[k]
            sw += w[k]
    want = sv / (sw * 8.0)
    got = _reduce(dict(zip(rays, vals)), 8.0)
    assert got == want
    assert abs(got - 2.5 / 8.0) < 0.05  # a weighted mean of 1, 2, 3, 4 over the scale (below the clamp at 1)
