"""Writes the committed golden fixtures under tests/golden/ (run once in the build container):

  models.npz      Thea's own test meshes Data/Models/{cube,sphere,teapot}.obj as float32 vertices
                  and int32 triangles (data files of the reference, converted; read from the
                  reference checkout only here, never at test time).
  testbvh.npz     the inputs of Thea's Code/Source/Test/TestBvh.cpp regenerated from its seeded RNG
                  (Random rng(0xFFFF), Random.cpp:49-132, restated in the oracle) -- 10,000 points,
                  a second 10,000 points, 1,000 triangles, a second 1,000 triangles and the ray
                  direction, in TestBvh's draw order -- plus the oracle's answers to TestBvh's queries.
                  TestBvh builds each Vector3 from three rand01() calls in one argument list, whose
                  evaluation order C++ leaves unspecified; GCC (Thea's Linux toolchain) evaluates
                  them right to left, so the third coordinate is drawn first. Both orders are stored
                  ("gcc" and "ltr"); tests use both.
  metrics_kat.json known answers for TestMetrics.cpp's fixed objects, in FP64 closed form.

Run: python tests/golden/make_fixtures.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
REF_MODELS = "/root/reference/Data/Models"

from oracle import oracle as O  # noqa: E402


def read_obj(path):
    verts, faces = [], []
    with open(path) as f:
        for line in f:
            parts = line.split()
            if not parts:
                continue
            if parts[0] == "v":
                verts.append([float(x) for x in parts[1:4]])
            elif parts[0] == "f":
                idx = [int(p.split("/")[0]) - 1 for p in parts[1:]]
                for k in range(1, len(idx) - 1):  # fan triangulation (all faces here are triangles)
                    faces.append([idx[0], idx[k], idx[k + 1]])
    return np.array(verts, np.float32), np.array(faces, np.int32)


def make_models():
    out = {}
    for name in ("cube", "sphere", "teapot"):
        v, t = read_obj(os.path.join(REF_MODELS, name + ".obj"))
        out[name + "_v"], out[name + "_t"] = v, t
    np.savez_compressed(os.path.join(HERE, "models.npz"), **out)
    return out


def testbvh_inputs(order):
    n = 30000 + 30000 + 9000 + 9000 + 3
    u = O.thea_mt_u01(0xFFFF, n)
    pos = [0]

    def vec():
        a, b, c = u[pos[0]], u[pos[0] + 1], u[pos[0] + 2]
        pos[0] += 3
        return (c, b, a) if order == "gcc" else (a, b, c)

    pts = np.array([vec() for _ in range(10000)], np.float32)
    pts2 = np.array([vec() for _ in range(10000)], np.float32)
    tris = np.array([[vec() for _ in range(3)] for _ in range(1000)], np.float32)
    tris2 = np.array([[vec() for _ in range(3)] for _ in range(1000)], np.float32)
    d = np.array(vec(), np.float32)
    # Eigen normalized(): v / sqrt(squaredNorm()), squaredNorm = x0 + (x1 + x2) in float
    sq = np.float32(d[0] * d[0] + (d[1] * d[1] + d[2] * d[2]))
    d = (d / np.sqrt(sq)).astype(np.float32)
    return pts, pts2, tris, tris2, d


def make_testbvh():
    out = {}
    q = np.array([[0.5, 0.5, 0.5]], np.float32)
    for order in ("gcc", "ltr"):
        pts, pts2, tris, tris2, d = testbvh_inputs(order)
        tv = tris.reshape(-1, 3)
        tt = np.arange(len(tv), dtype=np.int32).reshape(-1, 3)
        kp = O.KdTree(points=pts)
        nn_idx, _, nn_d2, _ = kp.closest(q, dist_bound=0.25, mode=0)          # TestBvh.cpp:204-214
        k3_idx, k3_d2, _ = kp.knn(q, 3, dist_bound=0.25, mode=0)              # :221-236
        pair_idx, _, pair_d2, _ = kp.closest(pts2, mode=0)                     # :243-264 (B vs A)
        kt = O.KdTree(tv, tt)
        tnn_idx, tnn_cp, tnn_d2, _ = kt.closest(q, dist_bound=0.25, mode=0)   # :338-360
        ray_idx, ray_t, _ = kt.ray(np.zeros((1, 3), np.float32), d[None], mode=0)  # :405-439
        # the same answers by brute force (pins the kd-tree oracle)
        b = O.brute_closest(tv, tt, q, dist_bound=0.25)
        assert b[0][0] == tnn_idx[0]
        bi, bt = O.brute_ray(tv, tt, np.zeros((1, 3), np.float32), d[None])
        assert bi[0] == ray_idx[0] and bt[0] == ray_t[0]
        for k, val in dict(points=pts, points2=pts2, tris=tris, tris2=tris2, ray_dir=d, nn_idx=nn_idx, nn_d2=nn_d2,
                           k3_idx=k3_idx, k3_d2=k3_d2, pair_idx=pair_idx, pair_d2=pair_d2, tnn_idx=tnn_idx,
                           tnn_cp=tnn_cp, tnn_d2=tnn_d2, ray_idx=ray_idx, ray_t=ray_t).items():
            out[f"{order}_{k}"] = val
    np.savez_compressed(os.path.join(HERE, "testbvh.npz"), **out)
    return out


def closest_on_triangle_f64(p, a, b, c):
    """Ericson, Real-Time Collision Detection, 5.1.5 (FP64)."""
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = ab @ ap, ac @ ap
    if d1 <= 0 and d2 <= 0:
        return a
    bp = p - b
    d3, d4 = ab @ bp, ac @ bp
    if d3 >= 0 and d4 <= d3:
        return b
    vc = d1 * d4 - d3 * d2
    if vc <= 0 and d1 >= 0 and d3 <= 0:
        return a + d1 / (d1 - d3) * ab
    cp = p - c
    d5, d6 = ab @ cp, ac @ cp
    if d6 >= 0 and d5 <= d6:
        return c
    vb = d5 * d2 - d1 * d6
    if vb <= 0 and d2 >= 0 and d6 <= 0:
        return a + d2 / (d2 - d6) * ac
    va = d3 * d6 - d5 * d4
    if va <= 0 and (d4 - d3) >= 0 and (d5 - d6) >= 0:
        return b + (d4 - d3) / ((d4 - d3) + (d5 - d6)) * (c - b)
    den = 1.0 / (va + vb + vc)
    return a + ab * (vb * den) + ac * (vc * den)


def make_metrics_kat():
    # Test/TestMetrics.cpp:39-56 fixed objects
    p = np.array([2.0, 2.0, 2.0])
    tri = [np.array(x, float) for x in ([-1, -1, -1], [0.5, 0.5, 0], [0, 0.5, 0.5])]
    cp = closest_on_triangle_f64(p, *tri)
    kat = {
        "point_point_3d": {"a": [0, 0, 0], "b": [1, 1, 1], "distance": float(np.sqrt(3.0))},
        "aabb_point": {"lo": [0, 0, 0], "hi": [1, 1, 1], "p": [2, 2, 2], "distance": float(np.sqrt(3.0))},
        "triangle_point": {"tri": [t.tolist() for t in tri], "p": p.tolist(), "closest_point": cp.tolist(),
                           "distance": float(np.linalg.norm(cp - p))},
    }
    with open(os.path.join(HERE, "metrics_kat.json"), "w") as f:
        json.dump(kat, f, indent=1)
    return kat


if __name__ == "__main__":
    O.build()
    m = make_models()
    print({k: v.shape for k, v in m.items()})
    tb = make_testbvh()
    print({k: (v.shape if hasattr(v, "shape") else v) for k, v in tb.items() if not k.endswith(("points", "points2", "tris", "tris2"))})
    print(make_metrics_kat())
