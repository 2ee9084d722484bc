"""The C++ host API (include/thea_gpu_bvh.hpp), driven as a Thea caller would drive it
(tests/shim/shim_example.cpp, built by the package build), gives the same results as the Python
binding over the same C ABI."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
VERTS = np.array([0, 0, 0, 1, 1, 0, 1, 0, 0, 0, 1, 0, 0, 1, 1, 1, 1, 1, 0, 0, 1, 1, 0, 1], np.float32).reshape(-1, 3)
TRIS = np.array([0, 1, 2, 0, 3, 1, 0, 4, 3, 0, 6, 4, 3, 5, 1, 3, 4, 5,
                 2, 1, 5, 2, 5, 7, 0, 2, 7, 0, 7, 6, 6, 7, 5, 6, 5, 4], np.int32).reshape(-1, 3)


def run_shim(*args):
    exe = os.path.join(ROOT, "tests", "shim", "shim_example")
    assert os.path.exists(exe), "tests/shim/shim_example is built by __graft_entry__.build()"
    out = subprocess.run([exe, *args], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = {}
    for line in out.stdout.splitlines():
        key, *rest = line.split()
        lines.setdefault(key, []).append(rest)
    return lines


def test_cpp_api_matches_python_binding(oracle):
    import thea_amd
    print("case C++ shim on the unit cube", flush=True)
    got = run_shim()
    mesh = thea_amd.Mesh(VERTS, TRIS)
    bvh = thea_amd.MeshBvh(mesh)
    f = np.float32  # the shim builds its queries in float arithmetic
    q = np.array([[f(-0.5) + f(0.5) * f(i), f(0.3), f(1.7) - f(0.4) * f(i)] for i in range(5)], np.float32)
    idx, dist, cp, _ = bvh.closest(q)
    for i, row in enumerate(got["closest"]):
        assert int(row[0]) == i and int(row[1]) == idx[i]
        assert np.float32(float(row[2])) == dist[i]
        assert np.array_equal(np.array([float(x) for x in row[3:]], np.float32), cp[i])
    kidx, kdist = bvh.kClosestElements(np.array([[0.5, 0.5, 2.0]], np.float32), 3)
    assert [int(r[0]) for r in got["knn"]] == list(kidx[0])
    assert np.array_equal(np.array([float(r[1]) for r in got["knn"]], np.float32), kdist[0])
    ridx, rt, _ = bvh.rayStructureIntersection(np.array([[0.25, 0.5, 3.0]], np.float32),
                                               np.array([[0.0, 0.0, -1.0]], np.float32))
    hit = got["ray"][0]
    assert int(hit[0]) == ridx[0] and np.float32(float(hit[1])) == rt[0] and hit[2] == "0"
    sampler = thea_amd.MeshSampler(mesh)
    pos, nrm, tri = sampler.sampleEvenlyByArea(4, seed=7, stream_id=0)
    assert [int(r[0]) for r in got["sample"]] == list(tri)
    assert np.array_equal(np.array([[float(x) for x in r[1:]] for r in got["sample"]], np.float32), pos)
    # MeshSampler::CountMode::APPROXIMATE through the C++ layer equals the binding's approximate mode
    apos, _, atri = sampler.sampleEvenlyByArea(40, seed=7, stream_id=5, count_mode="approximate")
    assert len(got["approx"]) == len(atri) and 30 <= len(atri) <= 50
    assert [int(r[0]) for r in got["approx"]] == list(atri)
    assert np.array_equal(np.array([[float(x) for x in r[1:]] for r in got["approx"]], np.float32), apos)
    # ICP of 2000 further samples shifted by +0.05 in x back onto the cube, least-squares fit
    surf, _, _ = sampler.sampleEvenlyByArea(2000, seed=7, stream_id=1)
    moved = surf.astype(np.float64) + np.array([0.05, 0.0, 0.0])
    tr, err = thea_amd.Icp3(kabsch=True).align(moved, bvh)
    tx, ty, tz, cerr = (float(x) for x in got["icp"][0])
    assert abs(tx - tr[0, 3]) <= 1e-6 and abs(ty - tr[1, 3]) <= 1e-6 and abs(tz - tr[2, 3]) <= 1e-6
    assert abs(cerr - err) <= 1e-3 * err
    assert -0.06 < tx < -0.02  # moved back towards the cube (ICP slides on the faces parallel to x)
    # double point BVH + Icp3<double> (ShapeAlign's form): the nearest target sample against the
    # oracle's double brute force, and the alignment against the double-precision Python path
    target = surf.astype(np.float64)
    i0, d0 = oracle.brute_closest_points_f64(target, np.array([[0.3, 0.2, 1.5]]))
    row = got["closest64"][0]
    assert int(row[0]) == i0[0] and float(row[1]) == np.sqrt(d0[0])
    assert [float(x) for x in row[2:]] == list(target[i0[0]])
    tb = thea_amd.PointBvhF64(target)
    tr_d, err_d = thea_amd.Icp3().align(moved, tb)
    r = [float(x) for x in got["icp64"][0]]
    assert r[:3] == [tr_d[0, 3], tr_d[1, 3], tr_d[2, 3]] and r[3] == err_d
    # ShapeDiameter with the default scale: the diagonal of the 1.05-padded unit cube
    scale = float(got["scale"][0][0])
    assert abs(scale - 1.05 * np.sqrt(3.0)) < 1e-6
    lo, hi = bvh.getBounds()
    assert np.allclose(lo, -0.025) and np.allclose(hi, 1.025)
    ref_n = bvh.shapeDiameter(pos, nrm)
    ref_p = bvh.shapeDiameter(pos)
    assert [float(r[1]) for r in got["sdf"]] == list(ref_n)
    assert [float(r[2]) for r in got["sdf"]] == list(ref_p)
    e = hi - lo
    oracle_sdf, _ = oracle.shape_diameter(oracle.KdTree(VERTS, TRIS), oracle.tri_props(VERTS, TRIS)[:, :3], pos, nrm,
                                          float(np.sqrt(e[0] * e[0] + (e[1] * e[1] + e[2] * e[2]))))
    assert np.array_equal(ref_n, oracle_sdf)
    assert np.all((ref_n > 0) & (ref_n <= 1.0))


def test_cpp_api_testbvh_call_sites(oracle, tmp_path):
    """Thea's TestBvh.cpp call sites (point NN with bound 0.25, 3-NN into a BoundedSortedArrayN<3>,
    BVH-vs-BVH closestPair and kClosestPairs, the triangle NN with closest point after
    enableNearestNeighborAcceleration, and the three ray queries) through the C++ layer on TestBvh's
    own regenerated inputs, against the committed oracle answers."""
    g = np.load(os.path.join(ROOT, "tests", "golden", "testbvh.npz"))
    for order in ("gcc", "ltr"):
        print(f"case TestBvh call sites through the C++ layer, {order} argument order", flush=True)
        pts, pts2, tris = g[f"{order}_points"], g[f"{order}_points2"], g[f"{order}_tris"]
        blob = np.concatenate([pts.ravel(), pts2.ravel(), tris.ravel(), g[f"{order}_ray_dir"].ravel()]).astype(np.float32)
        path = tmp_path / f"testbvh_{order}.f32"
        blob.tofile(path)
        got = run_shim(str(path))
        # closestElement with dist_bound 0.25 (TestBvh.cpp:204-215)
        nn = got["tb_nn"][0]
        assert int(nn[0]) == int(g[f"{order}_nn_idx"][0])
        assert np.float32(float(nn[1])) == np.sqrt(g[f"{order}_nn_d2"][0])
        # kClosestPairs into BoundedSortedArrayN<3, NeighborPair> (TestBvh.cpp:221-236)
        assert [int(r[1]) for r in got["tb_k3"]] == list(g[f"{order}_k3_idx"][0])
        assert all(int(r[0]) == 3 for r in got["tb_k3"])
        assert np.array_equal(np.array([float(r[2]) for r in got["tb_k3"]], np.float32), g[f"{order}_k3_d2"][0])
        # closestPair(new_bvh) (TestBvh.cpp:256-262): the least (d2, query) over the per-query answers
        pd2, pidx = g[f"{order}_pair_d2"], g[f"{order}_pair_idx"]
        q = int(np.argmin(pd2))
        pr = got["tb_pair"][0]
        assert (int(pr[0]), int(pr[1])) == (q, int(pidx[q]))
        assert np.float32(float(pr[2])) == pd2[q]
        assert np.array_equal(np.array([float(x) for x in pr[3:]], np.float32), pts[pidx[q]])
        # kClosestPairs(new_bvh, nbrs) (TestBvh.cpp:282-297): the 3 least pairs over all query points
        kidx, kd2, _ = oracle.KdTree(points=pts).knn(pts2, 3)
        cand = sorted((float(kd2[i, j]), i, int(kidx[i, j])) for i in range(len(pts2)) for j in range(3))[:3]
        assert [(int(r[1]), int(r[2])) for r in got["tb_kpairs"]] == [(c[1], c[2]) for c in cand]
        assert np.array_equal(np.array([float(r[3]) for r in got["tb_kpairs"]], np.float32),
                              np.array([c[0] for c in cand], np.float32))
        # k beyond the GPU limit: a clear THG_E_UNSUPPORTED (-5) error, not a silent truncation
        assert got["tb_bigk"][0][0] == "-5" and "exceeds" in " ".join(got["tb_bigk"][0])
        # triangles (TestBvh.cpp:338-360): closestElement with closest point
        t = got["tb_tnn"][0]
        assert int(t[0]) == int(g[f"{order}_tnn_idx"][0])
        assert np.float32(float(t[1])) == np.sqrt(g[f"{order}_tnn_d2"][0])
        assert np.array_equal(np.array([float(x) for x in t[2:]], np.float32), g[f"{order}_tnn_cp"][0])
        # rays from the origin (TestBvh.cpp:405-439): rayIntersects, rayIntersectionTime, rayStructureIntersection
        r = got["tb_ray"][0]
        assert r[0] == "1" and int(r[2]) == int(g[f"{order}_ray_idx"][0])
        assert np.float32(float(r[1])) == g[f"{order}_ray_t"][0] == np.float32(float(r[3]))
