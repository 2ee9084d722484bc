"""Mesh and point files (host code of the library, no device): Thea's OFF/OBJ reading rules, its
face triangulation (quad diagonal rule, ear clipping) checked against areas and orientations, and
MeshSample's point-file format checked against C's %g (the C++ stream's default float format)."""
import os
import struct

import numpy as np
import pytest

from thea_amd import io

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "models")


def write(tmp_path, name, text, mode="w"):
    p = tmp_path / name
    with open(p, mode) as f:
        f.write(text)
    return str(p)


def tri_area_normal(v, t):
    a, b, c = v[t[:, 0]].astype(np.float64), v[t[:, 1]].astype(np.float64), v[t[:, 2]].astype(np.float64)
    return 0.5 * np.cross(b - a, c - a)


def test_cube_fixture():
    v, t, f, nf = io.read_mesh(os.path.join(GOLD, "cube.obj"))
    assert v.shape == (8, 3) and t.shape == (12, 3) and nf == 12
    assert list(f) == list(range(12))
    assert list(t[0]) == [0, 1, 2] and list(t[11]) == [5, 6, 4]
    # a closed, consistently oriented surface: the signed normals sum to zero, total area 6
    an = tri_area_normal(v, t)
    assert np.allclose(an.sum(0), 0) and np.isclose(np.linalg.norm(an, axis=1).sum(), 6.0)


def test_sphere_fixture():
    path = os.path.join(GOLD, "sphere.obj")
    v, t, f, nf = io.read_mesh(path)
    # the same file parsed independently: decimal -> double -> float, 1-based triangle indices
    pv, pf = [], []
    for line in open(path):
        w = line.split()
        if w and w[0] == "v":
            pv.append([float(x) for x in w[1:4]])
        elif w and w[0] == "f":
            pf.append([int(x.split("/")[0]) - 1 for x in w[1:]])
    assert np.array_equal(v, np.array(pv, np.float64).astype(np.float32))
    assert all(len(x) == 3 for x in pf) and np.array_equal(t, np.array(pf, np.int32))
    assert len(t) == nf == 1888 and list(f) == list(range(1888))
    assert np.allclose(np.linalg.norm(v, axis=1), 1.0, atol=1e-2)


def test_obj_rules(tmp_path):
    text = """# comment
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
vn 0 0 1
vt 0.5 0.5
g first
f 1/1/1 2/1/1 3/1/1
f -4 -2 -1
f 1 5 2
f 0 1 2
p 1
f 1 2
o second
f 1//1 3//1 4//1
f 1 2 3 4
"""
    v, t, f, nf = io.read_mesh(write(tmp_path, "m.obj", text))
    assert v.shape == (4, 3)
    # malformed faces (index 5 of 4, index 0) are skipped without an index; "p 1" and "f 1 2" take
    # indices 2 and 3 but have too few vertices
    assert nf == 4
    assert [list(x) for x in t[:3]] == [[0, 1, 2], [0, 2, 3], [0, 2, 3]]
    assert list(f) == [0, 1, 4, 5, 5]
    assert [list(x) for x in t[3:]] == [[0, 1, 2], [2, 3, 0]]  # a convex quad split along 0-2


def off_ascii(verts, faces, extra=""):
    lines = ["OFF", "# comment", "", f"{len(verts)} {len(faces)} 0"]
    lines += [" ".join(f"{x:.9g}" for x in p) for p in verts]
    lines += [" ".join(str(x) for x in [len(fc)] + list(fc)) + extra for fc in faces]
    return "\n".join(lines) + "\n"


def off_binary(verts, faces):
    b = b"OFF BINARY\n" + struct.pack(">iii", len(verts), len(faces), 0)
    for p in verts:
        b += struct.pack(">fff", *p)
    for fc in faces:
        b += struct.pack(">i", len(fc)) + struct.pack(f">{len(fc)}i", *fc) + struct.pack(">i", 0)
    return b


def test_off_ascii_and_binary_agree(tmp_path):
    verts = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0.5, 0.5, 1)]
    faces = [(0, 1, 2), (0, 2, 3), (0, 1, 1), (0, 1, 9), (0, 1, 4), (1, 2, 4), (2, 3, 4), (3, 0, 4)]
    a = io.read_mesh(write(tmp_path, "a.off", off_ascii(verts, faces, extra=" 0.1 0.2 0.3")))
    b = io.read_mesh(write(tmp_path, "b.off", off_binary(verts, faces), "wb"))
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)
    assert a[3] == b[3] == 6
    assert list(a[2]) == [0, 1, 4, 5, 6, 7]  # faces 2 (repeated vertex) and 3 (index 9) skipped


def test_off_rejects_other_headers(tmp_path):
    from thea_amd import TheaGpuError
    with pytest.raises(TheaGpuError):
        io.read_mesh(write(tmp_path, "c.off", "COFF\n1 0 0\n0 0 0\n"))
    with pytest.raises(TheaGpuError):
        io.read_mesh(write(tmp_path, "c.ply", "ply\n"))


def test_quad_diagonal_rule(tmp_path):
    # a dart: splitting along 0-2 would put a triangle outside, so Thea splits along 1-3
    verts = [(0, 0, 0), (2, 0.5, 0), (0, 1, 0), (0.5, 0.5, 0)]
    v, t, _, _ = io.read_mesh(write(tmp_path, "q.off", off_ascii(verts, [(0, 1, 2, 3)])))
    assert [list(x) for x in t] == [[1, 2, 3], [3, 0, 1]]
    an = tri_area_normal(v, t)
    assert np.all(an[:, 2] > 0)  # both triangles keep the quad's winding
    # a quad with three collinear corners keeps only its non-degenerate triangle
    v, t, _, _ = io.read_mesh(write(tmp_path, "d.off", off_ascii([(0, 0, 0), (1, 0, 0), (2, 0, 0), (1, 1, 0)],
                                                                  [(0, 1, 2, 3)])))
    assert [list(x) for x in t] == [[2, 3, 0]]


def shoelace(p):
    x, y = p[:, 0], p[:, 1]
    return 0.5 * (np.dot(x, np.roll(y, -1)) - np.dot(np.roll(x, -1), y))


@pytest.mark.parametrize("shape", ["hexagon", "star", "comb", "clockwise"])
def test_ear_clipping_tiles_the_polygon(tmp_path, shape):
    if shape == "hexagon":
        a = np.arange(6) * np.pi / 3
        p2 = np.stack([np.cos(a), np.sin(a)], 1)
    elif shape == "star":
        a = np.arange(10) * np.pi / 5
        r = np.where(np.arange(10) % 2 == 0, 1.0, 0.4)
        p2 = np.stack([r * np.cos(a), r * np.sin(a)], 1)
    elif shape == "comb":
        p2 = np.array([[0, 0], [5, 0], [5, 2], [4, 2], [4, 1], [3, 1], [3, 2], [2, 2], [2, 1], [1, 1], [1, 2], [0, 2]],
                      float)
    else:
        a = -np.arange(7) * 2 * np.pi / 7
        p2 = np.stack([np.cos(a), np.sin(a)], 1)
    # place the polygon on a tilted plane in 3-D
    basis = np.array([[0.6, 0.0, 0.8], [0.0, 1.0, 0.0]])
    p3 = (p2 @ basis + [0.3, -0.2, 0.5]).astype(np.float32)
    v, t, f, _ = io.read_mesh(write(tmp_path, "p.off", off_ascii(p3, [tuple(range(len(p3)))])))
    n = len(p3)
    assert len(t) == n - 2 and np.all(f == 0)
    normal = np.cross(basis[0], basis[1])
    an = tri_area_normal(v, t) @ normal
    area = abs(shoelace(p2))
    # every triangle has the polygon's orientation, and together they cover exactly its area
    sign = np.sign(shoelace(p2))
    assert np.all(an * sign > 0)
    assert np.isclose(an.sum() * sign, area, rtol=1e-5)


def test_pts_format_and_round_trip(tmp_path):
    pos = np.array([[0.1, -2.5, 123456789.0], [1e-36, 3.0e-5, -0.0], [1.0 / 3, 2.0 / 3, 1e10]], np.float32)
    nrm = np.array([[0, 0, 1], [0.6, 0.8, 0], [-1, 0, 0]], np.float32)
    ids = np.array([4, 0, 12345678901], np.int64)
    path = str(tmp_path / "s.pts")
    io.write_pts(path, pos, nrm, ids)
    lines = open(path).read().splitlines()
    clean = lambda x: 0.0 if abs(x) < 1e-35 else float(x)  # noqa: E731
    for i, line in enumerate(lines):
        want = " ".join("%g" % clean(x) for x in list(pos[i]) + list(nrm[i])) + " %d" % ids[i]
        assert line == want, (line, want)
    back = io.read_pts(path)
    assert np.array_equal(back, np.array([[float("%g" % clean(x)) for x in r] for r in pos], np.float32))
    io.write_pts(path, pos)
    assert open(path).read().splitlines()[0] == "0.1 -2.5 1.23457e+08"
    with open(path, "a") as fh:
        fh.write("\n# trailing comment\n\n")
    assert io.read_pts(path).shape == (3, 3)


def test_mesh_sample_options_without_device(tmp_path):
    """Argument handling of the MeshSample program that ends before any device work."""
    import subprocess
    app = os.path.join(os.path.dirname(GOLD), "..", "..", "thea_amd", "apps", "mesh_sample")
    if not os.path.exists(app):
        pytest.skip("mesh_sample is built by __graft_entry__.build()")
    r = subprocess.run([app], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "Usage" in r.stdout
    r = subprocess.run([app, "-v", "a.obj", "b.pts"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "not provided" in r.stderr
    r = subprocess.run([app, "-n-3", "a.obj", "b.pts"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "invalid number" in r.stderr
