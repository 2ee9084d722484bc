"""The MeshSample program over the C ABI (thea_amd/apps/mesh_sample.cpp, built by the package
build) on Thea's own model files: every output line equals what the Python binding computes for the
same mesh, seed and mode, in MeshSample's text format."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "thea_amd", "apps", "mesh_sample")
GOLD = os.path.join(ROOT, "tests", "golden", "models")


def run(*args):
    assert os.path.exists(APP), "thea_amd/apps/mesh_sample is built by __graft_entry__.build()"
    r = subprocess.run([APP] + [str(a) for a in args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r


def fmt(x):
    return "%g" % (0.0 if abs(x) < 1e-35 else float(x))


def expected_lines(pos, nrm, ids=None):
    out = []
    for i in range(len(pos)):
        s = " ".join(fmt(x) for x in list(pos[i]) + list(nrm[i]))
        out.append(s + ("" if ids is None else " %d" % ids[i]))
    return out


@pytest.mark.parametrize("model,n", [("cube.obj", 2000), ("sphere.obj", 5000)])
def test_area_sampling_with_ids(tmp_path, model, n):
    import thea_amd
    from thea_amd import io
    print(f"case mesh_sample area {model}", flush=True)
    out = tmp_path / "s.pts"
    run(f"-n{n}", "-id", "-seed7", os.path.join(GOLD, model), out)
    v, t, faces, _ = io.read_mesh(os.path.join(GOLD, model))
    pos, nrm, tri = thea_amd.MeshSampler(thea_amd.Mesh(v, t)).sampleEvenlyByArea(n, seed=7, stream_id=0)
    assert open(out).read().splitlines() == expected_lines(pos, nrm, faces[tri])


def test_separation_and_presampled(tmp_path):
    import thea_amd
    from thea_amd import io
    print("case mesh_sample separation", flush=True)
    sphere = os.path.join(GOLD, "sphere.obj")
    out = tmp_path / "sep.pts"
    run("-n300", "-s", "-seed3", sphere, out)
    v, t, _, _ = io.read_mesh(sphere)
    pos, nrm, _ = thea_amd.MeshSampler(thea_amd.Mesh(v, t)).sampleEvenlyBySeparation(300, seed=3, stream_id=0)
    assert open(out).read().splitlines() == expected_lines(pos, nrm)
    # -s -i: subsample an existing point file, keeping the chosen lines verbatim
    dense = tmp_path / "dense.pts"
    run("-n3000", "-seed5", sphere, dense)
    sub = tmp_path / "sub.pts"
    run("-n200", "-s", "-i", dense, sphere, sub)
    lines = open(dense).read().splitlines()
    sel = thea_amd.furthestPointSampling(io.read_pts(dense), 200)
    assert open(sub).read().splitlines() == [lines[i] for i in sel]
