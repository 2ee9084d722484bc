"""Invariants of the per-triangle bounding sphere behind the closest-point pruning bound
(thea_amd/csrc/tri_math.hpp make_trirec), checked on the host: the sphere holds every vertex, a
record marked tight has a unit normal whose plane holds the vertices and the centre to 2^-20 of the
coordinates, and the others keep the box centre."""
import pathlib
import subprocess

ROOT = pathlib.Path(__file__).resolve().parents[1]


def test_trirec_sphere_invariants(tmp_path):
    exe = tmp_path / "trirec_check"
    subprocess.run(["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-ffp-contract=off", "--offload-arch=gfx950",
                    "-I", str(ROOT / "thea_amd" / "csrc"), str(ROOT / "tests" / "trirec_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe), "200000"], check=True, capture_output=True, text=True).stdout.split()
    trials, bad, tight, loose = (int(x) for x in out)
    assert trials == 200000 and bad == 0
    assert tight > 0 and loose > 0  # both kinds of record occur
