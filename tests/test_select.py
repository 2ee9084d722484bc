"""The device restatement of std::nth_element (thea_amd/csrc/select.hpp) leaves exactly the
permutation the host's libstdc++ leaves; Thea's shape diameter reduction depends on it."""
import pathlib
import subprocess

ROOT = pathlib.Path(__file__).resolve().parents[1]


def test_nth_element_matches_libstdcxx(tmp_path):
    exe = tmp_path / "select_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", str(ROOT / "thea_amd" / "csrc"),
                    str(ROOT / "tests" / "select_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), "2000000"], check=True, capture_output=True, text=True).stdout.split()
    trials, bad, heap = (int(x) for x in out)
    assert trials == 2000000 and bad == 0
    assert heap > 0  # the depth-limited heap-select branch was exercised
