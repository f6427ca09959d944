"""Guide tables of the area-light triangle search (nanogi_amd/csrc/ngi_guide.h).

The GPU samplers find the emitting triangle of an area light through a guide
table instead of a binary search over the whole CDF; the index must be the
one Distribution1D::SampleReuse takes (basic.hpp:469-475: upper_bound - 1).
tests/guide_check.cpp compiles the same header on the host and compares the
guided search with std::upper_bound on every CDF entry and bucket edge (and
their FP32/FP64 neighbours) plus random u, for the scene loader's CDFs.  The
mutant run lowers the guide entries by one and must be caught.
"""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _build(tmp_path):
    exe = tmp_path / "guide_check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "nanogi_amd", "csrc"),
                    os.path.join(ROOT, "tests", "guide_check.cpp"), "-o", str(exe)], check=True)
    return str(exe)


def test_guided_search_is_upper_bound(tmp_path):
    exe = _build(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok ") and int(r.stdout.split()[1]) > 100000


def test_a_shifted_guide_is_caught(tmp_path):
    exe = _build(tmp_path)
    r = subprocess.run([exe, "mutant"], capture_output=True, text=True)
    assert r.returncode == 0 and "mutant caught" in r.stdout, r.stdout + r.stderr
