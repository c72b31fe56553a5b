"""The occlusion walks' box test, pinned on the CPU (host build of optics.hpp,
tests/slab_check.cpp): slabs() folding with fminf/fmaxf decides every box as
the reference's compare-and-select folds do (minimum_lbvh.h:82-96), and
slabs_ordered() with entry and exit planes picked by the direction's signs
decides as slabs() does, on random rays and boxes with zero, signed-zero,
denormal and huge direction components, flat and point boxes and origins on
box planes.  A planted mistake (entry and exit swapped on one axis) must show."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def slab_check(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not found")
    exe = str(tmp_path_factory.mktemp("slab") / "slab_check")
    cmd = [HIPCC, "-x", "hip", "--offload-host-only", "-std=c++17", "-O2", "-ffp-contract=off",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "palzilla_amd", "csrc"),
           os.path.join(HERE, "slab_check.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def _run(exe, *args):
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    f = r.stdout.split()
    return {f[i]: int(f[i + 1]) for i in range(0, len(f), 2)}


def test_box_decisions_equal_the_reference(slab_check):
    out = _run(slab_check, "2000000")
    assert out == {"trials": 2000000, "fold": 0, "ordered": 0}


def test_planted_plane_swap_is_caught(slab_check):
    out = _run(slab_check, "200000", "swap")
    assert out["fold"] == 0 and out["ordered"] > 1000
