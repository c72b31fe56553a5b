"""Affine enclosures of the admissibility residual (palzilla_amd/csrc/affine.hpp).

The host checker tests/affine/enclosure_driver.cpp evaluates, for random
chains and random parameter rectangles, the affine residual of every oriented
vertex, and samples the same residual in float64 inside the rectangles and
the refractive-index range: every sample must lie in the affine ranges.  The
affine ranges must also be clearly narrower than the interval ranges of the
same rectangles (the point of the affine test).  A copy of the header with
the product's cross-term bound dropped must fail the same check, so the
sampling is dense enough to see a missing error term.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "palzilla_amd", "csrc")
DRIVER = os.path.join(HERE, "affine", "enclosure_driver.cpp")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CROSS_TERM = "e = add_ru(e, mul_ru(rad(x), rad(y)));"


def _build(exe, include_dir):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc is not installed")
    cmd = [HIPCC, "-x", "hip", "--cuda-host-only", "-std=c++17", "-O0", "-ffp-contract=off",
           "-I", include_dir, "-I", os.path.join(ROOT, "include"), "-I", CSRC, DRIVER, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]


def _run(exe, cases=3000):
    r = subprocess.run([exe, str(cases)], capture_output=True, text=True, timeout=120)
    cases, samples, violations, ratio = r.stdout.strip().splitlines()[-1].split()
    return int(cases), int(samples), int(violations), float(ratio), r.stdout


def test_affine_residual_encloses_sampled_residuals(tmp_path):
    exe = str(tmp_path / "enclosure")
    _build(exe, CSRC)
    cases, samples, violations, ratio, out = _run(exe)
    print(out)
    assert violations == 0, out
    assert cases > 5000 and samples > 300_000
    assert ratio < 0.8, f"affine ranges are {ratio:.3f} of the interval widths"


def test_dropped_cross_term_is_caught(tmp_path):
    src = open(os.path.join(CSRC, "affine.hpp")).read()
    assert src.count(CROSS_TERM) == 1
    mutated = tmp_path / "inc"
    mutated.mkdir()
    (mutated / "affine.hpp").write_text(src.replace(CROSS_TERM, "", 1))
    exe = str(tmp_path / "enclosure_mutated")
    _build(exe, str(mutated))
    _, _, violations, _, out = _run(exe)
    assert violations > 0, out
