"""The oracle against the reference's own code (CPU only).

tests/golden/ref_vectors.npz holds the outputs of the reference's sobol.h,
saka.h and interval.h, compiled unmodified from the reference tree by
oracle/ref_build.py and driven by tests/golden/make_ref_vectors.py.  Where
oracle/_ref/libref.so is built (the reference tree is present) the same
comparisons also run on fresh random inputs.

  * Sobol (sobol.h:17-129): the oracle's shuffled_scrambled_sobol_2d equals
    the reference's bit for bit at the photon tracer's seed patterns
    (main_gpu.cu:84, :88, :143) and at random seeds.  The GPU photon records
    equal the oracle's, so this ties every photon direction to the reference.
  * saka duals (saka.h:192-208): reflection and refraction_norm_free, value
    and tangent, bit for bit (the zero vector on total internal reflection).
  * Intervals (interval.h:46-713): the reference rounds to nearest; the
    oracle rounds outward.  Each oracle bound lies on the outer side of the
    reference's (containment) and within a few ulps of it.
"""
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_vectors.npz")


@pytest.fixture(scope="module")
def gold():
    with np.load(GOLD) as z:
        return {k: z[k] for k in z.files}


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_sobol_equals_the_reference_bit_for_bit(gold):
    got = np.array([oracle.sobol2d(*map(int, row)) for row in gold["sobol_in"]], np.float32)
    assert np.array_equal(_bits(got), _bits(gold["sobol_out"]))
    # the photon tracer's three seed patterns are all present
    p2 = gold["sobol_in"][:, 3]
    assert (p2 == 789).sum() > 100 and (p2 == 178).sum() > 100


def _sobol_lk(index, p0, p1, p2, mults):
    """sobol.h's scrambled Sobol with the four Laine-Karras multipliers as
    arguments: a planted-mistake variant for the discrimination check."""
    M = 0xFFFFFFFF

    def rev(v):
        return int(f"{v:032b}"[::-1], 2)

    def lk(x, seed):
        x = (x + seed) & M
        for m in mults:
            x ^= (x * m) & M
        return x

    def nus(x, seed):
        return rev(lk(rev(x), seed))

    def P(v):
        v ^= (v << 16) & M
        v ^= ((v & 0x00FF00FF) << 8) & M
        v ^= ((v & 0x0F0F0F0F) << 4) & M
        v ^= ((v & 0x33333333) << 2) & M
        v ^= ((v & 0x55555555) << 1) & M
        return v

    def pcg3d(x, y, z):
        x = (x * 1664525 + 1013904223) & M
        y = (y * 1664525 + 1013904223) & M
        z = (z * 1664525 + 1013904223) & M
        x = (x + y * z) & M
        y = (y + z * x) & M
        z = (z + x * y) & M
        x ^= x >> 16
        y ^= y >> 16
        z ^= z >> 16
        x = (x + y * z) & M
        y = (y + z * x) & M
        z = (z + x * y) & M
        return x, y, z

    p0, p1, p2 = pcg3d(p0, p1, p2)
    v = lk(rev(index), p0)
    f = lambda u: np.float32(u >> 8) / np.float32(1 << 24)
    return f(nus(v, p1)), f(nus(P(v), p2))


def test_sobol_fixture_rejects_a_wrong_scramble_constant(gold):
    """The fixture separates sobol.h from a variant with one even multiplier
    changed (which keeps the (0, 2)-net property the oracle was pinned by
    before) and from swapped seeds."""
    right = (0x6C50B47C, 0xB82F1E52, 0xC7AFE638, 0x8D22F6E6)
    wrong = (0x6C50B47C, 0xB82F1E52, 0xC7AFE638, 0x8D22F6E4)
    rows = gold["sobol_in"][:64].tolist()
    want = gold["sobol_out"][:64]
    ok = np.array([_sobol_lk(*map(int, r), right) for r in rows], np.float32)
    bad = np.array([_sobol_lk(*map(int, r), wrong) for r in rows], np.float32)
    swapped = np.array([oracle.sobol2d(int(a), int(b), int(d), int(c)) for a, b, c, d in rows], np.float32)
    assert np.array_equal(_bits(ok), _bits(want))
    assert (np.any(_bits(bad) != _bits(want), axis=1)).mean() > 0.9
    assert (np.any(_bits(swapped) != _bits(want), axis=1)).mean() > 0.9


@pytest.mark.parametrize("op,name", [(0, "refl"), (1, "refr")])
def test_saka_duals_equal_the_reference_bit_for_bit(gold, op, name):
    wi, n, eta = gold["saka_wi"], gold["saka_n"], gold["saka_eta"]
    want = gold[f"saka_{name}_out"]
    got = np.stack([oracle.dual_optics(op, wi[i], n[i], float(eta[i])) for i in range(len(wi))])
    assert np.array_equal(_bits(got), _bits(want))
    if op == 1:  # the inputs include total internal reflection
        tir = np.all(want == 0.0, axis=1)
        assert 20 < tir.sum() < len(want) - 20


def _ulp(x):
    x = np.abs(np.asarray(x, np.float32))
    return np.spacing(np.maximum(x, np.float32(1e-30)))


@pytest.mark.parametrize("op", range(9))
def test_interval_ops_enclose_the_reference(gold, op):
    a, b, want = gold[f"iv{op}_a"], gold[f"iv{op}_b"], gold[f"iv{op}_out"]
    got = np.stack([oracle.iv_ops(op, a[i], b[i]) for i in range(len(a))]).reshape(len(a), -1, 2)
    want = want.reshape(len(a), -1, 2)
    lo, hi = got[..., 0], got[..., 1]
    rlo, rhi = want[..., 0], want[..., 1]
    if op == 3:
        # the reference divides as a * [1/b.l, 1/b.u]: two roundings, so its
        # bounds may sit up to two ulps either side of the correctly rounded
        # quotient bounds the oracle produces
        assert np.all(np.abs(lo - rlo) <= 2 * _ulp(rlo)) and np.all(np.abs(hi - rhi) <= 2 * _ulp(rhi))
        return
    assert np.all(lo <= rlo) and np.all(hi >= rhi), "the oracle's bounds must lie outside round-to-nearest ones"
    # and not far outside: one ulp per rounding of the operation's terms
    terms = {0: 1, 1: 1, 2: 1, 4: 1, 5: 1, 6: 6, 7: 3, 8: 5}[op]
    mag = np.maximum(np.abs(rlo), np.abs(rhi)) + (np.max(np.abs(a), axis=1, keepdims=True) ** 2 if op >= 6 else 0)
    assert np.all(rlo - lo <= terms * _ulp(mag)) and np.all(hi - rhi <= terms * _ulp(mag))


def test_against_the_built_reference_on_fresh_inputs():
    """Larger random comparisons where oracle/_ref/libref.so is built."""
    ref = oracle.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref/libref.so is not built (no reference tree here)")
    rng = np.random.default_rng(7)
    rows = rng.integers(0, 1 << 32, size=(20000, 4), dtype=np.uint64).astype(np.uint32)
    out = np.zeros(2, np.float32)
    for r in rows.tolist():
        ref.ref_sobol2d(*r, oracle._ptr(out))
        got = np.array(oracle.sobol2d(*r), np.float32)
        assert np.array_equal(_bits(got), _bits(out)), r
    res = np.zeros(6, np.float32)
    for _ in range(4000):
        wi = rng.normal(size=6).astype(np.float32)
        n = rng.normal(size=6).astype(np.float32)
        eta = np.float32(rng.uniform(0.5, 2.0))
        for op in (0, 1):
            ref.ref_saka_optics(op, oracle._ptr(wi), oracle._ptr(n), float(eta), oracle._ptr(res))
            assert np.array_equal(_bits(oracle.dual_optics(op, wi, n, float(eta))), _bits(res))
