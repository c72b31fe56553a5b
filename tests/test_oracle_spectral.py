"""The spectral-resolve oracle pinned by what the colorimetry fixes: the fitted
CIE 2015 10-degree functions integrate to the reference's normalising constant
and to equal areas (equal-energy white), the sRGB matrix maps D65 white to
(1, 1, 1), the transfer function is continuous and fixes 0 and 1, packRGBA's
known answers, and linearity of the resolve."""
import numpy as np
import pytest

from oracle import spectral as S

GRID = np.linspace(360.0, 830.0, 47001)


def test_cmf_y_integral_matches_reference_constant():
    area = np.trapezoid(S.cmf_y(GRID).astype(np.float64), GRID)
    assert abs(area / float(S.INTEGRAL_OF_CMF_Y_IN_NM) - 1.0) < 1e-3


@pytest.mark.parametrize("f", [S.cmf_x, S.cmf_z])
def test_cmfs_have_equal_areas(f):
    # equal-energy white has x = y = z: the three functions enclose equal areas
    ay = np.trapezoid(S.cmf_y(GRID).astype(np.float64), GRID)
    assert abs(np.trapezoid(f(GRID).astype(np.float64), GRID) / ay - 1.0) < 2e-3


def test_cmf_peaks():
    assert abs(GRID[np.argmax(S.cmf_y(GRID))] - 557.0) < 3.0   # the 10-degree V(lambda) peak
    assert abs(GRID[np.argmax(S.cmf_z(GRID))] - 446.0) < 3.0
    assert np.all(S.cmf_y(GRID) >= 0.0) and np.all(S.cmf_z(GRID) >= 0.0)


def test_srgb_matrix_maps_d65_to_white():
    assert np.allclose(S.xyz2srgblinear([0.95047, 1.0, 1.08883]), 1.0, atol=2e-4)


def test_oetf_fixed_points_and_continuity():
    assert S.srgb_oetf(0.0) == 0.0 and abs(S.srgb_oetf(1.0) - 1.0) < 1e-6
    lo, hi = S.srgb_oetf(np.float32(0.003130)), S.srgb_oetf(np.nextafter(np.float32(0.003130), np.float32(1)))
    assert abs(hi - lo) < 1e-5
    assert S.srgb_oetf(-1.0) == 0.0


def test_pack_known_answers():
    assert int(S.pack_rgba([0.5, 0.5, 0.5])) == 0xFF808080
    assert int(S.pack_rgba([1.2, -0.1, 0.25])) == 0xFF4000FF
    assert int(S.pack_rgba([0.0, 0.0, 1.0])) == 0xFFFF0000


def test_resolve_is_linear_in_the_images():
    rng = np.random.default_rng(1)
    L, n = 16, 257
    img = rng.random((L, n), dtype=np.float32)
    lam = np.linspace(380, 720, L, dtype=np.float32)
    w = np.full(L, 340.0 / 15.0, dtype=np.float32)
    R = rng.random((n, 3), dtype=np.float32)
    a1, _ = S.resolve(img, lam, w, R, 1.0)
    a2, _ = S.resolve(img * np.float32(2.0), lam, w, R, 1.0)
    assert np.allclose(a2[:, :3], 2.0 * a1[:, :3], rtol=1e-6)


def test_flat_spectrum_resolves_to_illuminant_e():
    # a flat unit spectrum over the visible range, trapezoid weights and albedo
    # pi (cancelling the diffuse 1/pi): XYZ = (1, 1, 1), so linear sRGB = the
    # matrix's row sums (1.2048, 0.9484, 0.9087)
    lam = np.linspace(360.0, 830.0, 4701, dtype=np.float32)
    w = np.full(lam.size, (830.0 - 360.0) / (lam.size - 1), dtype=np.float32)
    w[0] *= 0.5
    w[-1] *= 0.5
    acc, _ = S.resolve(np.ones((lam.size, 1), np.float32), lam, w, np.full((1, 3), np.pi, np.float32), 1.0)
    assert np.allclose(acc[0, :3], S.M_XYZ_TO_SRGB.sum(axis=1), rtol=2e-3)
