"""GPU parity of the spectral resolve (palz_spectral_resolve) with the numpy
oracle: accumulators to FP32 round-off (the GPU's expf/powf against numpy's),
packed pixels equal up to one step where a channel sits on a rounding edge."""
import numpy as np
import pytest

from oracle import spectral as S
from palzilla_amd import caustics as CA
from palzilla_amd import scenes
from palzilla_amd.device import DeviceBuffer

pytestmark = pytest.mark.gpu


def _points(R):
    pts = np.zeros(R.shape[0], dtype=scenes.FIRST_DIFFUSE)
    pts["R"] = R
    return pts


def _compare(acc, pix, want_acc, want_pix):
    assert np.allclose(acc, want_acc, rtol=2e-5, atol=1e-7)
    a = pix.view(np.uint8).reshape(-1, 4).astype(np.int32)
    b = want_pix.view(np.uint8).reshape(-1, 4).astype(np.int32)
    assert np.abs(a - b).max() <= 1
    assert (a != b).any(axis=1).mean() < 1e-3


@pytest.mark.parametrize("L,n,frames", [(1, 1, 1.0), (16, 4099, 1.0), (64, 20000, 7.0)])
def test_resolve_random_images(L, n, frames):
    rng = np.random.default_rng(L * n)
    img = (rng.random((L, n), dtype=np.float32) * 3).astype(np.float32)
    lam = np.sort(rng.uniform(380, 720, L)).astype(np.float32)
    w = rng.uniform(0.5, 30.0, L).astype(np.float32)
    R = rng.random((n, 3), dtype=np.float32)
    acc, pix = CA.spectral_resolve(DeviceBuffer.from_array(img), DeviceBuffer.from_array(lam),
                                   DeviceBuffer.from_array(w), DeviceBuffer.from_array(_points(R)), n, L, frames)
    want_acc, want_pix = S.resolve(img, lam, w, R, frames)
    _compare(acc.download(np.float32, 4 * n).reshape(n, 4), pix.download(np.uint32, n), want_acc, want_pix)


def test_resolve_of_a_solved_image():
    # the irradiance images of a small spectral slab, solved by the batched solver
    s = scenes.config2(cells=32, width=96, photons=1 << 14, n_lambdas=4)
    g = CA.GpuScene(s)
    g.trace()
    g.build_table()
    d_p, n = g.lookup(0, capacity=1 << 20)
    bs = CA.BatchedSolver(g.sorter)
    bs.prepare(d_p, n)
    g.clear_image()
    bs.solve(CA.upload_scene_descs([g.desc()]), 1, s.lambdas.shape[0], CA.REFINE_VERTICES, None)
    img = g.image().reshape(s.lambdas.shape[0], -1)
    assert img.max() > 0
    L, m = img.shape
    w = np.full(L, 340.0 / max(L - 1, 1), dtype=np.float32)
    acc, pix = CA.spectral_resolve(g.d_image, g.d_lambdas, DeviceBuffer.from_array(w), g.d_points, m, L)
    want_acc, want_pix = S.resolve(img, s.lambdas.astype(np.float32), w, s.points["R"], 1.0)
    _compare(acc.download(np.float32, 4 * m).reshape(m, 4), pix.download(np.uint32, m), want_acc, want_pix)
