"""GPU test of the irradiance adjoint (palz_adjoint_irradiance) through a thin
torch.autograd.Function: d(L2 image loss)/d(heights) of a refractive slab
heightfield against FP64 central finite differences (h = 1e-4) of the
oracle's FP64 restatement, on 64 random height parameters."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
import torch

import oracle
from oracle import adjoint as AD
from palzilla_amd import caustics as CA
from palzilla_amd import scenes
from palzilla_amd.device import DeviceBuffer

pytestmark = pytest.mark.gpu

CELLS, WIDTH, PHOTONS = 16, 48, 1 << 14


class Render:
    """Forward state of one height vector: scene, batched solve, image."""

    def __init__(self, heights: np.ndarray, recompute: bool = False, cells: int = CELLS, width: int = WIDTH,
                 photons: int = PHOTONS):
        self.cells = cells
        self.s = scenes.config5(heights, cells=cells, width=width, photons=photons)
        self.g = CA.GpuScene(self.s)
        self.g.trace()
        self.g.build_table()
        _, n = self.g.lookup(0, capacity=1)
        self.d_paths, self.n = self.g.lookup(0, capacity=max(n, 1))
        self.L = self.s.lambdas.shape[0]
        self.bs = CA.BatchedSolver(self.g.sorter)
        self.bs.prepare(self.d_paths, self.n)
        self.d_scenes = CA.upload_scene_descs([self.g.desc()])
        self.d_res = DeviceBuffer(max(self.n * self.L, 1) * CA.PATH_RESULT.itemsize)
        self.g.clear_image()
        self.bs.solve(self.d_scenes, 1, self.L, CA.REFINE_VERTICES, self.d_res)
        self.image = self.g.image().reshape(-1).astype(np.float64)
        tv, nb, coef = scenes.height_param_tables(cells, self.s.n_tris)
        self.tv, self.nb = tv, nb
        self.hp = CA.HeightParams(tv, nb, coef)
        self.recompute = recompute

    def backward(self, weight: np.ndarray) -> np.ndarray:
        d_w = DeviceBuffer.from_array(np.ascontiguousarray(weight, dtype=np.float32))
        d_g = DeviceBuffer(self.hp.n_params * 8)
        d_g.zero()
        _, p_sorted, _ = self.bs.pointers()
        CA.adjoint_irradiance(self.g, self.hp, p_sorted, self.n, None if self.recompute else self.d_res.ptr,
                              d_w, d_g)
        return d_g.download(np.float64, self.hp.n_params)


class Irradiance(torch.autograd.Function):
    @staticmethod
    def forward(ctx, heights, state):
        ctx.state = state
        return torch.from_numpy(state.image.copy())

    @staticmethod
    def backward(ctx, grad_image):
        return torch.from_numpy(ctx.state.backward(grad_image.detach().numpy())), None


@pytest.fixture(scope="module")
def problem():
    h = scenes.slab_heights(CELLS, 5)
    target = Render(scenes.slab_heights(CELLS, 6)).image
    r = Render(h)
    return h, target, r


def _autograd(r):
    heights = torch.tensor(r.h, dtype=torch.float64, requires_grad=True)
    img = Irradiance.apply(heights, r)
    loss = ((img - torch.from_numpy(r.target)) ** 2).sum()
    loss.backward()
    return heights.grad.numpy()


def test_adjoint_matches_fp64_finite_differences(problem):
    h, target, r = problem
    r.recompute, r.h, r.target = False, h, target
    grad = _autograd(r)
    # FP64 finite differences on the forward solve's item set
    sorted_paths = r.bs.sorted_paths()
    res = CA.download_results(r.d_res, r.n * r.L)
    prob = AD.Problem(r.s, sorted_paths, res, r.tv, r.nb, CELLS, CA.LIGHT_INTENSITY, CA.RADIANCE_CLAMP)
    # the FP64 image of the item set is the FP32 image to FP32 accuracy
    assert np.allclose(prob.image(h), r.image, rtol=2e-3, atol=1e-3)
    rng = np.random.default_rng(7)
    live = np.nonzero(np.abs(grad) > 0)[0]
    assert live.size > 64
    pick = rng.choice(live, 64, replace=False)
    ad = grad[pick]
    fd4 = np.array([prob.loss_fd(h, target, int(j), 1e-4) for j in pick])
    fd6 = np.array([prob.loss_fd(h, target, int(j), 1e-6) for j in pick])
    scale = np.max(np.abs(fd6))
    rel4 = np.abs(ad - fd4) / (np.abs(fd4) + 1e-3 * scale)
    rel6 = np.abs(ad - fd6) / (np.abs(fd6) + 1e-3 * scale)
    print(f"64 params, |dL/dh| up to {scale:.4g}: vs FD h=1e-4 median {np.median(rel4):.2e}, "
          f"{np.sum(rel4 < 1e-2)}/64 within 1e-2; vs FD h=1e-6 median {np.median(rel6):.2e}, "
          f"{np.sum(rel6 < 1e-3)}/64 within 1e-3")
    # h = 1e-4 carries the truncation error of a strongly curved caustic image
    # and straddles clamp kinks for a few parameters; h = 1e-6 converges
    assert np.median(rel4) < 1e-3 and np.sum(rel4 < 1e-2) >= 60
    assert np.median(rel6) < 1e-5 and np.sum(rel6 < 1e-3) >= 62


def test_adjoint_recomputed_equals_stored(problem):
    h, target, r = problem
    r.h, r.target = h, target
    r.recompute = False
    stored = _autograd(r)
    r.recompute = True
    recomputed = _autograd(r)
    r.recompute = False
    assert np.allclose(recomputed, stored, rtol=1e-6, atol=1e-9 * np.abs(stored).max())


def test_adjoint_of_zero_weight_is_zero(problem):
    h, target, r = problem
    r.recompute = False
    g = r.backward(np.zeros_like(r.image))
    assert np.all(g == 0)


@pytest.mark.timeout(900)
def test_adjoint_full_size_matches_fp64_finite_differences():
    # the adjoint workload at its real size: 128^2 cells (16,641 heights),
    # 512^2 receiver, 2^21 photons, target from a second seed.  The loss is
    # taken around the GPU image, as the adjoint's weights 2 (I - T) are; each
    # central difference moves only the items that depend on the height, with
    # their FP64 contributions.  64 random live heights, differences at
    # h = 1e-4 (the configured step), 1e-5, 1e-6 and 1e-7.
    import time

    cells, width, photons = 128, 512, 1 << 21
    h = scenes.slab_heights(cells, 5)
    target = Render(scenes.slab_heights(cells, 6), cells=cells, width=width, photons=photons).image
    r = Render(h, cells=cells, width=width, photons=photons)
    r.h, r.target = h, target
    print(f"C5: {r.n} paths, {r.hp.n_params} heights", flush=True)
    grad = _autograd(r)
    t0 = time.perf_counter()
    res = CA.download_results(r.d_res, r.n * r.L)
    sorted_paths = r.bs.sorted_paths()
    prob = AD.Problem(r.s, sorted_paths, res, r.tv, r.nb, cells, CA.LIGHT_INTENSITY, CA.RADIANCE_CLAMP)
    # the FP64 contributions of a 200k-item sample against the GPU's
    rng = np.random.default_rng(7)
    for st in prob.sets:
        sel = rng.choice(st.item.size, min(st.item.size, 200_000), replace=False)
        c64 = st.contributions(h, sel)
        # a set's item numbers count the items of its own chain length's paths
        own = np.nonzero((sorted_paths["event"] >> 24) == st.K)[0]
        glob = own[st.item[sel] // r.L] * r.L + st.item[sel] % r.L
        c32 = res["contribution"][glob].astype(np.float64)
        # an FP64 re-solve that ends in total internal reflection has no
        # finite value: counted apart, the rest compared
        fin = np.isfinite(c64)
        c64, c32 = c64[fin], c32[fin]
        close = np.isclose(c64, c32, rtol=1e-3, atol=1e-6)
        far = ~close
        rel = np.abs(c64 - c32) / np.maximum(np.abs(c64), 1e-30)
        share = np.abs(c64[far]).sum() / np.abs(c64).sum()
        sums = abs(c64.sum() - c32.sum()) / abs(c64.sum())
        print(f"K={st.K}: {sel.size} sampled items, {int((~fin).sum())} without a finite FP64 value, "
              f"{int(far.sum())} beyond rtol 1e-3 (largest {rel.max():.2e}, {share:.2e} of the sample's "
              f"irradiance); sums differ by {sums:.2e}", flush=True)
        # K <= 2: the FP32 roots are the FP64 roots to round-off, up to a few
        # items near a caustic fold.  K = 3 (T-R-T through the slab's bottom)
        # is ill-conditioned: FP32 Newton stops at an FP64 residual of ~5e-6,
        # and the FP64 re-solve moves some chains' parameters by ~5e-4.
        if st.K <= 2:
            assert fin.all() and far.sum() <= 1e-3 * sel.size and sums <= 1e-4
    live = np.nonzero(np.abs(grad) > 0)[0]
    assert live.size > 1000
    base = r.image
    print(f"finite: image {np.isfinite(base).all()}, target {np.isfinite(target).all()}, "
          f"gradient {np.isfinite(grad).all()}", flush=True)
    # The FP64 central differences of 64 random live heights at h = 1e-4 ..
    # 1e-7.  Rule: a height whose stencil puts one of its items' FP64 chains at
    # a square root of zero (grazing refraction or total internal reflection
    # within the step: the chain's value is not finite) has no central
    # difference there; it is left out, counted, and another height drawn.
    steps = (1e-4, 1e-5, 1e-6, 1e-7)
    for st in prob.sets:
        st.depends(0, prob.nb)  # build the dependency index before the threads share it
    order = rng.permutation(live)

    def ladder(j):
        return prob.loss_fd_ladder(h, target, int(j), steps, base)

    picked, fds, excluded = [], [], 0
    with ThreadPoolExecutor(8) as ex:
        at = 0
        while len(picked) < 64:
            batch = order[at:at + 64 - len(picked)]
            at += batch.size
            assert batch.size, "too few heights with finite central differences"
            for j, (fd, finite) in zip(batch, ex.map(ladder, batch)):
                if finite:
                    picked.append(int(j))
                    fds.append(fd)
                else:
                    excluded += 1
            print(f"{len(picked)}/64 heights, {excluded} left out, {time.perf_counter() - t0:.1f} s", flush=True)
    fd = np.array(fds)
    assert np.all(np.isfinite(fd)), "a kept height has a non-finite central difference"
    ad = grad[np.array(picked)]
    scale = np.max(np.abs(fd[:, -1]))
    # A step has resolved a height when the difference no longer moves: within
    # 1e-3 of the next step's (a tenth).  The truncation error of a resolved
    # difference is then about that change, so the adjoint must agree within
    # 2.5e-3 at every step on every height that step resolves.
    ok_counts = []
    for i, hstep in enumerate(steps[:3]):
        resolved = np.abs(fd[:, i] - fd[:, i + 1]) <= 1e-3 * np.abs(fd[:, i + 1])
        rel = np.abs(ad - fd[:, i]) / (np.abs(fd[:, i]) + 1e-6 * scale)
        print(f"h={hstep:g}: {int(resolved.sum())}/64 heights resolved; adjoint vs FD on them: median "
              f"{np.median(rel[resolved]) if resolved.any() else float('nan'):.2e}, max "
              f"{rel[resolved].max() if resolved.any() else float('nan'):.2e}; over all 64: median "
              f"{np.median(rel):.2e}", flush=True)
        assert np.all(rel[resolved] <= 2.5e-3), f"h={hstep:g}: adjoint differs from a resolved difference"
        ok_counts.append(int(resolved.sum()))
    # At 128^2 cells the caustic image's curvature leaves h = 1e-4 unresolved for
    # most heights (measured on 85 heights: 27 resolved at 1e-4, 55 at 1e-5, 73
    # at 1e-6); the heights no step resolves sit within 1e-6 of a radiance-clamp
    # kink or a total-internal-reflection cut.
    assert ok_counts[0] >= 10 and ok_counts[1] >= 28 and ok_counts[2] >= 44, ok_counts
