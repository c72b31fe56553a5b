// Per-tuple path-cuts culling (pk.h:2101-2207 with interval.h's operations,
// in directed rounding; see interval.hpp), optionally followed by the same
// test in affine arithmetic (affine.hpp), as a kernel and as C entry points.
#include <algorithm>

#include "affine.hpp"
#include "cull.hpp"
#include "interval.hpp"

namespace palz {

namespace {

// The wavelengths' refractive indices, as the solvers compute them.
__device__ __forceinline__ iv::I eta_range(const palz_scene_desc& sc) {
    float lo = iv::kInf, hi = -iv::kInf;
    for (uint32_t l = 0; l < sc.n_lambdas; ++l) {
        const float lam = sc.lambdas[l];
        const float eta = sc.cauchy_a + sc.cauchy_b / (lam * lam);
        lo = pmin(eta, lo);
        hi = pmax(eta, hi);
    }
    return {lo, hi};
}

template <int K>
__device__ __forceinline__ bool admissible_path(const palz_scene_desc& sc, const palz_path& path, v3 pb, v3 pe,
                                                iv::I eta) {
    ChainTri tri[K];
#pragma unroll
    for (int k = 0; k < K; ++k) tri[k] = load_chain_tri(sc.tris, sc.attribs, path.tris[k]);
    return iv::chain_admissible<K>(pb, pe, tri, path.event, eta);
}

template <int K>
__device__ __forceinline__ bool admissible_path_affine(const palz_scene_desc& sc, const palz_path& path, v3 pb,
                                                       v3 pe, iv::I eta) {
    ChainTri tri[K];
#pragma unroll
    for (int k = 0; k < K; ++k) tri[k] = load_chain_tri(sc.tris, sc.attribs, path.tris[k]);
    return aa::chain_admissible<K>(pb, pe, tri, path.event, eta);
}

__global__ void __launch_bounds__(256) k_cull_paths(const palz_scene_desc* __restrict__ scenes, uint32_t n_scenes,
                                                    const palz_path* __restrict__ paths, uint32_t n,
                                                    const uint32_t* __restrict__ n_dev,
                                                    uint8_t* __restrict__ admissible) {
    const uint32_t m = n_dev ? min(n, *n_dev) : n;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < m; p += gridDim.x * blockDim.x) {
        const palz_path path = paths[p];
        const int K = ev_size(path.event);
        bool ok = true;
        if (K >= 1 && K <= 3 && path.scene < n_scenes) {
            const palz_scene_desc& sc = scenes[path.scene];
            const iv::I eta = eta_range(sc);
            const v3 pb = mk(sc.p_light);
            const v3 pe = mk(sc.points[path.pixel].p);
            if (K == 1) ok = admissible_path<1>(sc, path, pb, pe, eta);
            if (K == 2) ok = admissible_path<2>(sc, path, pb, pe, eta);
            if (K == 3) ok = admissible_path<3>(sc, path, pb, pe, eta);
        }
        admissible[p] = ok ? 1 : 0;
    }
}

// The affine test on the paths the interval test kept (flag 1): clears the
// flag of those it excludes.
__global__ void __launch_bounds__(256) k_affine_paths(const palz_scene_desc* __restrict__ scenes, uint32_t n_scenes,
                                                      const palz_path* __restrict__ paths, uint32_t n,
                                                      const uint32_t* __restrict__ n_dev,
                                                      uint8_t* __restrict__ admissible) {
    const uint32_t m = n_dev ? min(n, *n_dev) : n;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < m; p += gridDim.x * blockDim.x) {
        if (!admissible[p]) continue;
        const palz_path path = paths[p];
        const int K = ev_size(path.event);
        if (K < 1 || K > 3 || path.scene >= n_scenes) continue;
        const palz_scene_desc& sc = scenes[path.scene];
        const iv::I eta = eta_range(sc);
        const v3 pb = mk(sc.p_light);
        const v3 pe = mk(sc.points[path.pixel].p);
        bool ok = true;
        if (K == 1) ok = admissible_path_affine<1>(sc, path, pb, pe, eta);
        if (K == 2) ok = admissible_path_affine<2>(sc, path, pb, pe, eta);
        if (K == 3) ok = admissible_path_affine<3>(sc, path, pb, pe, eta);
        if (!ok) admissible[p] = 0;
    }
}

}  // namespace

void launch_cull_paths(const palz_scene_desc* scenes, uint32_t n_scenes, const palz_path* paths, uint32_t n,
                       const uint32_t* n_dev, uint8_t* admissible, hipStream_t stream, int mode) {
    if (n == 0) return;
    const uint32_t blocks = std::min(div_up(n, 256u), 65536u);
    k_cull_paths<<<blocks, 256, 0, stream>>>(scenes, n_scenes, paths, n, n_dev, admissible);
    if (mode == 2) k_affine_paths<<<blocks, 256, 0, stream>>>(scenes, n_scenes, paths, n, n_dev, admissible);
}

}  // namespace palz

extern "C" {

int palz_interval_cull(const palz_scene_desc* d_scenes, uint32_t n_scenes, const palz_path* d_paths,
                       uint32_t n_paths, const uint32_t* d_count, uint8_t* d_admissible, void* stream) {
    using namespace palz;
    if (!d_scenes || n_scenes == 0 || (n_paths && (!d_paths || !d_admissible)))
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_interval_cull: NULL argument");
    launch_cull_paths(d_scenes, n_scenes, d_paths, n_paths, d_count, d_admissible, (hipStream_t)stream, 1);
    PALZ_LAUNCH_CHECK();
    return PALZ_OK;
}

int palz_affine_cull(const palz_scene_desc* d_scenes, uint32_t n_scenes, const palz_path* d_paths, uint32_t n_paths,
                     const uint32_t* d_count, uint8_t* d_admissible, void* stream) {
    using namespace palz;
    if (!d_scenes || n_scenes == 0 || (n_paths && (!d_paths || !d_admissible)))
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_affine_cull: NULL argument");
    launch_cull_paths(d_scenes, n_scenes, d_paths, n_paths, d_count, d_admissible, (hipStream_t)stream, 2);
    PALZ_LAUNCH_CHECK();
    return PALZ_OK;
}

}  // extern "C"
