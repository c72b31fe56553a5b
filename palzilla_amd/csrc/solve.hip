// Specular-chain solvers.
//
// Plain: solveSpecularPath_K (main_gpu.cu:418-497) — one lane per (path,
// wavelength), templated on the chain length, the reference's 2K dual passes
// per Newton iteration, throughput with K+1 occlusion rays, dA/dw and a float
// atomic into the receiver image.
#include <algorithm>

#include "optics.hpp"

namespace palz {

namespace {

constexpr int kSolveThreads = 256;
constexpr int kMaxIterations = 32;         // main_gpu.cu:456
constexpr float kCostTolerance = 1.0e-10f;  // main_gpu.cu:456
constexpr int kWarmUp = 1;                 // main_gpu.cu:456

__device__ __forceinline__ float cauchy(const palz_scene_desc& sc, float nm) {
    return sc.cauchy_a + sc.cauchy_b / (nm * nm);
}

// Throughput, footprint and irradiance of a converged chain (main_gpu.cu:458-471).
template <int K>
__device__ __forceinline__ float finish_path(const palz_scene_desc& sc, uint32_t root, const ChainTri* tri,
                                             const float* params, uint32_t event, float eta, v3 p_end, v3 n_end,
                                             bool* contributes, float* footprint) {
    const v3 p_light = mk(sc.p_light);
    const float thr = contributable_throughput<K>(params, p_light, p_end, tri, event, sc.nodes, sc.tris, root, eta);
    *contributes = false;
    if (!(0.0f < thr)) return 0.0f;
    *contributes = true;
    const float area = dadw<K>(p_light, chain_vertex(tri[0], params[0], params[1]) - p_light, p_end, tri, event, eta);
    *footprint = area;
    const v3 last = chain_vertex(tri[K - 1], params[2 * (K - 1)], params[2 * (K - 1) + 1]);
    float contrib = thr * sc.light_intensity / area * hmax(dot(normalize(last - p_end), n_end), 0.0f);
    return hmin(contrib, sc.radiance_clamp);
}

template <int K>
__global__ void __launch_bounds__(kSolveThreads)
    k_solve_plain(palz_scene_desc sc, const palz_path* __restrict__ paths, uint32_t n_paths,
                  const uint32_t* __restrict__ d_count, palz_path_result* __restrict__ results) {
    const uint32_t L = sc.n_lambdas;
    const uint64_t item = (uint64_t)blockIdx.x * kSolveThreads + threadIdx.x;
    const uint32_t np = d_count ? min(n_paths, *d_count) : n_paths;
    if (item >= (uint64_t)np * L) return;
    const uint32_t pi = (uint32_t)(item / L);
    const uint32_t l = (uint32_t)(item % L);
    const palz_path path = paths[pi];
    palz_path_result res;
    res.status = PALZ_PATH_NOT_CONVERGED;
    res.iterations = 0;
    res.contribution = 0.0f;
    res.footprint = 0.0f;
    // The chain code takes the event descriptor from the path, a per-lane
    // value, and the kernel checks only its length.  Made wave-uniform (the
    // launch's descriptor), the R/T selects become scalar branches inside the
    // divergent finishing region, and hipcc (ROCm 7.2) then computes dA/dw
    // wrongly for whole waves; tests/test_gpu_caustics.py guards this.
    const uint32_t event = path.event;
    if (ev_size(event) == K) {
        const palz_first_diffuse q = sc.points[path.pixel];
        const v3 p_end = mk(q.p);
        ChainTri tri[K];
#pragma unroll
        for (int k = 0; k < K; ++k) tri[k] = load_chain_tri(sc.tris, sc.attribs, path.tris[k]);
        float params[2 * K];
#pragma unroll
        for (int i = 0; i < 2 * K; ++i) params[i] = path.params[i];
        const float eta = cauchy(sc, sc.lambdas[l]);
        int iters = 0;
        const bool converged = solve_constraints<K, 1, float, v3>(params, mk(sc.p_light), p_end, tri, eta, event,
                                                                 kMaxIterations, kCostTolerance, kWarmUp, &iters);
        res.iterations = iters;
        if (converged) {
            res.status = PALZ_PATH_CONVERGED;
            bool contributes = false;
            const float c = finish_path<K>(sc, *sc.root, tri, params, event, eta, p_end, mk(q.ng), &contributes,
                                           &res.footprint);
            if (contributes) {
                res.status = PALZ_PATH_CONTRIBUTES;
                res.contribution = c;
                add_contribution(sc, (size_t)l * sc.n_points + path.pixel, c);
            }
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) res.params[i] = i < 2 * K ? params[i] : 0.0f;
    } else {
#pragma unroll
        for (int i = 0; i < 8; ++i) res.params[i] = 0.0f;
    }
    if (results) results[item] = res;
}

}  // namespace

}  // namespace palz

extern "C" {

int palz_solve_specular(int K, uint32_t event, const palz_scene_desc* scene, const palz_path* d_paths, uint32_t n_paths,
                        const uint32_t* d_count, palz_path_result* d_results, void* stream) {
    using namespace palz;
    if (!scene || (!d_paths && n_paths)) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_solve_specular: NULL argument");
    if (K < 1 || K > 4 || ev_size(event) != K)
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_solve_specular: K must be 1..4 and match the event size");
    if (scene->n_lambdas == 0 || !scene->lambdas || !scene->image || !scene->points)
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_solve_specular: scene needs points, wavelengths and an image");
    if (scene->image_fixed && !(scene->radiance_clamp > 0.0f && scene->radiance_clamp < 2147483648.0f))
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_solve_specular: a fixed-point image needs radiance_clamp in (0, 2^31)");
    const uint64_t items = (uint64_t)n_paths * scene->n_lambdas;
    if (items == 0) return PALZ_OK;
    if (items >= (1ull << 31)) return fail(PALZ_ERR_CAPACITY, "palz_solve_specular: more than 2^31 items");
    const uint32_t blocks = (uint32_t)((items + kSolveThreads - 1) / kSolveThreads);
    hipStream_t st = (hipStream_t)stream;
    // `event` names the descriptor the path list was flattened for; the
    // kernels read each path's own descriptor (see k_solve_plain).
    switch (K) {
        case 1: k_solve_plain<1><<<blocks, kSolveThreads, 0, st>>>(*scene, d_paths, n_paths, d_count, d_results); break;
        case 2: k_solve_plain<2><<<blocks, kSolveThreads, 0, st>>>(*scene, d_paths, n_paths, d_count, d_results); break;
        case 3: k_solve_plain<3><<<blocks, kSolveThreads, 0, st>>>(*scene, d_paths, n_paths, d_count, d_results); break;
        default: k_solve_plain<4><<<blocks, kSolveThreads, 0, st>>>(*scene, d_paths, n_paths, d_count, d_results); break;
    }
    PALZ_LAUNCH_CHECK();
    return PALZ_OK;
}

}  // extern "C"
