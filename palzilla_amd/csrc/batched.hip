// Batched persistent solver: many scenes, chain lengths 1..3, one batch.
//
// The reference launches solveSpecularPath_K once per event descriptor and
// reads the path count back to the host before each launch (pk.h:1480-1586,
// :1502-1503); inside a launch a wave runs as many Newton iterations as its
// slowest lane.  Here:
//   * paths are partitioned by chain length on the device (stable, so each
//     length keeps its tuple-coherent order), and every count stays on the
//     device;
//   * per chain length one persistent kernel keeps every lane busy: lanes
//     whose chain converged, left the triangle or ran out of iterations take
//     the next (path, wavelength) item from a per-wave chunk of a global work
//     queue, so a wave never idles behind its slowest lane;
//   * the Jacobian of an iteration comes from one multi-tangent dual pass
//     (bit-identical to the reference's 2K single-tangent passes);
//   * converged chains queue up for a finish kernel: throughput with K+1
//     occlusion rays decides on the FP32 solution (the same decisions as the
//     plain kernel), then one FP64 Newton step refines the vertices that feed
//     dA/dw and the reported solution;
//   * launch shapes are chosen per chain length by an autotuner, and a whole
//     solve can be captured in a hipGraph (counters are reset by a kernel, no
//     memset nodes).
#include <algorithm>
#include <cstring>

#include "optics.hpp"
#include "sort.hpp"

namespace palz {

namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kMaxIterations = 32;          // main_gpu.cu:456
constexpr float kCostTolerance = 1.0e-10f;  // main_gpu.cu:456
constexpr int kWarmUp = 1;                  // main_gpu.cu:456
constexpr int kQueueStride = 6;             // floats per queued solution (K <= 3)

struct Counters {
    uint32_t ticket[3];
    uint32_t queue[3];
    unsigned long long stats[16];
};

__global__ void k_reset(Counters* c) {
    if (threadIdx.x < 3) {
        c->ticket[threadIdx.x] = 0u;
        c->queue[threadIdx.x] = 0u;
    }
    if (threadIdx.x < 16) c->stats[threadIdx.x] = 0ull;
}

// words[i] = chain length | i << 32
__global__ void __launch_bounds__(256) k_size_words(const palz_path* __restrict__ paths, uint32_t n,
                                                    const uint32_t* __restrict__ d_count, uint64_t* __restrict__ words) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t m = d_count ? min(n, *d_count) : n;
    if (i >= m) return;
    words[i] = (uint64_t)min(ev_size(paths[i].event), 7) | ((uint64_t)i << 32);
}

__global__ void __launch_bounds__(256) k_gather_sorted(const uint64_t* __restrict__ words,
                                                       const palz_path* __restrict__ paths, uint32_t n,
                                                       const uint32_t* __restrict__ d_count,
                                                       palz_path* __restrict__ sorted, uint32_t* __restrict__ order,
                                                       uint32_t* __restrict__ n_out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t m = d_count ? min(n, *d_count) : n;
    if (i == 0) *n_out = m;
    if (i >= m) return;
    const uint32_t src = (uint32_t)(words[i] >> 32);
    sorted[i] = paths[src];
    order[i] = src;
}

__global__ void k_empty(uint32_t* ranges) {
    if (threadIdx.x < 6) ranges[threadIdx.x] = 0u;
}

// ranges[k] = number of paths with chain length <= k (k = 0..4), by binary
// search of the length-sorted words.
__global__ void k_ranges(const uint64_t* __restrict__ words, const uint32_t* __restrict__ n_dev,
                         uint32_t* __restrict__ ranges) {
    const int k = threadIdx.x;
    if (k > 4) return;
    const uint32_t n = *n_dev;
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((uint32_t)(words[mid] & 7u) <= (uint32_t)k)
            lo = mid + 1;
        else
            hi = mid;
    }
    ranges[k] = lo;
}

__device__ __forceinline__ float item_eta(const palz_scene_desc* sc, uint32_t l) {
    const float lam = sc->lambdas[l];
    return sc->cauchy_a + sc->cauchy_b / (lam * lam);
}

__device__ __forceinline__ void write_unconverged(palz_path_result* results, uint32_t item, int iters,
                                                  const float* prm, int P) {
    if (!results) return;
    palz_path_result r;
    r.status = PALZ_PATH_NOT_CONVERGED;
    r.iterations = iters;
    r.contribution = 0.0f;
    r.footprint = 0.0f;
#pragma unroll
    for (int i = 0; i < 8; ++i) r.params[i] = i < P ? prm[i] : 0.0f;
    results[item] = r;
}

__device__ __forceinline__ void wave_add_stat(unsigned long long* dst, unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(dst, v);
}

// Persistent Newton iterations of every item of chain length K.
template <int K>
__global__ void __launch_bounds__(256)
    k_newton(const palz_scene_desc* __restrict__ scenes, uint32_t n_scenes, const palz_path* __restrict__ paths,
             const uint32_t* __restrict__ ranges, uint32_t L, uint32_t chunk, Counters* __restrict__ ctr,
             uint32_t* __restrict__ q_item, float* __restrict__ q_params, int* __restrict__ q_iters,
             palz_path_result* __restrict__ results) {
    constexpr int P = 2 * K;
    constexpr int R = 3 * K;
    const uint32_t first_path = ranges[K - 1];
    const uint32_t total = (ranges[K] - first_path) * L;
    const uint32_t item_base = first_path * L;
    const int lane = threadIdx.x & 63;
    const uint64_t lt = __lanemask_lt();
    uint32_t* ticket = &ctr->ticket[K - 1];
    uint32_t* q_count = &ctr->queue[K - 1];

    uint32_t item = kNone;
    int iter = 0;
    float edge_eps = 0.25f, eta = 0.0f;
    uint32_t event = 0u;
    v3 pb = mk(0.0f, 0.0f, 0.0f), pe = mk(0.0f, 0.0f, 0.0f);
    float prm[P];
    ChainTri tri[K];
    uint32_t c_next = 0u, c_end = 0u;  // this wave's window of the work queue
    bool exhausted = total == 0u;
    unsigned long long n_iter = 0ull, n_conv = 0ull, n_items = 0ull;

    for (;;) {
        // refill idle lanes, in lane order, from the wave's queue window
        const uint64_t idle = __ballot(item == kNone);
        if (idle != 0ull && !exhausted) {
            const uint32_t need = (uint32_t)__popcll(idle);
            const uint32_t rank = (uint32_t)__popcll(idle & lt);
            uint32_t served = 0u;
            while (served < need) {
                if (c_next >= c_end) {
                    uint32_t base = 0u;
                    if (lane == 0) base = atomicAdd(ticket, chunk);
                    base = (uint32_t)__shfl((int)base, 0, 64);
                    if (base >= total) {
                        exhausted = true;
                        break;
                    }
                    c_next = base;
                    c_end = min(base + chunk, total);
                }
                const uint32_t take = min(need - served, c_end - c_next);
                if (item == kNone && rank >= served && rank < served + take) {
                    item = item_base + c_next + (rank - served);
                    const uint32_t pi = item / L, l = item % L;
                    const palz_path path = paths[pi];
#pragma unroll
                    for (int i = 0; i < P; ++i) prm[i] = path.params[i];
                    ++n_items;
                    if (path.scene < n_scenes) {
                        const palz_scene_desc* sc = scenes + path.scene;
#pragma unroll
                        for (int k = 0; k < K; ++k) tri[k] = load_chain_tri(sc->tris, sc->attribs, path.tris[k]);
                        pe = mk(sc->points[path.pixel].p);
                        pb = mk(sc->p_light);
                        eta = item_eta(sc, l);
                        event = path.event;  // per lane (see k_solve_plain)
                        iter = 0;
                        edge_eps = 0.25f;
                    } else {
                        // a path naming a scene outside the batch is not solved
                        write_unconverged(results, item, 0, prm, P);
                        item = kNone;
                        atomicAdd(&ctr->stats[14], 1ull);
                    }
                }
                c_next += take;
                served += take;
            }
        }
        if (__ballot(item != kNone) == 0ull) {
            if (exhausted) break;  // queue drained and every lane done
            continue;              // every item of this refill was rejected up front
        }

        // one iteration of solveConstraints (pk.h:477-590) per active lane
        bool converged = false;
        if (item != kNone) {
            float A[R * P];
            float b[R];
            float cost = 0.0f;
            chain_residual<K, P, float, v3>(prm, 0, pb, pe, tri, eta, event, A, b, &cost);
            ++n_iter;
            if (cost < kCostTolerance) {
                converged = true;
            } else {
                float dp[P];
                solve_qr<R, P, float>(A, b, dp);
#pragma unroll
                for (int i = 0; i < P; ++i) prm[i] = prm[i] - dp[i];
                bool outside = false;
                if (kWarmUp <= iter) {
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        const float u = prm[2 * k], v = prm[2 * k + 1];
                        if (!(-edge_eps < u && -edge_eps < v && u + v < 1.0f + edge_eps)) outside = true;
                    }
                    edge_eps *= edge_eps;
                }
                if (outside) {
                    write_unconverged(results, item, iter, prm, P);
                    item = kNone;
                } else if (++iter == kMaxIterations) {
                    write_unconverged(results, item, kMaxIterations, prm, P);
                    item = kNone;
                }
            }
        }
        // converged chains join the finish queue (one atomic per wave)
        const uint64_t cm = __ballot(converged);
        if (cm != 0ull) {
            const int leader = __ffsll((unsigned long long)cm) - 1;
            uint32_t base = 0u;
            if (lane == leader) base = atomicAdd(q_count, (uint32_t)__popcll(cm));
            base = (uint32_t)__shfl((int)base, leader, 64);
            if (converged) {
                const uint32_t slot = base + (uint32_t)__popcll(cm & lt);
                q_item[slot] = item;
                q_iters[slot] = iter;
#pragma unroll
                for (int i = 0; i < P; ++i) q_params[(size_t)slot * kQueueStride + i] = prm[i];
                ++n_conv;
                item = kNone;
            }
        }
    }
    wave_add_stat(&ctr->stats[(K - 1) * 4 + 0], n_items);
    wave_add_stat(&ctr->stats[(K - 1) * 4 + 1], n_conv);
    wave_add_stat(&ctr->stats[(K - 1) * 4 + 3], n_iter);
}

// Throughput, FP64 refinement, dA/dw and irradiance of the converged chains.
template <int K>
__global__ void __launch_bounds__(256)
    k_finish(const palz_scene_desc* __restrict__ scenes, const palz_path* __restrict__ paths, uint32_t L,
             Counters* __restrict__ ctr, const uint32_t* __restrict__ q_item, const float* __restrict__ q_params,
             const int* __restrict__ q_iters, int refine, palz_path_result* __restrict__ results,
             double* __restrict__ refined) {
    constexpr int P = 2 * K;
    constexpr int R = 3 * K;
    const uint32_t n = ctr->queue[K - 1];
    unsigned long long n_contrib = 0ull;
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        const uint32_t item = q_item[s];
        const uint32_t pi = item / L, l = item % L;
        const palz_path path = paths[pi];
        const palz_scene_desc* sc = scenes + path.scene;
        ChainTri tri[K];
#pragma unroll
        for (int k = 0; k < K; ++k) tri[k] = load_chain_tri(sc->tris, sc->attribs, path.tris[k]);
        float prm[P];
#pragma unroll
        for (int i = 0; i < P; ++i) prm[i] = q_params[(size_t)s * kQueueStride + i];
        const palz_first_diffuse q = sc->points[path.pixel];
        const v3 pe = mk(q.p);
        const v3 pb = mk(sc->p_light);
        const float eta = item_eta(sc, l);
        const uint32_t event = path.event;
        const float thr =
            contributable_throughput<K>(prm, pb, pe, tri, event, sc->nodes, sc->tris, *sc->root, eta);
        float pf[P];
#pragma unroll
        for (int i = 0; i < P; ++i) pf[i] = prm[i];
        if (refine != PALZ_REFINE_NONE) {
            // one Newton step in FP64 from the FP32 solution (two tangents per pass)
            double p64[P];
#pragma unroll
            for (int i = 0; i < P; ++i) p64[i] = (double)prm[i];
            double A[R * P];
            double b[R];
            double cost = 0.0;
#pragma unroll
            for (int first = 0; first < P; first += 2)
                chain_residual<K, 2, double, d3>(p64, first, mkd(pb), mkd(pe), tri, (double)eta, event, A, b, &cost);
            double dp[P];
            solve_qr<R, P, double>(A, b, dp);
#pragma unroll
            for (int i = 0; i < P; ++i) {
                p64[i] = p64[i] - dp[i];
                pf[i] = (float)p64[i];
            }
            if (refined) {
#pragma unroll
                for (int i = 0; i < 8; ++i) refined[(size_t)item * 8 + i] = i < P ? p64[i] : 0.0;
            }
        }
        palz_path_result r;
        r.status = PALZ_PATH_CONVERGED;
        r.iterations = q_iters[s];
        r.contribution = 0.0f;
        r.footprint = 0.0f;
#pragma unroll
        for (int i = 0; i < 8; ++i) r.params[i] = i < P ? pf[i] : 0.0f;
        if (0.0f < thr) {
            // PALZ_REFINE_VERTICES keeps the irradiance on the FP32 chain (the
            // reference's arithmetic); PALZ_REFINE_ALL feeds the refined one.
            const float* pa = refine == PALZ_REFINE_ALL ? pf : prm;
            const float area = dadw<K>(pb, chain_vertex(tri[0], pa[0], pa[1]) - pb, pe, tri, event, eta);
            const v3 last = chain_vertex(tri[K - 1], pa[P - 2], pa[P - 1]);
            float c = thr * sc->light_intensity / area * hmax(dot(normalize(last - pe), mk(q.ng)), 0.0f);
            c = hmin(c, sc->radiance_clamp);
            atomicAdd(sc->image + (size_t)l * sc->n_points + path.pixel, c);
            r.status = PALZ_PATH_CONTRIBUTES;
            r.contribution = c;
            r.footprint = area;
            ++n_contrib;
        }
        if (results) results[item] = r;
    }
    wave_add_stat(&ctr->stats[(K - 1) * 4 + 2], n_contrib);
}

// Items of chain length 4 are not part of a batch: mark and count them.
__global__ void __launch_bounds__(256) k_skipped(const uint32_t* __restrict__ ranges, uint32_t L,
                                                 Counters* __restrict__ ctr, palz_path_result* __restrict__ results) {
    const uint32_t b = ranges[3] * L, e = ranges[4] * L;
    if (blockIdx.x == 0 && threadIdx.x == 0) ctr->stats[15] = (unsigned long long)(e - b);
    if (!results) return;
    for (uint32_t i = b + blockIdx.x * blockDim.x + threadIdx.x; i < e; i += gridDim.x * blockDim.x) {
        palz_path_result r = {};
        results[i] = r;
    }
}

template <int K>
void launch_newton(const palz_launch_config& c, int cus, hipStream_t st, const palz_scene_desc* scenes,
                   uint32_t n_scenes, const palz_path* paths, const uint32_t* ranges, uint32_t L, Counters* ctr,
                   uint32_t* qi, float* qp, int* qit, palz_path_result* results) {
    k_newton<K><<<c.blocks_per_cu * cus, c.block, 0, st>>>(scenes, n_scenes, paths, ranges, L, (uint32_t)c.chunk,
                                                          ctr, qi, qp, qit, results);
}
template <int K>
void launch_finish(const palz_launch_config& c, int cus, hipStream_t st, const palz_scene_desc* scenes,
                   const palz_path* paths, uint32_t L, Counters* ctr, const uint32_t* qi, const float* qp,
                   const int* qit, int refine, palz_path_result* results, double* refined) {
    k_finish<K><<<cus * 8, c.finish_block, 0, st>>>(scenes, paths, L, ctr, qi, qp, qit, refine, results, refined);
}

}  // namespace

}  // namespace palz

struct palz_batched {
    palz_path* paths = nullptr;
    uint32_t* order = nullptr;
    uint64_t* words = nullptr;
    uint64_t* words_tmp = nullptr;
    uint32_t* ranges = nullptr;  // [5] + [1] path count
    uint32_t cap_paths = 0;
    palz::Counters* ctr = nullptr;
    uint32_t* q_item = nullptr;
    float* q_params = nullptr;
    int* q_iters = nullptr;
    size_t cap_items = 0;
    palz_launch_config cfg[3] = {{256, 4, 256, 128}, {256, 2, 256, 128}, {128, 2, 256, 128}};
    int cus = 256;
    // hipGraph cache
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t capture = nullptr;  // graphs are captured here (never the null stream)
    struct Key {
        const void *scenes, *results, *refined;
        uint32_t n_scenes, L;
        int refine;
        palz_launch_config cfg[3];
    } key{};

    int ensure_paths(uint32_t n) {
        if (n <= cap_paths) return PALZ_OK;
        for (void* p : {(void*)paths, (void*)order, (void*)words, (void*)words_tmp})
            if (p) PALZ_HIP(hipFree(p));
        PALZ_HIP(hipMalloc(&paths, sizeof(palz_path) * n));
        PALZ_HIP(hipMalloc(&order, sizeof(uint32_t) * n));
        PALZ_HIP(hipMalloc(&words, sizeof(uint64_t) * n));
        PALZ_HIP(hipMalloc(&words_tmp, sizeof(uint64_t) * n));
        cap_paths = n;
        return PALZ_OK;
    }
    int ensure_items(size_t n) {
        if (n <= cap_items) return PALZ_OK;
        for (void* p : {(void*)q_item, (void*)q_params, (void*)q_iters})
            if (p) PALZ_HIP(hipFree(p));
        PALZ_HIP(hipMalloc(&q_item, sizeof(uint32_t) * n));
        PALZ_HIP(hipMalloc(&q_params, sizeof(float) * palz::kQueueStride * n));
        PALZ_HIP(hipMalloc(&q_iters, sizeof(int) * n));
        cap_items = n;
        return PALZ_OK;
    }
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~palz_batched() {
        drop_graph();
        if (capture) (void)hipStreamDestroy(capture);
        for (void* p : {(void*)paths, (void*)order, (void*)words, (void*)words_tmp, (void*)ranges, (void*)ctr,
                        (void*)q_item, (void*)q_params, (void*)q_iters})
            if (p) (void)hipFree(p);
    }

    int launch_all(const palz_scene_desc* scenes, uint32_t n_scenes, uint32_t L, int refine,
                   palz_path_result* results, double* refined, hipStream_t st) {
        using namespace palz;
        k_reset<<<1, 64, 0, st>>>(ctr);
        const palz_path* P = paths;
        launch_newton<1>(cfg[0], cus, st, scenes, n_scenes, P, ranges, L, ctr, q_item, q_params, q_iters, results);
        launch_finish<1>(cfg[0], cus, st, scenes, P, L, ctr, q_item, q_params, q_iters, refine, results, refined);
        launch_newton<2>(cfg[1], cus, st, scenes, n_scenes, P, ranges, L, ctr, q_item, q_params, q_iters, results);
        launch_finish<2>(cfg[1], cus, st, scenes, P, L, ctr, q_item, q_params, q_iters, refine, results, refined);
        launch_newton<3>(cfg[2], cus, st, scenes, n_scenes, P, ranges, L, ctr, q_item, q_params, q_iters, results);
        launch_finish<3>(cfg[2], cus, st, scenes, P, L, ctr, q_item, q_params, q_iters, refine, results, refined);
        k_skipped<<<64, 256, 0, st>>>(ranges, L, ctr, results);
        PALZ_LAUNCH_CHECK();
        return PALZ_OK;
    }
};

namespace {
int check_solve_args(palz_batched* b, const palz_scene_desc* d_scenes, uint32_t n_scenes, uint32_t L,
                     int refine) {
    using namespace palz;
    if (!b || !d_scenes || n_scenes == 0 || L == 0)
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_batched_solve: need a batch, scenes and n_lambdas >= 1");
    if (!b->paths) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_batched_solve: call palz_batched_prepare first");
    if (refine < PALZ_REFINE_NONE || refine > PALZ_REFINE_ALL)
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_batched_solve: refine must be a PALZ_REFINE_* value");
    if ((uint64_t)b->cap_paths * L >= (1ull << 32))
        return fail(PALZ_ERR_CAPACITY, "palz_batched_solve: more than 2^32 items");
    return b->ensure_items((size_t)b->cap_paths * L);
}
}  // namespace

extern "C" {

int palz_batched_create(palz_batched** out) {
    if (!out) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_batched_create: NULL");
    palz_batched* b = new palz_batched();
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
        b->cus = std::max(1, prop.multiProcessorCount);
    if (hipMalloc(&b->ranges, sizeof(uint32_t) * 8) != hipSuccess ||
        hipMalloc(&b->ctr, sizeof(palz::Counters)) != hipSuccess ||
        hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return palz::fail(PALZ_ERR_OUT_OF_MEMORY, "palz_batched_create: hipMalloc failed");
    }
    (void)hipMemset(b->ranges, 0, sizeof(uint32_t) * 8);
    (void)hipMemset(b->ctr, 0, sizeof(palz::Counters));
    *out = b;
    return PALZ_OK;
}

int palz_batched_destroy(palz_batched* b) {
    delete b;
    return PALZ_OK;
}

int palz_batched_prepare(palz_batched* b, const palz_path* d_paths, uint32_t n_paths, const uint32_t* d_count,
                         palz_sorter* sorter, void* stream) {
    using namespace palz;
    if (!b || !sorter || (!d_paths && n_paths)) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_batched_prepare: NULL");
    if (n_paths > kSortMaxN) return fail(PALZ_ERR_CAPACITY, "palz_batched_prepare: n_paths must be < 2^30");
    const uint32_t cap = std::max(n_paths, 1u);
    int rc = b->ensure_paths(cap);
    if (rc != PALZ_OK) return rc;
    rc = sorter->reserve(cap);
    if (rc != PALZ_OK) return rc;
    b->drop_graph();
    hipStream_t st = (hipStream_t)stream;
    const uint32_t nb = div_up(cap, 256u);
    if (n_paths == 0) {
        k_empty<<<1, 64, 0, st>>>(b->ranges);
        PALZ_LAUNCH_CHECK();
        return PALZ_OK;
    }
    k_size_words<<<nb, 256, 0, st>>>(d_paths, n_paths, d_count, b->words);
    const SortScratch& s = sorter->scratch;
    launch_sort_prepare(s, nullptr, st);
    launch_sort_count(s, b->words, n_paths, 0, 3, st, d_count);
    const int npass = launch_sort_passes(s, b->words, b->words_tmp, nullptr, nullptr, n_paths, 0, 3, st, nullptr,
                                         d_count);
    const uint64_t* sorted = (npass & 1) ? b->words_tmp : b->words;
    k_gather_sorted<<<nb, 256, 0, st>>>(sorted, d_paths, n_paths, d_count, b->paths, b->order, b->ranges + 5);
    k_ranges<<<1, 64, 0, st>>>(sorted, b->ranges + 5, b->ranges);
    PALZ_LAUNCH_CHECK();
    return PALZ_OK;
}

int palz_batched_order(palz_batched* b, const uint32_t** d_order, const palz_path** d_sorted_paths,
                       const uint32_t** d_ranges) {
    if (!b) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_batched_order: NULL");
    if (d_order) *d_order = b->order;
    if (d_sorted_paths) *d_sorted_paths = b->paths;
    if (d_ranges) *d_ranges = b->ranges;
    return PALZ_OK;
}

int palz_batched_set_config(palz_batched* b, int K, const palz_launch_config* cfg) {
    using namespace palz;
    if (!b || !cfg || K < 1 || K > 3) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_batched_set_config: bad argument");
    const palz_launch_config& c = *cfg;
    auto ok_block = [](int v) { return v == 64 || v == 128 || v == 256; };
    if (!ok_block(c.block) || !ok_block(c.finish_block) || c.blocks_per_cu < 1 || c.blocks_per_cu > 32 ||
        c.chunk < 1 || c.chunk > (1 << 16))
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_batched_set_config: block/finish_block in {64,128,256}, "
                                               "blocks_per_cu in 1..32, chunk in 1..65536");
    b->cfg[K - 1] = c;
    b->drop_graph();
    return PALZ_OK;
}

int palz_batched_get_config(palz_batched* b, int K, palz_launch_config* cfg) {
    if (!b || !cfg || K < 1 || K > 3) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_batched_get_config: bad argument");
    *cfg = b->cfg[K - 1];
    return PALZ_OK;
}

int palz_batched_solve(palz_batched* b, const palz_scene_desc* d_scenes, uint32_t n_scenes, uint32_t n_lambdas,
                       int refine, palz_path_result* d_results, double* d_refined, void* stream) {
    int rc = check_solve_args(b, d_scenes, n_scenes, n_lambdas, refine);
    if (rc != PALZ_OK) return rc;
    return b->launch_all(d_scenes, n_scenes, n_lambdas, refine, d_results, d_refined, (hipStream_t)stream);
}

int palz_batched_solve_graph(palz_batched* b, const palz_scene_desc* d_scenes, uint32_t n_scenes,
                             uint32_t n_lambdas, int refine, palz_path_result* d_results, double* d_refined,
                             void* stream) {
    int rc = check_solve_args(b, d_scenes, n_scenes, n_lambdas, refine);
    if (rc != PALZ_OK) return rc;
    hipStream_t st = (hipStream_t)stream;
    palz_batched::Key k{d_scenes, d_results, d_refined, n_scenes, n_lambdas, refine, {}};
    std::copy(b->cfg, b->cfg + 3, k.cfg);
    const bool same = b->exec && std::memcmp(&k, &b->key, sizeof(k)) == 0;
    if (!same) {
        b->drop_graph();
        PALZ_HIP(hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal));
        rc = b->launch_all(d_scenes, n_scenes, n_lambdas, refine, d_results, d_refined, b->capture);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->capture, &g);
        if (rc != PALZ_OK) return rc;
        PALZ_HIP(e);
        b->graph = g;
        PALZ_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
        b->key = k;
    }
    PALZ_HIP(hipGraphLaunch(b->exec, st));
    return PALZ_OK;
}

int palz_batched_stats(palz_batched* b, uint64_t* stats16, void* stream) {
    if (!b || !stats16) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_batched_stats: NULL");
    PALZ_HIP(hipStreamSynchronize((hipStream_t)stream));
    palz::Counters h;
    PALZ_HIP(hipMemcpy(&h, b->ctr, sizeof(h), hipMemcpyDeviceToHost));
    for (int i = 0; i < 16; ++i) stats16[i] = h.stats[i];
    return PALZ_OK;
}

int palz_batched_autotune(palz_batched* b, const palz_scene_desc* d_scenes, uint32_t n_scenes, uint32_t n_lambdas,
                          int refine, int32_t* report, void* stream) {
    using namespace palz;
    int rc = check_solve_args(b, d_scenes, n_scenes, n_lambdas, refine);
    if (rc != PALZ_OK) return rc;
    hipStream_t st = (hipStream_t)stream;
    uint32_t ranges[6];
    PALZ_HIP(hipStreamSynchronize(st));
    PALZ_HIP(hipMemcpy(ranges, b->ranges, sizeof(ranges), hipMemcpyDeviceToHost));
    hipEvent_t e0, e1;
    PALZ_HIP(hipEventCreate(&e0));
    PALZ_HIP(hipEventCreate(&e1));
    const int blocks[3] = {64, 128, 256};
    const int per_cu[4] = {1, 2, 4, 8};
    const int chunks[3] = {64, 256, 1024};
    for (int K = 1; K <= 3; ++K) {
        palz_launch_config best = b->cfg[K - 1];
        float best_ms = 0.0f, best_fin = 0.0f;
        if (ranges[K] > ranges[K - 1]) {
            best_ms = 1e30f;
            auto time_newton = [&](const palz_launch_config& c) -> float {
                k_reset<<<1, 64, 0, st>>>(b->ctr);
                (void)hipEventRecord(e0, st);
                if (K == 1) launch_newton<1>(c, b->cus, st, d_scenes, n_scenes, b->paths, b->ranges, n_lambdas, b->ctr, b->q_item, b->q_params, b->q_iters, nullptr);
                if (K == 2) launch_newton<2>(c, b->cus, st, d_scenes, n_scenes, b->paths, b->ranges, n_lambdas, b->ctr, b->q_item, b->q_params, b->q_iters, nullptr);
                if (K == 3) launch_newton<3>(c, b->cus, st, d_scenes, n_scenes, b->paths, b->ranges, n_lambdas, b->ctr, b->q_item, b->q_params, b->q_iters, nullptr);
                (void)hipEventRecord(e1, st);
                (void)hipEventSynchronize(e1);
                float ms = 0.0f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                return ms;
            };
            for (int blk : blocks)
                for (int pc : per_cu) {
                    if (blk * pc > 2048) continue;  // at most 32 waves resident per CU
                    for (int ch : chunks) {
                        palz_launch_config c = best;
                        c.block = blk;
                        c.blocks_per_cu = pc;
                        c.chunk = ch;
                        const float ms = time_newton(c);
                        if (ms < best_ms) {
                            best_ms = ms;
                            best = c;
                        }
                    }
                }
            // finish kernel block size, on the queue of the best Newton shape
            (void)time_newton(best);
            best_fin = 1e30f;
            for (int fb : blocks) {
                palz_launch_config c = best;
                c.finish_block = fb;
                (void)hipEventRecord(e0, st);
                if (K == 1) launch_finish<1>(c, b->cus, st, d_scenes, b->paths, n_lambdas, b->ctr, b->q_item, b->q_params, b->q_iters, refine, nullptr, nullptr);
                if (K == 2) launch_finish<2>(c, b->cus, st, d_scenes, b->paths, n_lambdas, b->ctr, b->q_item, b->q_params, b->q_iters, refine, nullptr, nullptr);
                if (K == 3) launch_finish<3>(c, b->cus, st, d_scenes, b->paths, n_lambdas, b->ctr, b->q_item, b->q_params, b->q_iters, refine, nullptr, nullptr);
                (void)hipEventRecord(e1, st);
                (void)hipEventSynchronize(e1);
                float ms = 0.0f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                if (ms < best_fin) {
                    best_fin = ms;
                    best.finish_block = fb;
                }
            }
            b->cfg[K - 1] = best;
        }
        if (report) {
            int32_t* r = report + (K - 1) * 6;
            r[0] = best.block;
            r[1] = best.blocks_per_cu;
            r[2] = best.chunk;
            r[3] = best.finish_block;
            r[4] = (int32_t)(best_ms * 1000.0f);
            r[5] = (int32_t)(best_fin * 1000.0f);
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->drop_graph();
    PALZ_LAUNCH_CHECK();
    return PALZ_OK;
}

}  // extern "C"
