// Batched persistent solver: many scenes, chain lengths 1..3, one batch.
//
// The reference launches solveSpecularPath_K once per event descriptor and
// reads the path count back to the host before each launch (pk.h:1480-1586,
// :1502-1503); inside a launch a wave runs as many Newton iterations as its
// slowest lane.  Here:
//   * paths are sorted on the device by chain length, scene and first
//     triangle (stable, so a tuple's paths stay together), so that waves see
//     one scene's neighbouring chains, and every count stays on the device;
//   * per chain length one persistent kernel keeps every lane busy: lanes
//     whose chain converged, left the triangle or ran out of iterations take
//     the next (path, wavelength) item from a per-wave chunk of a global work
//     queue, so a wave never idles behind its slowest lane;
//   * the Jacobian of an iteration comes from one multi-tangent dual pass
//     (bit-identical to the reference's 2K single-tangent passes);
//   * converged chains are compacted in item order (so the finish kernel's
//     waves keep the tuple-coherent order of the path list for their
//     occlusion rays) and finished: throughput with K+1 occlusion rays decides
//     on the FP32 solution (the same decisions as the plain kernel), then one
//     FP64 Newton step refines the converged vertices;
//   * launch shapes are chosen per chain length by an autotuner, and a whole
//     solve can be captured in a hipGraph (counters are reset by a kernel, no
//     memset nodes).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "cull.hpp"
#include "optics.hpp"
#include "prims.hpp"
#include "sort.hpp"

namespace palz {

namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kMaxIterations = 32;          // main_gpu.cu:456
constexpr float kCostTolerance = 1.0e-10f;  // main_gpu.cu:456
constexpr int kWarmUp = 1;                  // main_gpu.cu:456
constexpr int kQueueStride = 6;             // floats per converged solution (K <= 3)

constexpr uint32_t kXcds = 8;  // MI355X: workgroups go round-robin over 8 XCDs, each with its own L2

struct Counters {
    uint32_t ticket[3];
    uint32_t n_items;    // items of every chain length (scan length)
    uint32_t converged;  // scan total
    uint32_t pad_;
    unsigned long long stats[16];
    uint32_t occlude[3][kXcds];  // per chain length and XCD: list positions handed out
    unsigned long long conv_paths[3];  // per chain length: paths with at least one converged item
};

__global__ void k_reset(Counters* c, const uint32_t* __restrict__ ranges, uint32_t L) {
    if (threadIdx.x < 3) c->ticket[threadIdx.x] = 0u;
    if (threadIdx.x == 0) c->n_items = ranges[4] * L;
    if (threadIdx.x < 16) c->stats[threadIdx.x] = 0ull;
    if (threadIdx.x < 3) c->conv_paths[threadIdx.x] = 0ull;
}

__global__ void k_reset_bounds(uint32_t* ranges) {
    if (threadIdx.x == 0) {
        ranges[6] = 0u;
        ranges[7] = 0u;
    }
}

// The largest scene index and first-triangle index over the paths (ranges[6],
// [7]): a grid-stride loop reduced per block, one atomic per block and word
// (atomics from every wave on the same two words serialise at the L2).
__global__ void __launch_bounds__(256) k_key_bounds(const palz_path* __restrict__ paths, uint32_t n,
                                                    const uint32_t* __restrict__ d_count, uint32_t* __restrict__ ranges) {
    __shared__ uint32_t s_max[2][256 / 64];
    const uint32_t m = d_count ? min(n, *d_count) : n;
    uint32_t s = 0u, t = 0u;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        s = max(s, paths[i].scene);
        t = max(t, (uint32_t)paths[i].tris[0]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        s = max(s, (uint32_t)__shfl_xor((int)s, o, 64));
        t = max(t, (uint32_t)__shfl_xor((int)t, o, 64));
    }
    const uint32_t wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_max[0][wave] = s;
        s_max[1][wave] = t;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        uint32_t v = 0u;
        for (uint32_t w = 0; w < blockDim.x / 64u; ++w) v = max(v, s_max[threadIdx.x][w]);
        atomicMax(ranges + 6 + threadIdx.x, v);
    }
}

__device__ __forceinline__ uint32_t bits_for(uint32_t v) { return v ? 32u - (uint32_t)__clz((int)v) : 0u; }

// words[i] = key | i << 32, key = chain length (bits 29-31) | scene | first
// triangle (the low 29 bits, the triangle index shifted down to fit)
__global__ void __launch_bounds__(256) k_size_words(const palz_path* __restrict__ paths, uint32_t n,
                                                    const uint32_t* __restrict__ d_count,
                                                    const uint32_t* __restrict__ ranges, uint64_t* __restrict__ words) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t m = d_count ? min(n, *d_count) : n;
    if (i >= m) return;
    const palz_path p = paths[i];
    const uint32_t sb = min(bits_for(ranges[6]), 29u);
    const uint32_t avail = 29u - sb;
    const uint32_t tb = bits_for(ranges[7]);
    const uint32_t shift = tb > avail ? tb - avail : 0u;
    const uint32_t tri = shift >= 32u ? 0u : ((uint32_t)p.tris[0] >> shift);
    const uint32_t scene = p.scene & ((1u << sb) - 1u);
    const uint32_t low = avail ? ((scene << avail) | (tri & ((1u << avail) - 1u))) : scene;
    const uint32_t key = (uint32_t)min(ev_size(p.event), 7) << 29 | low;
    words[i] = (uint64_t)key | ((uint64_t)i << 32);
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
        if (((uint32_t)words[mid] >> 29) <= (uint32_t)k)
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
                                                  const float* prm, int P,
                                                  uint32_t status = PALZ_PATH_NOT_CONVERGED) {
    if (!results) return;
    palz_path_result r;
    r.status = status;
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

// Persistent Newton iterations of every item of chain length K, run by the
// calling wave until chain length K's queue is drained.
template <int K>
__device__ __forceinline__ void newton_persistent(const palz_scene_desc* __restrict__ scenes, uint32_t n_scenes,
                                                  const palz_path* __restrict__ paths,
                                                  const uint32_t* __restrict__ ranges, uint32_t L, uint32_t chunk,
                                                  Counters* __restrict__ ctr, uint32_t* __restrict__ flags,
                                                  float* __restrict__ conv_params, int* __restrict__ conv_iters,
                                                  const uint8_t* __restrict__ admissible,
                                                  palz_path_result* __restrict__ results) {
    constexpr int P = 2 * K;
    constexpr int R = 3 * K;
    const uint32_t first_path = ranges[K - 1];
    const uint32_t total = (ranges[K] - first_path) * L;
    const uint32_t item_base = first_path * L;
    const int lane = threadIdx.x & 63;
    const uint64_t lt = __lanemask_lt();
    uint32_t* ticket = &ctr->ticket[K - 1];

    uint32_t item = kNone;
    int iter = 0;
    float edge_eps = 0.25f, eta = 0.0f;
    uint32_t event = 0u;
    v3 pb = mk(0.0f, 0.0f, 0.0f), pe = mk(0.0f, 0.0f, 0.0f);
    float prm[P];
    ChainTri tri[K];
    uint32_t c_next = 0u, c_end = 0u;  // this wave's window of the work queue
    bool exhausted = total == 0u;
    unsigned long long n_iter = 0ull, n_conv = 0ull, n_items = 0ull, n_culled = 0ull;

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
                    if (admissible && !admissible[pi]) {
                        // no point of the tuple's boxes can satisfy the chain
                        write_unconverged(results, item, 0, prm, P, PALZ_PATH_CULLED);
                        flags[item] = 0u;
                        item = kNone;
                        ++n_culled;
                    } else if (path.scene < n_scenes) {
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
                        flags[item] = 0u;
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
                    flags[item] = 0u;
                    item = kNone;
                } else if (++iter == kMaxIterations) {
                    write_unconverged(results, item, kMaxIterations, prm, P);
                    flags[item] = 0u;
                    item = kNone;
                }
            }
        }
        // converged chains are stored per item for the compaction
        if (converged) {
            flags[item] = 1u;
            conv_iters[item] = iter;
#pragma unroll
            for (int i = 0; i < P; ++i) conv_params[(size_t)item * kQueueStride + i] = prm[i];
            ++n_conv;
            item = kNone;
        }
    }
    wave_add_stat(&ctr->stats[(K - 1) * 4 + 0], n_items);
    wave_add_stat(&ctr->stats[(K - 1) * 4 + 1], n_conv);
    wave_add_stat(&ctr->stats[(K - 1) * 4 + 3], n_iter);
    wave_add_stat(&ctr->stats[12], n_culled);
}

// The Newton kernel of chain length K.  For K = 2 the register budget is held
// at 168 so that three waves share a SIMD (K = 3 would spill).
template <int K>
__global__ void __launch_bounds__(256, K == 2 ? 3 : 1)
    k_newton(const palz_scene_desc* __restrict__ scenes, uint32_t n_scenes, const palz_path* __restrict__ paths,
             const uint32_t* __restrict__ ranges, uint32_t L, uint32_t chunk, Counters* __restrict__ ctr,
             uint32_t* __restrict__ flags, float* __restrict__ conv_params, int* __restrict__ conv_iters,
             const uint8_t* __restrict__ admissible, palz_path_result* __restrict__ results) {
    newton_persistent<K>(scenes, n_scenes, paths, ranges, L, chunk, ctr, flags, conv_params, conv_iters, admissible,
                         results);
}

// One persistent kernel over every chain length: each wave drains the K = 3
// queue, then K = 2's, then K = 1's (the longest chains first, so that the
// tail is made of the cheapest items).  A wave that finds one queue empty
// moves on while others still work on it, so no chain length waits for
// another's last wave.  One register allocation serves every body: budgeted
// for kWaves waves per SIMD, it is K = 3's 244 VGPRs at 2 waves (K = 2 then
// runs at 2 waves too) or 168 at 3 waves (K = 2's own budget; K = 3 spills).
template <int kWaves>
__global__ void __launch_bounds__(256, kWaves)
    k_newton_mixed(const palz_scene_desc* __restrict__ scenes, uint32_t n_scenes,
                   const palz_path* __restrict__ paths, const uint32_t* __restrict__ ranges, uint32_t L,
                   uint32_t chunk1, uint32_t chunk2, uint32_t chunk3, Counters* __restrict__ ctr,
                   uint32_t* __restrict__ flags, float* __restrict__ conv_params, int* __restrict__ conv_iters,
                   const uint8_t* __restrict__ admissible, palz_path_result* __restrict__ results) {
    newton_persistent<3>(scenes, n_scenes, paths, ranges, L, chunk3, ctr, flags, conv_params, conv_iters, admissible,
                         results);
    newton_persistent<2>(scenes, n_scenes, paths, ranges, L, chunk2, ctr, flags, conv_params, conv_iters, admissible,
                         results);
    newton_persistent<1>(scenes, n_scenes, paths, ranges, L, chunk1, ctr, flags, conv_params, conv_iters, admissible,
                         results);
}

// Throughput, FP64 refinement, dA/dw and irradiance of the converged chains.
__device__ __forceinline__ uint32_t compact_pos(const uint32_t* excl, uint32_t j, uint32_t n, uint32_t total) {
    return j < n ? excl[j] : total;
}

// One refinement step from the FP32 solution, mixed-precision iterative
// refinement: the residual evaluated in FP64 at the FP32 parameters (the
// precision lives there), the Jacobian of the FP32 iteration at the same
// point (one multi-tangent FP32 pass, widened to double) and the
// least-squares correction solved in FP64.  The FP32 Jacobian's relative
// error (~1e-7) scales a correction that is itself ~1e-5 of the parameters.
template <int K>
__device__ __forceinline__ void refine_step(const float* prm, v3 pb, v3 pe, const ChainTri* tri, float eta,
                                            uint32_t event, double* p64) {
    constexpr int P = 2 * K;
    constexpr int R = 3 * K;
#pragma unroll
    for (int i = 0; i < P; ++i) p64[i] = (double)prm[i];
    // the FP64 residual first: only it is live through the FP32 pass
    double A[R * P];
    double b[R];
    double cost = 0.0;
    chain_residual<K, 0, double, d3>(p64, 0, mkd(pb), mkd(pe), tri, (double)eta, event, A, b, &cost);
    float A32[R * P];
    float b32[R];
    float cost32 = 0.0f;
    chain_residual<K, P, float, v3>(prm, 0, pb, pe, tri, eta, event, A32, b32, &cost32);
#pragma unroll
    for (int i = 0; i < R * P; ++i) A[i] = (double)A32[i];
    double dp[P];
    solve_qr<R, P, double>(A, b, dp);
#pragma unroll
    for (int i = 0; i < P; ++i) p64[i] = p64[i] - dp[i];
}

// contributableThroughput (pk.h:760-842) of the converged chains of length K,
// in item order: every test but the occlusion rays per lane, then the K + 1
// occlusion rays as wave packets (per lane when the wave's items belong to
// more than one scene, and past the first segment when they belong to more
// than kPacketPaths paths).  Writes the throughput (0 when any test fails) per
// list position.  A kernel of its own so that the traversal runs at the
// occupancy of its few live registers, not at that of dA/dw's.
//
// The list keeps the batch order (chain length, scene, first triangle), so
// neighbouring positions trace neighbouring rays.  Each XCD (the workgroups with blockIdx % 8 equal)
// takes a contiguous eighth of the positions, so that its L2 holds the nodes
// and triangles of one part of the scene rather than of all of it; its waves
// take kOccludeGrab positions at a time from that eighth's counter, then help
// the other XCDs' eighths.  A grab is 64 to 256 positions (one to four wave
// passes): large enough to keep the counters cold on big lists, small enough
// that a short list still spreads over every wave.
constexpr uint32_t kOccludeGrabMax = 256;
// Most distinct paths in a wave whose segments past the first walk as packets.
constexpr uint32_t kPacketPaths = 16;

// Clears chain length K's per-XCD counters before its k_occlude, or with
// K = 0 every length's (a kernel, not a memset, so that it can be captured in
// a graph).
__global__ void k_clear_occlude(Counters* c, int K) {
    if (threadIdx.x < kXcds * 3u && (K == 0 || (int)(threadIdx.x / kXcds) == K - 1))
        c->occlude[threadIdx.x / kXcds][threadIdx.x % kXcds] = 0u;
}

// The occlusion rays of chain length K's converged items, run by the calling
// wave until that length's list is drained.  `pstack`: the wave's LDS stack.
template <int K>
__device__ __forceinline__ void occlude_persistent(const palz_scene_desc* __restrict__ scenes,
                                                   const palz_path* __restrict__ paths,
                                                   const uint32_t* __restrict__ ranges, uint32_t L,
                                                   Counters* __restrict__ ctr, const uint32_t* __restrict__ excl,
                                                   const uint32_t* __restrict__ list,
                                                   const float* __restrict__ conv_params,
                                                   float* __restrict__ thr_out, PacketStack* pstack) {
    constexpr int P = 2 * K;
    const uint32_t n_items = ctr->n_items, total = ctr->converged;
    const uint32_t begin = compact_pos(excl, ranges[K - 1] * L, n_items, total);
    const uint32_t end = compact_pos(excl, ranges[K] * L, n_items, total);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t waves = gridDim.x * (blockDim.x / 64u);
    const uint32_t grab_size = 64u * min(kOccludeGrabMax / 64u, max(1u, (end - begin) / (waves * 64u * 8u)));
    const uint32_t per_xcd = div_up(end - begin, kXcds);
    const uint32_t home = blockIdx.x % kXcds;
    for (uint32_t k = 0; k < kXcds; ++k) {
        const uint32_t x = (home + k) % kXcds;
        const uint32_t r0 = begin + min(x * per_xcd, end - begin);
        const uint32_t r1 = begin + min((x + 1u) * per_xcd, end - begin);
        for (;;) {
            uint32_t grab = 0u;
            if (lane == 0u) grab = atomicAdd(&ctr->occlude[K - 1][x], grab_size);
            grab = __builtin_amdgcn_readfirstlane(grab);
            if (r0 + grab >= r1 || grab >= r1 - r0) break;
            for (uint32_t j = 0; j < grab_size; j += 64u) {
                const uint32_t s = r0 + grab + j + lane;
                if (__ballot(s < r1) == 0ull) break;
                const bool in_range = s < r1;
                const uint32_t item = list[in_range ? s : r0];
                const uint32_t pi = item / L, l = item % L;
                const palz_path path = paths[pi];
                const palz_scene_desc* sc = scenes + path.scene;
                Segment seg[K + 1] = {};
                float thr = 0.0f;
                if (in_range) {
                    ChainTri tri[K];
#pragma unroll
                    for (int t = 0; t < K; ++t) tri[t] = load_chain_tri(sc->tris, sc->attribs, path.tris[t]);
                    float prm[P];
#pragma unroll
                    for (int i = 0; i < P; ++i) prm[i] = conv_params[(size_t)item * kQueueStride + i];
                    const v3 pe = mk(sc->points[path.pixel].p);
                    const v3 pb = mk(sc->p_light);
                    thr = throughput_unoccluded<K>(prm, pb, pe, tri, path.event, item_eta(sc, l), seg);
                }
                const uint32_t scene0 = __builtin_amdgcn_readfirstlane(path.scene);
                const bool one_scene = __ballot(path.scene != scene0) == 0ull;
                // A packet walk visits the union of its lanes' nodes.  The first
                // segments of a one-scene wave leave one light point for
                // neighbouring first vertices and stay close; the later ones
                // of a wave that holds many paths (one wavelength each) spread
                // and walk per lane.
                const uint32_t prev_pi = (uint32_t)__shfl_up((int)pi, 1, 64);
                const uint32_t n_paths_wave =
                    (uint32_t)__popcll(__ballot(in_range && (lane == 0u || prev_pi != pi)));
                bool visible = 0.0f < thr;
#pragma unroll
                for (int i = 0; i < K + 1; ++i) {
                    const v3 ird = inv_rd(seg[i].rd);
                    bool hit;
                    const bool packet = one_scene && (i == 0 || n_paths_wave <= kPacketPaths);
                    if (packet) {
                        const palz_scene_desc* us = scenes + scene0;
                        if (us->wide) {
                            const WaveSlabs ws = wave_slabs(ird, __ballot(visible));
                            hit = ws.ordered ? any_hit_wide<true>(us->nodes, us->wide, us->tris, *us->root, visible,
                                                                  seg[i].ro, seg[i].rd, ird, ws, 1.0f, pstack)
                                             : any_hit_wide<false>(us->nodes, us->wide, us->tris, *us->root,
                                                                   visible, seg[i].ro, seg[i].rd, ird, ws, 1.0f,
                                                                   pstack);
                        } else {
                            hit = any_hit_packet(us->nodes, us->tris, *us->root, visible, seg[i].ro, seg[i].rd, ird,
                                                 1.0f, pstack);
                        }
                    } else {
                        hit = visible &&
                              any_hit_escape(sc->nodes, sc->tris, *sc->root, seg[i].ro, seg[i].rd, ird, 1.0f);
                    }
                    if (hit) visible = false;
                }
                if (in_range) thr_out[s] = visible ? thr : 0.0f;
            }
        }
    }
}

template <int K>
__global__ void __launch_bounds__(256, 8)  // 8 waves per SIMD: few live registers
    k_occlude(const palz_scene_desc* __restrict__ scenes, const palz_path* __restrict__ paths,
              const uint32_t* __restrict__ ranges, uint32_t L, Counters* __restrict__ ctr,
              const uint32_t* __restrict__ excl, const uint32_t* __restrict__ list,
              const float* __restrict__ conv_params, float* __restrict__ thr_out) {
    __shared__ PacketStack packet_stacks[256 / 64];
    occlude_persistent<K>(scenes, paths, ranges, L, ctr, excl, list, conv_params, thr_out,
                          &packet_stacks[threadIdx.x >> 6]);
}

// The occlusion rays of every chain length in one persistent kernel: each
// wave drains the K = 3 list, then K = 2's, then K = 1's.
__global__ void __launch_bounds__(256, 8)
    k_occlude_mixed(const palz_scene_desc* __restrict__ scenes, const palz_path* __restrict__ paths,
                    const uint32_t* __restrict__ ranges, uint32_t L, Counters* __restrict__ ctr,
                    const uint32_t* __restrict__ excl, const uint32_t* __restrict__ list,
                    const float* __restrict__ conv_params, float* __restrict__ thr_out) {
    __shared__ PacketStack packet_stacks[256 / 64];
    PacketStack* ps = &packet_stacks[threadIdx.x >> 6];
    occlude_persistent<3>(scenes, paths, ranges, L, ctr, excl, list, conv_params, thr_out, ps);
    occlude_persistent<2>(scenes, paths, ranges, L, ctr, excl, list, conv_params, thr_out, ps);
    occlude_persistent<1>(scenes, paths, ranges, L, ctr, excl, list, conv_params, thr_out, ps);
}

// dA/dw and irradiance of the converged chains of length K, in item order,
// from the throughput k_occlude left per list position.  kMode: the refinement
// mode.  PALZ_REFINE_VERTICES and PALZ_REFINE_ALL run the FP64 step here on
// the item's data already in registers; with ALL its chain feeds dA/dw,
// otherwise dA/dw uses the FP32 chain.
template <int K, int kMode>
__global__ void __launch_bounds__(256)
    k_finish(const palz_scene_desc* __restrict__ scenes, const palz_path* __restrict__ paths,
             const uint32_t* __restrict__ ranges, uint32_t L, Counters* __restrict__ ctr,
             const uint32_t* __restrict__ excl, const uint32_t* __restrict__ list,
             const float* __restrict__ conv_params, const int* __restrict__ conv_iters,
             const float* __restrict__ thr_in, palz_path_result* __restrict__ results,
             double* __restrict__ refined) {
    constexpr int P = 2 * K;
    const uint32_t n_items = ctr->n_items, total = ctr->converged;
    const uint32_t begin = compact_pos(excl, ranges[K - 1] * L, n_items, total);
    const uint32_t end = compact_pos(excl, ranges[K] * L, n_items, total);
    unsigned long long n_contrib = 0ull;
    for (uint32_t s = begin + blockIdx.x * blockDim.x + threadIdx.x; s < end; s += gridDim.x * blockDim.x) {
        const uint32_t item = list[s];
        const uint32_t pi = item / L, l = item % L;
        const palz_path path = paths[pi];
        const palz_scene_desc* sc = scenes + path.scene;
        ChainTri tri[K];
#pragma unroll
        for (int k = 0; k < K; ++k) tri[k] = load_chain_tri(sc->tris, sc->attribs, path.tris[k]);
        float prm[P];
#pragma unroll
        for (int i = 0; i < P; ++i) prm[i] = conv_params[(size_t)item * kQueueStride + i];
        const palz_first_diffuse q = sc->points[path.pixel];
        const v3 pe = mk(q.p);
        const v3 pb = mk(sc->p_light);
        const float eta = item_eta(sc, l);
        const uint32_t event = path.event;  // per lane (see k_solve_plain)
        const float thr = thr_in[s];        // contributableThroughput, from k_occlude
        float pf[P];  // the reported parameters
        float pd[P];  // the chain dA/dw and the receiver cosine take
#pragma unroll
        for (int i = 0; i < P; ++i) pf[i] = pd[i] = prm[i];
        // The FP64 step: before dA/dw when dA/dw takes its chain (ALL), else
        // after it, so that the step's registers and dA/dw's are not live at
        // once.
        auto refine = [&]() {
            double p64[P];
            refine_step<K>(prm, pb, pe, tri, eta, event, p64);
#pragma unroll
            for (int i = 0; i < P; ++i) pf[i] = (float)p64[i];
            if (refined) {
#pragma unroll
                for (int i = 0; i < 8; ++i) refined[(size_t)item * 8 + i] = i < P ? p64[i] : 0.0;
            }
        };
        if (kMode == PALZ_REFINE_ALL) {
            refine();
#pragma unroll
            for (int i = 0; i < P; ++i) pd[i] = pf[i];
        }
        palz_path_result r;
        r.status = PALZ_PATH_CONVERGED;
        r.iterations = conv_iters[item];
        r.contribution = 0.0f;
        r.footprint = 0.0f;
        if (0.0f < thr) {
            const float area = dadw<K>(pb, chain_vertex(tri[0], pd[0], pd[1]) - pb, pe, tri, event, eta);
            const v3 last = chain_vertex(tri[K - 1], pd[P - 2], pd[P - 1]);
            float c = thr * sc->light_intensity / area * hmax(dot(normalize(last - pe), mk(q.ng)), 0.0f);
            c = hmin(c, sc->radiance_clamp);
            add_contribution(*sc, (size_t)l * sc->n_points + path.pixel, c);
            r.status = PALZ_PATH_CONTRIBUTES;
            r.contribution = c;
            r.footprint = area;
            ++n_contrib;
        }
        if (kMode == PALZ_REFINE_VERTICES) refine();
#pragma unroll
        for (int i = 0; i < 8; ++i) r.params[i] = i < P ? pf[i] : 0.0f;
        if (results) results[item] = r;
    }
    wave_add_stat(&ctr->stats[(K - 1) * 4 + 2], n_contrib);
}

// Items of chain length 4 (and of none) are not part of a batch: mark and
// count them; their flags stay clear for the compaction.
__global__ void __launch_bounds__(256) k_skipped(const uint32_t* __restrict__ ranges, uint32_t L,
                                                 Counters* __restrict__ ctr, uint32_t* __restrict__ flags,
                                                 palz_path_result* __restrict__ results) {
    const uint32_t b = ranges[3] * L, e = ranges[4] * L, z = ranges[0] * L;
    if (blockIdx.x == 0 && threadIdx.x == 0) ctr->stats[15] = (unsigned long long)(e - b);
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < z + (e - b); i += stride) {
        const uint32_t item = i < z ? i : b + (i - z);
        flags[item] = 0u;
        if (results) {
            palz_path_result r = {};
            results[item] = r;
        }
    }
}

// Converged items, listed in item order from the exclusive scan of their flags.
// Also counts, per chain length, the paths with at least one converged item
// (an item is its path's first converged one when the scan has not moved
// since the path's first item).
__global__ void __launch_bounds__(256) k_compact_scatter(const uint32_t* __restrict__ excl,
                                                         Counters* __restrict__ ctr,
                                                         const uint32_t* __restrict__ ranges, uint32_t L,
                                                         uint32_t* __restrict__ list) {
    const uint32_t n = ctr->n_items, total = ctr->converged;
    unsigned long long first[3] = {0ull, 0ull, 0ull};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t p = excl[i];
        const uint32_t next = i + 1 < n ? excl[i + 1] : total;
        if (next != p) {
            list[p] = i;
            const uint32_t pi = i / L;
            if (excl[pi * L] == p) {
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    if (ranges[k] <= pi && pi < ranges[k + 1]) ++first[k];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) wave_add_stat(&ctr->conv_paths[k], first[k]);
}

template <int K>
void launch_newton(const palz_launch_config& c, int cus, hipStream_t st, const palz_scene_desc* scenes,
                   uint32_t n_scenes, const palz_path* paths, const uint32_t* ranges, uint32_t L, Counters* ctr,
                   uint32_t* flags, float* cp, int* cit, const uint8_t* admissible, palz_path_result* results) {
    k_newton<K><<<c.blocks_per_cu * cus, c.block, 0, st>>>(scenes, n_scenes, paths, ranges, L, (uint32_t)c.chunk,
                                                          ctr, flags, cp, cit, admissible, results);
}
template <int K>
void launch_occlude(const palz_launch_config& c, int cus, hipStream_t st, const palz_scene_desc* scenes,
                    const palz_path* paths, const uint32_t* ranges, uint32_t L, Counters* ctr, const uint32_t* excl,
                    const uint32_t* list, const float* cp, float* thr) {
    k_clear_occlude<<<1, 64, 0, st>>>(ctr, K);
    k_occlude<K><<<cus * 8, c.finish_block, 0, st>>>(scenes, paths, ranges, L, ctr, excl, list, cp, thr);
}
void launch_newton_mixed(const palz_launch_config& m, const palz_launch_config* cfg, int cus, hipStream_t st,
                         const palz_scene_desc* scenes, uint32_t n_scenes, const palz_path* paths,
                         const uint32_t* ranges, uint32_t L, Counters* ctr, uint32_t* flags, float* cp, int* cit,
                         const uint8_t* admissible, palz_path_result* results) {
    auto kernel = m.chunk == 3 ? k_newton_mixed<3> : k_newton_mixed<2>;
    kernel<<<m.blocks_per_cu * cus, m.block, 0, st>>>(scenes, n_scenes, paths, ranges, L, (uint32_t)cfg[0].chunk,
                                                      (uint32_t)cfg[1].chunk, (uint32_t)cfg[2].chunk, ctr, flags, cp,
                                                      cit, admissible, results);
}
void launch_occlude_mixed(const palz_launch_config& m, int cus, hipStream_t st, const palz_scene_desc* scenes,
                          const palz_path* paths, const uint32_t* ranges, uint32_t L, Counters* ctr,
                          const uint32_t* excl, const uint32_t* list, const float* cp, float* thr) {
    k_clear_occlude<<<1, 64, 0, st>>>(ctr, 0);
    k_occlude_mixed<<<cus * 8, m.finish_block, 0, st>>>(scenes, paths, ranges, L, ctr, excl, list, cp, thr);
}
template <int K>
void launch_finish(const palz_launch_config& c, int cus, hipStream_t st, const palz_scene_desc* scenes,
                   const palz_path* paths, const uint32_t* ranges, uint32_t L, Counters* ctr, const uint32_t* excl,
                   const uint32_t* list, const float* cp, const int* cit, const float* thr, int refine,
                   palz_path_result* results, double* refined) {
    const dim3 grid(cus * 8), block(c.finish_block);
    if (refine == PALZ_REFINE_ALL)
        k_finish<K, PALZ_REFINE_ALL><<<grid, block, 0, st>>>(scenes, paths, ranges, L, ctr, excl, list, cp, cit, thr,
                                                             results, refined);
    else if (refine == PALZ_REFINE_VERTICES)
        k_finish<K, PALZ_REFINE_VERTICES><<<grid, block, 0, st>>>(scenes, paths, ranges, L, ctr, excl, list, cp, cit,
                                                                  thr, results, refined);
    else
        k_finish<K, PALZ_REFINE_NONE><<<grid, block, 0, st>>>(scenes, paths, ranges, L, ctr, excl, list, cp, cit, thr,
                                                              results, nullptr);
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
    // per item: converged flag (scanned in place to list positions), FP32
    // solution and iteration count; the compacted list of converged items
    uint32_t* flags = nullptr;
    float* conv_params = nullptr;
    int* conv_iters = nullptr;
    uint32_t* list = nullptr;
    float* thr = nullptr;  // per list position: the throughput after the occlusion rays
    uint32_t* scan_sums = nullptr;
    size_t cap_items = 0;
    // Shapes until set or tuned: the autotuner's choice for a full-GPU list of
    // 16 wavelengths per path (the spectral slab), which fills every SIMD with
    // waves: the Newton kernels at their register budgets (5 / 3 / 2 waves per
    // SIMD for K = 1 / 2 / 3), the occlusion kernels at 8 (256-lane blocks).
    palz_launch_config cfg[3] = {{256, 8, 1024, 256}, {256, 4, 1024, 256}, {128, 8, 256, 256}};
    // the mixed-length kernels' shape (block, blocks_per_cu, the Newton
    // kernel's register budget in waves per SIMD (2 or 3), occlusion block);
    // each length keeps its own queue chunk from cfg
    palz_launch_config cfg_mixed = {256, 4, 3, 256};
    // fused mode, a bit mask: 1 = one persistent Newton kernel over every chain
    // length, 2 = one occlusion kernel over every chain length (3 = both, 0 =
    // one of each per length); -1 = chosen by the autotuner (per length until
    // it has run)
    int fused = -1;
    int auto_fused = 0;  // the autotuner's choice in mode -1
    int fused_mask() const { return fused == -1 ? auto_fused : fused; }
    std::string tune_cache;  // the autotuner's tile-shape cache file ("" = none)
    int cus = 256;
    // hipGraph cache
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t capture = nullptr;  // graphs are captured here (never the null stream)
    // Captured graphs run the chain lengths as three branches: K = 2 on the
    // capture stream, K = 1 and K = 3 on these two (forked and joined with
    // events), so that short kernels of one length overlap the others'.
    hipStream_t side[2] = {nullptr, nullptr};
    hipEvent_t fork_ev[6] = {};
    int branches = -1;  // -1: decided at capture from the chain-length mix; 0: off; 1: on
    bool timing = false;
    bool timed_solves = false;
    hipEvent_t ev[13] = {};
    int cull = 0;  // 0 off, 1 interval test, 2 interval then affine test
    uint8_t* admissible = nullptr;  // per sorted path, when culling
    struct Key {
        const void *scenes, *results, *refined;
        uint32_t n_scenes, L;
        int refine;
        palz_launch_config cfg[3];
        int cull;
        int fused;
        palz_launch_config cfg_mixed;
    } key{};

    int ensure_paths(uint32_t n) {
        if (n <= cap_paths) return PALZ_OK;
        for (void* p : {(void*)paths, (void*)order, (void*)words, (void*)words_tmp, (void*)admissible})
            if (p) PALZ_HIP(hipFree(p));
        PALZ_HIP(hipMalloc(&admissible, n));
        PALZ_HIP(hipMalloc(&paths, sizeof(palz_path) * n));
        PALZ_HIP(hipMalloc(&order, sizeof(uint32_t) * n));
        PALZ_HIP(hipMalloc(&words, sizeof(uint64_t) * n));
        PALZ_HIP(hipMalloc(&words_tmp, sizeof(uint64_t) * n));
        cap_paths = n;
        return PALZ_OK;
    }
    int ensure_items(size_t n) {
        if (n <= cap_items) return PALZ_OK;
        drop_graph();  // a captured graph holds the old buffers
        for (void* p : {(void*)flags, (void*)conv_params, (void*)conv_iters, (void*)list, (void*)thr,
                        (void*)scan_sums})
            if (p) PALZ_HIP(hipFree(p));
        PALZ_HIP(hipMalloc(&thr, sizeof(float) * n));
        PALZ_HIP(hipMalloc(&flags, sizeof(uint32_t) * n));
        PALZ_HIP(hipMalloc(&conv_params, sizeof(float) * palz::kQueueStride * n));
        PALZ_HIP(hipMalloc(&conv_iters, sizeof(int) * n));
        PALZ_HIP(hipMalloc(&list, sizeof(uint32_t) * n));
        PALZ_HIP(hipMalloc(&scan_sums, sizeof(uint32_t) * palz::scan_scratch_entries((uint32_t)n)));
        cap_items = n;
        return PALZ_OK;
    }
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    int ensure_branches() {
        for (hipStream_t& x : side)
            if (!x) PALZ_HIP(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
        for (hipEvent_t& e : fork_ev)
            if (!e) PALZ_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        return PALZ_OK;
    }
    ~palz_batched() {
        drop_graph();
        for (hipStream_t x : side)
            if (x) (void)hipStreamDestroy(x);
        for (hipEvent_t e : fork_ev)
            if (e) (void)hipEventDestroy(e);
        if (capture) (void)hipStreamDestroy(capture);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        for (void* p : {(void*)paths, (void*)order, (void*)words, (void*)words_tmp, (void*)admissible,
                        (void*)ranges, (void*)ctr,
                        (void*)flags, (void*)conv_params, (void*)conv_iters, (void*)list, (void*)thr,
                        (void*)scan_sums})
            if (p) (void)hipFree(p);
    }

    template <int K>
    void newton(const palz_launch_config& c, const palz_scene_desc* scenes, uint32_t n_scenes, uint32_t L,
                palz_path_result* results, hipStream_t st) {
        palz::launch_newton<K>(c, cus, st, scenes, n_scenes, paths, ranges, L, ctr, flags, conv_params, conv_iters,
                               cull ? admissible : nullptr, results);
    }
    template <int K>
    void occlude(const palz_launch_config& c, const palz_scene_desc* scenes, uint32_t L, hipStream_t st) {
        palz::launch_occlude<K>(c, cus, st, scenes, paths, ranges, L, ctr, flags, list, conv_params, thr);
    }
    template <int K>
    void finish(const palz_launch_config& c, const palz_scene_desc* scenes, uint32_t L, int refine,
                palz_path_result* results, double* refined, hipStream_t st) {
        palz::launch_finish<K>(c, cus, st, scenes, paths, ranges, L, ctr, flags, list, conv_params, conv_iters, thr,
                               refine, results, refined);
    }
    void newton_mixed(const palz_launch_config& m, const palz_scene_desc* scenes, uint32_t n_scenes, uint32_t L,
                      palz_path_result* results, hipStream_t st) {
        palz::launch_newton_mixed(m, cfg, cus, st, scenes, n_scenes, paths, ranges, L, ctr, flags, conv_params,
                                  conv_iters, cull ? admissible : nullptr, results);
    }
    void occlude_mixed(const palz_launch_config& m, const palz_scene_desc* scenes, uint32_t L, hipStream_t st) {
        palz::launch_occlude_mixed(m, cus, st, scenes, paths, ranges, L, ctr, flags, list, conv_params, thr);
    }
    // K = 4 marks, the scan of the converged flags and the scatter into the list
    void compact(uint32_t L, palz_path_result* results, hipStream_t st) {
        using namespace palz;
        k_skipped<<<256, 256, 0, st>>>(ranges, L, ctr, flags, results);
        launch_exclusive_scan(flags, flags, (uint32_t)cap_items, &ctr->n_items, &ctr->converged, scan_sums, st);
        k_compact_scatter<<<cus * 8, 256, 0, st>>>(flags, ctr, ranges, L, list);
    }

    // The whole solve as three branches (see `side`): fork after the reset,
    // join before the compaction (which scans every chain length's flags),
    // fork again for the finishing kernels and join at the end.
    int launch_branches(const palz_scene_desc* scenes, uint32_t n_scenes, uint32_t L, int refine,
                        palz_path_result* results, double* refined, hipStream_t st) {
        using namespace palz;
        auto fork = [&](hipEvent_t e) -> int {
            PALZ_HIP(hipEventRecord(e, st));
            for (hipStream_t x : side) PALZ_HIP(hipStreamWaitEvent(x, e, 0));
            return PALZ_OK;
        };
        auto join = [&](hipEvent_t e0, hipEvent_t e1) -> int {
            PALZ_HIP(hipEventRecord(e0, side[0]));
            PALZ_HIP(hipEventRecord(e1, side[1]));
            PALZ_HIP(hipStreamWaitEvent(st, e0, 0));
            PALZ_HIP(hipStreamWaitEvent(st, e1, 0));
            return PALZ_OK;
        };
        k_reset<<<1, 64, 0, st>>>(ctr, ranges, L);
        if (cull) launch_cull_paths(scenes, n_scenes, paths, cap_paths, ranges + 5, admissible, st, cull);
        int rc = fork(fork_ev[0]);
        if (rc != PALZ_OK) return rc;
        newton<1>(cfg[0], scenes, n_scenes, L, results, side[0]);
        newton<3>(cfg[2], scenes, n_scenes, L, results, side[1]);
        newton<2>(cfg[1], scenes, n_scenes, L, results, st);
        if ((rc = join(fork_ev[1], fork_ev[2])) != PALZ_OK) return rc;
        compact(L, results, st);
        if ((rc = fork(fork_ev[3])) != PALZ_OK) return rc;
        occlude<1>(cfg[0], scenes, L, side[0]);
        finish<1>(cfg[0], scenes, L, refine, results, refined, side[0]);
        occlude<3>(cfg[2], scenes, L, side[1]);
        finish<3>(cfg[2], scenes, L, refine, results, refined, side[1]);
        occlude<2>(cfg[1], scenes, L, st);
        finish<2>(cfg[1], scenes, L, refine, results, refined, st);
        if ((rc = join(fork_ev[4], fork_ev[5])) != PALZ_OK) return rc;
        PALZ_LAUNCH_CHECK();
        return PALZ_OK;
    }

    int launch_all(const palz_scene_desc* scenes, uint32_t n_scenes, uint32_t L, int refine,
                   palz_path_result* results, double* refined, hipStream_t st, bool timed) {
        using namespace palz;
        const bool t = timed && timing;
        k_reset<<<1, 64, 0, st>>>(ctr, ranges, L);
        if (t) PALZ_HIP(hipEventRecord(ev[0], st));
        if (cull) launch_cull_paths(scenes, n_scenes, paths, cap_paths, ranges + 5, admissible, st, cull);
        if (t) PALZ_HIP(hipEventRecord(ev[9], st));
        const int f = fused_mask();
        // a fused kernel's time goes to its K = 1 slot; the K = 2 and 3 slots stay ~0
        if (f & 1) {
            newton_mixed(cfg_mixed, scenes, n_scenes, L, results, st);
            for (int i : {1, 2, 3})
                if (t) PALZ_HIP(hipEventRecord(ev[i], st));
        } else {
            newton<1>(cfg[0], scenes, n_scenes, L, results, st);
            if (t) PALZ_HIP(hipEventRecord(ev[1], st));
            newton<2>(cfg[1], scenes, n_scenes, L, results, st);
            if (t) PALZ_HIP(hipEventRecord(ev[2], st));
            newton<3>(cfg[2], scenes, n_scenes, L, results, st);
            if (t) PALZ_HIP(hipEventRecord(ev[3], st));
        }
        compact(L, results, st);
        if (t) PALZ_HIP(hipEventRecord(ev[4], st));
        if (f & 2) occlude_mixed(cfg_mixed, scenes, L, st);
        if (!(f & 2)) occlude<1>(cfg[0], scenes, L, st);
        if (t) PALZ_HIP(hipEventRecord(ev[10], st));
        finish<1>(cfg[0], scenes, L, refine, results, refined, st);
        if (t) PALZ_HIP(hipEventRecord(ev[5], st));
        if (!(f & 2)) occlude<2>(cfg[1], scenes, L, st);
        if (t) PALZ_HIP(hipEventRecord(ev[11], st));
        finish<2>(cfg[1], scenes, L, refine, results, refined, st);
        if (t) PALZ_HIP(hipEventRecord(ev[6], st));
        if (!(f & 2)) occlude<3>(cfg[2], scenes, L, st);
        if (t) PALZ_HIP(hipEventRecord(ev[12], st));
        finish<3>(cfg[2], scenes, L, refine, results, refined, st);
        if (t) PALZ_HIP(hipEventRecord(ev[7], st));
        if (t) PALZ_HIP(hipEventRecord(ev[8], st));
        PALZ_LAUNCH_CHECK();
        if (t) timed_solves = true;
        return PALZ_OK;
    }
};

namespace {
bool valid_config(const palz_launch_config& c, bool mixed = false) {
    auto ok_block = [](int v) { return v == 64 || v == 128 || v == 256; };
    const bool ok_chunk = mixed ? (c.chunk == 2 || c.chunk == 3) : (c.chunk >= 1 && c.chunk <= (1 << 16));
    return ok_block(c.block) && ok_block(c.finish_block) && c.blocks_per_cu >= 1 && c.blocks_per_cu <= 32 && ok_chunk;
}

// The autotuner's tile-shape cache.  A workload is keyed by the device's CU
// count, the batch's scene count (rounded up to a power of two), wavelength
// count, refinement, culling and fused modes and, per chain length, its item
// count rounded up to a power of two; workloads with equal keys take the same
// launch shapes.  The file holds one line per tuning,
//     palz-tune <10 key fields> : <cfg K=1,2,3 and mixed: 4 x 4 fields> <6 times in us> <fused mask chosen>
// and the last line with a matching key wins.
constexpr int kTuneKeyFields = 10;
constexpr int kTuneFields = 4 * 4 + 6 + 1;
struct TuneKey {
    int v[kTuneKeyFields];
};
int ceil_log2_plus1(uint64_t x) {
    int r = 0;
    while (r < 63 && (1ull << r) < x + 1ull) ++r;
    return r;
}
TuneKey tune_key(int cus, const uint32_t* ranges, uint32_t n_scenes, uint32_t L, int refine, int cull, int fused) {
    TuneKey k;
    k.v[0] = 3;  // format version
    k.v[1] = cus;
    k.v[2] = ceil_log2_plus1(n_scenes);
    k.v[3] = (int)L;
    k.v[4] = refine;
    k.v[5] = cull;
    k.v[6] = fused;
    for (int K = 1; K <= 3; ++K) k.v[6 + K] = ceil_log2_plus1((uint64_t)(ranges[K] - ranges[K - 1]) * L);
    return k;
}
bool parse_ints(const char** p, int n, long* out) {
    for (int i = 0; i < n; ++i) {
        char* e = nullptr;
        const long x = strtol(*p, &e, 10);
        if (e == *p) return false;
        out[i] = x;
        *p = e;
    }
    return true;
}
bool tune_cache_lookup(const std::string& path, const TuneKey& key, int32_t* vals) {
    FILE* f = fopen(path.c_str(), "r");
    if (!f) return false;
    char line[1024];
    bool found = false;
    while (fgets(line, sizeof line, f)) {
        const char* p = line;
        if (strncmp(p, "palz-tune ", 10) != 0) continue;
        p += 10;
        long kv[kTuneKeyFields], v[kTuneFields];
        if (!parse_ints(&p, kTuneKeyFields, kv)) continue;
        while (*p == ' ') ++p;
        if (*p != ':') continue;
        ++p;
        if (!parse_ints(&p, kTuneFields, v)) continue;
        bool same = true;
        for (int i = 0; i < kTuneKeyFields; ++i) same = same && kv[i] == key.v[i];
        if (!same) continue;
        for (int i = 0; i < kTuneFields; ++i) vals[i] = (int32_t)v[i];
        found = true;
    }
    fclose(f);
    return found;
}
bool tune_cache_append(const std::string& path, const TuneKey& key, const int32_t* vals) {
    FILE* f = fopen(path.c_str(), "a");
    if (!f) return false;
    fprintf(f, "palz-tune");
    for (int i = 0; i < kTuneKeyFields; ++i) fprintf(f, " %d", key.v[i]);
    fprintf(f, " :");
    for (int i = 0; i < kTuneFields; ++i) fprintf(f, " %d", (int)vals[i]);
    fprintf(f, "\n");
    return fclose(f) == 0;
}

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
    k_reset_bounds<<<1, 64, 0, st>>>(b->ranges);
    k_key_bounds<<<std::min(nb, 1024u), 256, 0, st>>>(d_paths, n_paths, d_count, b->ranges);
    k_size_words<<<nb, 256, 0, st>>>(d_paths, n_paths, d_count, b->ranges, b->words);
    const SortScratch& s = sorter->scratch;
    launch_sort_prepare(s, nullptr, st);
    launch_sort_count(s, b->words, n_paths, 0, 32, st, d_count);
    const int npass = launch_sort_passes(s, b->words, b->words_tmp, nullptr, nullptr, n_paths, 0, 32, st, nullptr,
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
    if (!b || !cfg || K < 0 || K > 3) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_batched_set_config: bad argument");
    const palz_launch_config& c = *cfg;
    if (!valid_config(c, K == 0))
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_batched_set_config: block/finish_block in {64,128,256}, "
                                               "blocks_per_cu in 1..32, chunk in 1..65536 (K = 0: 2 or 3)");
    (K == 0 ? b->cfg_mixed : b->cfg[K - 1]) = c;
    b->drop_graph();
    return PALZ_OK;
}

int palz_batched_get_config(palz_batched* b, int K, palz_launch_config* cfg) {
    if (!b || !cfg || K < 0 || K > 3) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_batched_get_config: bad argument");
    *cfg = K == 0 ? b->cfg_mixed : b->cfg[K - 1];
    return PALZ_OK;
}

int palz_batched_set_fused(palz_batched* b, int mode) {
    if (!b || mode < -1 || mode > 3)
        return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_batched_set_fused: mode must be -1, 0, 1, 2 or 3");
    b->fused = mode;
    b->drop_graph();
    return PALZ_OK;
}

int palz_batched_get_fused(palz_batched* b, int* mode, int* active) {
    if (!b) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_batched_get_fused: NULL batch");
    if (mode) *mode = b->fused;
    if (active) *active = b->fused_mask();
    return PALZ_OK;
}

int palz_batched_solve(palz_batched* b, const palz_scene_desc* d_scenes, uint32_t n_scenes, uint32_t n_lambdas,
                       int refine, palz_path_result* d_results, double* d_refined, void* stream) {
    int rc = check_solve_args(b, d_scenes, n_scenes, n_lambdas, refine);
    if (rc != PALZ_OK) return rc;
    return b->launch_all(d_scenes, n_scenes, n_lambdas, refine, d_results, d_refined, (hipStream_t)stream, true);
}

int palz_batched_solve_graph(palz_batched* b, const palz_scene_desc* d_scenes, uint32_t n_scenes,
                             uint32_t n_lambdas, int refine, palz_path_result* d_results, double* d_refined,
                             void* stream) {
    int rc = check_solve_args(b, d_scenes, n_scenes, n_lambdas, refine);
    if (rc != PALZ_OK) return rc;
    hipStream_t st = (hipStream_t)stream;
    palz_batched::Key k{d_scenes, d_results, d_refined, n_scenes, n_lambdas, refine, {}, b->cull, b->fused_mask(),
                        b->cfg_mixed};
    std::copy(b->cfg, b->cfg + 3, k.cfg);
    const bool same = b->exec && std::memcmp(&k, &b->key, sizeof(k)) == 0;
    if (!same) {
        b->drop_graph();
        // Branches pay when no chain length dominates (many scenes of mixed
        // lengths); when one length holds most paths its kernels fill the GPU
        // and the others only compete with them.
        bool branch = b->branches == 1 && b->fused_mask() == 0;
        if (b->branches < 0 && b->fused_mask() == 0) {
            uint32_t r[6];
            PALZ_HIP(hipMemcpy(r, b->ranges, sizeof(r), hipMemcpyDeviceToHost));
            const uint32_t most = std::max(r[1] - r[0], std::max(r[2] - r[1], r[3] - r[2]));
            branch = (uint64_t)most * 4u < (uint64_t)(r[3] - r[0]) * 3u;
        }
        if (branch) {
            rc = b->ensure_branches();
            if (rc != PALZ_OK) return rc;
        }
        PALZ_HIP(hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal));
        rc = branch ? b->launch_branches(d_scenes, n_scenes, n_lambdas, refine, d_results, d_refined, b->capture)
                    : b->launch_all(d_scenes, n_scenes, n_lambdas, refine, d_results, d_refined, b->capture, false);
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

int palz_batched_set_branches(palz_batched* b, int mode) {
    if (!b || mode < -1 || mode > 1)
        return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_batched_set_branches: mode must be -1, 0 or 1");
    b->branches = mode;
    b->drop_graph();
    return PALZ_OK;
}

int palz_batched_set_timing(palz_batched* b, int enable) {
    if (!b) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_batched_set_timing: NULL");
    if (enable && !b->ev[0])
        for (hipEvent_t& e : b->ev) PALZ_HIP(hipEventCreate(&e));
    b->timing = enable != 0;
    b->timed_solves = false;
    return PALZ_OK;
}

int palz_batched_timings(palz_batched* b, float* ms12) {
    if (!b || !ms12) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_batched_timings: NULL");
    if (!b->timing || !b->timed_solves)
        return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_batched_timings: no timed solve since timing was enabled");
    hipEvent_t* ev = b->ev;
    PALZ_HIP(hipEventSynchronize(ev[8]));
    // (from, to) event pairs in the order of the reported durations
    const int pairs[12][2] = {{9, 1}, {1, 2}, {2, 3}, {3, 4}, {4, 10}, {10, 5},
                              {5, 11}, {11, 6}, {6, 12}, {12, 7}, {7, 8}, {0, 9}};
    for (int i = 0; i < 12; ++i) PALZ_HIP(hipEventElapsedTime(&ms12[i], ev[pairs[i][0]], ev[pairs[i][1]]));
    return PALZ_OK;
}

int palz_batched_set_culling(palz_batched* b, int enable) {
    if (!b) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_batched_set_culling: NULL");
    if (enable < 0 || enable > 2)
        return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_batched_set_culling: mode must be 0, 1 or 2");
    b->cull = enable;
    b->drop_graph();
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

int palz_batched_path_stats(palz_batched* b, uint64_t* converged_paths3, void* stream) {
    if (!b || !converged_paths3) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_batched_path_stats: NULL");
    PALZ_HIP(hipStreamSynchronize((hipStream_t)stream));
    palz::Counters h;
    PALZ_HIP(hipMemcpy(&h, b->ctr, sizeof(h), hipMemcpyDeviceToHost));
    for (int k = 0; k < 3; ++k) converged_paths3[k] = h.conv_paths[k];
    return PALZ_OK;
}

int palz_batched_set_tune_cache(palz_batched* b, const char* path) {
    if (!b) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_batched_set_tune_cache: NULL batch");
    b->tune_cache = path ? path : "";
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
    const TuneKey key = tune_key(b->cus, ranges, n_scenes, n_lambdas, refine, b->cull, b->fused);
    float newton_ms[3] = {0.0f, 0.0f, 0.0f}, finish_ms[3] = {0.0f, 0.0f, 0.0f};
    int32_t times_us[6] = {0, 0, 0, 0, 0, 0};  // Newton and finish per chain length, as reported and cached
    auto write_report = [&]() {
        if (!report) return;
        for (int K = 1; K <= 3; ++K) {
            const palz_launch_config& c = b->cfg[K - 1];
            int32_t* r = report + (K - 1) * 6;
            r[0] = c.block;
            r[1] = c.blocks_per_cu;
            r[2] = c.chunk;
            r[3] = c.finish_block;
            r[4] = times_us[2 * (K - 1)];
            r[5] = times_us[2 * (K - 1) + 1];
        }
    };
    if (!b->tune_cache.empty()) {
        int32_t v[kTuneFields];
        if (tune_cache_lookup(b->tune_cache, key, v)) {
            palz_launch_config c[4];
            for (int i = 0; i < 4; ++i) c[i] = {v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]};
            const bool ok = valid_config(c[0]) && valid_config(c[1]) && valid_config(c[2]) &&
                            valid_config(c[3], true) && v[22] >= 0 && v[22] <= 3;
            if (ok) {
                std::copy(c, c + 3, b->cfg);
                b->cfg_mixed = c[3];
                b->auto_fused = v[22];
                for (int i = 0; i < 6; ++i) times_us[i] = v[16 + i];
                write_report();
                b->drop_graph();
                return PALZ_OK;
            }
        }
    }
    hipEvent_t e0, e1;
    PALZ_HIP(hipEventCreate(&e0));
    PALZ_HIP(hipEventCreate(&e1));
    const int blocks[3] = {64, 128, 256};
    const int per_cu[4] = {1, 2, 4, 8};
    const int chunks[3] = {64, 256, 1024};
    auto elapsed = [&]() -> float {
        (void)hipEventRecord(e1, st);
        (void)hipEventSynchronize(e1);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        return ms;
    };
    auto run_newton = [&](int K, const palz_launch_config& c) {
        if (K == 1) b->newton<1>(c, d_scenes, n_scenes, n_lambdas, nullptr, st);
        if (K == 2) b->newton<2>(c, d_scenes, n_scenes, n_lambdas, nullptr, st);
        if (K == 3) b->newton<3>(c, d_scenes, n_scenes, n_lambdas, nullptr, st);
    };
    if (b->cull)
        launch_cull_paths(d_scenes, n_scenes, b->paths, b->cap_paths, b->ranges + 5, b->admissible, st, b->cull);
    // Newton kernels: block size, resident blocks per CU and queue chunk
    for (int K = 1; K <= 3; ++K) {
        if (ranges[K] == ranges[K - 1]) continue;
        palz_launch_config best = b->cfg[K - 1];
        float best_ms = 1e30f;
        for (int blk : blocks)
            for (int pc : per_cu) {
                // at least one wave per SIMD, at most 32 waves resident per CU
                if (blk * pc < 256 || blk * pc > 2048) continue;
                for (int ch : chunks) {
                    palz_launch_config c = best;
                    c.block = blk;
                    c.blocks_per_cu = pc;
                    c.chunk = ch;
                    k_reset<<<1, 64, 0, st>>>(b->ctr, b->ranges, n_lambdas);
                    (void)hipEventRecord(e0, st);
                    run_newton(K, c);
                    const float ms = elapsed();
                    if (ms < best_ms) {
                        best_ms = ms;
                        best = c;
                    }
                }
            }
        b->cfg[K - 1] = best;
        newton_ms[K - 1] = best_ms;
    }
    // the mixed-length Newton kernel: register budget (waves per SIMD), block
    // size and resident blocks per CU, with each length's tuned queue chunk
    if (b->fused == -1 || (b->fused & 1)) {
        float best_ms = 1e30f;
        palz_launch_config best = b->cfg_mixed;
        for (int waves : {2, 3})
            for (int blk : blocks)
                for (int pc : per_cu) {
                    if (blk * pc < 256 || blk * pc > 2048) continue;
                    palz_launch_config c = best;
                    c.block = blk;
                    c.blocks_per_cu = pc;
                    c.chunk = waves;
                    k_reset<<<1, 64, 0, st>>>(b->ctr, b->ranges, n_lambdas);
                    (void)hipEventRecord(e0, st);
                    b->newton_mixed(c, d_scenes, n_scenes, n_lambdas, nullptr, st);
                    const float ms = elapsed();
                    if (ms < best_ms) {
                        best_ms = ms;
                        best = c;
                    }
                }
        b->cfg_mixed = best;
    }
    // finish kernels: block size, on the converged list of the tuned Newton shapes
    k_reset<<<1, 64, 0, st>>>(b->ctr, b->ranges, n_lambdas);
    for (int K = 1; K <= 3; ++K) run_newton(K, b->cfg[K - 1]);
    b->compact(n_lambdas, nullptr, st);
    for (int K = 1; K <= 3; ++K) {
        if (ranges[K] == ranges[K - 1]) continue;
        float best_fin = 1e30f;
        palz_launch_config& best = b->cfg[K - 1];
        for (int fb : blocks) {
            palz_launch_config c = best;
            c.finish_block = fb;
            (void)hipEventRecord(e0, st);
            if (K == 1) b->occlude<1>(c, d_scenes, n_lambdas, st);
            if (K == 2) b->occlude<2>(c, d_scenes, n_lambdas, st);
            if (K == 3) b->occlude<3>(c, d_scenes, n_lambdas, st);
            if (K == 1) b->finish<1>(c, d_scenes, n_lambdas, refine, nullptr, nullptr, st);
            if (K == 2) b->finish<2>(c, d_scenes, n_lambdas, refine, nullptr, nullptr, st);
            if (K == 3) b->finish<3>(c, d_scenes, n_lambdas, refine, nullptr, nullptr, st);
            const float ms = elapsed();
            if (ms < best_fin) {
                best_fin = ms;
                best.finish_block = fb;
            }
        }
        finish_ms[K - 1] = best_fin;
    }
    // the mixed-length occlusion kernel's block size
    if (b->fused == -1 || (b->fused & 2)) {
        float best_ms = 1e30f;
        for (int fb : blocks) {
            palz_launch_config c = b->cfg_mixed;
            c.finish_block = fb;
            (void)hipEventRecord(e0, st);
            b->occlude_mixed(c, d_scenes, n_lambdas, st);
            const float ms = elapsed();
            if (ms < best_ms) {
                best_ms = ms;
                b->cfg_mixed.finish_block = fb;
            }
        }
    }
    // mode -1: the whole solve in each of the four structures (Newton and
    // occlusion each per length or fused) on the tuned shapes, the fastest kept
    if (b->fused == -1) {
        float best[4] = {1e30f, 1e30f, 1e30f, 1e30f};
        for (int rep = 0; rep < 2; ++rep)
            for (int f = 0; f < 4; ++f) {
                b->auto_fused = f;
                (void)hipEventRecord(e0, st);
                rc = b->launch_all(d_scenes, n_scenes, n_lambdas, refine, nullptr, nullptr, st, false);
                if (rc != PALZ_OK) return rc;
                best[f] = std::min(best[f], elapsed());
            }
        int pick = 0;
        for (int f = 1; f < 4; ++f)
            if (best[f] < best[pick]) pick = f;
        b->auto_fused = pick;
    }
    for (int K = 0; K < 3; ++K) {
        times_us[2 * K] = (int32_t)(newton_ms[K] * 1000.0f);
        times_us[2 * K + 1] = (int32_t)(finish_ms[K] * 1000.0f);
    }
    write_report();
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->drop_graph();
    PALZ_LAUNCH_CHECK();
    if (!b->tune_cache.empty()) {
        int32_t v[kTuneFields];
        const palz_launch_config* c4[4] = {&b->cfg[0], &b->cfg[1], &b->cfg[2], &b->cfg_mixed};
        for (int i = 0; i < 4; ++i) {
            v[4 * i] = c4[i]->block;
            v[4 * i + 1] = c4[i]->blocks_per_cu;
            v[4 * i + 2] = c4[i]->chunk;
            v[4 * i + 3] = c4[i]->finish_block;
        }
        for (int i = 0; i < 6; ++i) v[16 + i] = times_us[i];
        v[22] = b->auto_fused;
        if (!tune_cache_append(b->tune_cache, key, v))
            return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_batched_autotune: cannot append to the tile-shape cache file");
    }
    return PALZ_OK;
}

}  // extern "C"
