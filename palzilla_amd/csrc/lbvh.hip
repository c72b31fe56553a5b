// LBVH builder for gfx950: scene AABB -> Morton keys (+ fused digit counts)
// -> onesweep sort -> deltas -> Apetrei bottom-up build.
//
// Follows BVHGPUBuilder::build (minimum_lbvh.h:798-886) and the kernels of
// minimum_lbvh.cu:38-121, with these MI355X-first changes:
//   * triangle tiles are staged through LDS with 16-byte loads, so the 36-byte
//     AoS records are read with fully coalesced dwordx4 transactions;
//   * the Morton kernel also counts the sort's digits (saves the count pass's
//     8 B/key re-read);
//   * the build climbs on a compact 4-byte arrival word per node instead of the
//     node's `context` field, so the 64-byte node array needs no 0xFF memset;
//     the escape link is still written into `context`;
//   * the build launches one thread per leaf (the reference launches n-1
//     threads, minimum_lbvh.h:875, and skips the last leaf when n % 256 == 1).
#include "prims.hpp"
#include "sort.hpp"

#include <algorithm>

namespace palz {

namespace {

constexpr int kTriTile = 256;  // triangles per block in the staged kernels
constexpr uint32_t kNoArrival = 0xFFFFFFFFu;

constexpr int kBuildChunk = 512;  // sorted leaves per build workgroup
constexpr int kChunkRoots = 64;   // staged subtree roots per chunk or group
constexpr int kGroupChunks = 16;  // lower chunks per workgroup of the levels above
constexpr int kGroupItems = kGroupChunks * kChunkRoots;
constexpr uint32_t kDeviceChunk = 0xFFFFFFFFu;  // staged count of a chunk whose roots climbed device-wide

// A finished subtree whose parent lies across chunks: its node, leaf range and box.
struct alignas(16) SubtreeRoot {
    uint32_t node, lo, hi, pad_;
    float box[6];
    uint32_t pad2_[2];
};

// Per-build control word: k_morton_hist bumps the epoch that tags this build's
// arrival words, so they never need clearing.
struct BuildCtl {
    uint32_t epoch;
    uint32_t pad_[3];
};

// Stage `count` (<= kTriTile) triangles starting at `first` into LDS.  With
// `aligned` the block reads its 9216-byte window as float4s.
__device__ __forceinline__ void stage_triangles(const palz_triangle* __restrict__ tris, uint32_t first,
                                                uint32_t count, bool aligned, float* s_tri) {
    const float* src = reinterpret_cast<const float*>(tris + first);
    const uint32_t nfloat = count * 9u;
    if (aligned) {
        const float4* src4 = reinterpret_cast<const float4*>(src);
        const uint32_t n4 = nfloat / 4u;
        for (uint32_t i = threadIdx.x; i < n4; i += blockDim.x) {
            const float4 q = src4[i];
            s_tri[4 * i + 0] = q.x;
            s_tri[4 * i + 1] = q.y;
            s_tri[4 * i + 2] = q.z;
            s_tri[4 * i + 3] = q.w;
        }
        for (uint32_t i = n4 * 4u + threadIdx.x; i < nfloat; i += blockDim.x) s_tri[i] = src[i];
    } else {
        for (uint32_t i = threadIdx.x; i < nfloat; i += blockDim.x) s_tri[i] = src[i];
    }
    __syncthreads();
}

__device__ __forceinline__ float wave_cmin(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = cmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ float wave_cmax(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = cmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Float min/max through integer atomics (order-free; -0 < +0), the scheme of
// minimum_lbvh.cu:21-36.
__device__ __forceinline__ void atomic_min_float(float* addr, float v) {
    if (!__builtin_signbit(v))
        atomicMin(reinterpret_cast<int*>(addr), __float_as_int(v));
    else
        atomicMax(reinterpret_cast<unsigned int*>(addr), __float_as_uint(v));
}
__device__ __forceinline__ void atomic_max_float(float* addr, float v) {
    if (!__builtin_signbit(v))
        atomicMax(reinterpret_cast<int*>(addr), __float_as_int(v));
    else
        atomicMin(reinterpret_cast<unsigned int*>(addr), __float_as_uint(v));
}

// getSceneAABB (minimum_lbvh.cu:38-66): one pass over all vertices, read as a
// flat float stream; the component of float k is k % 3.
__global__ void __launch_bounds__(256) k_scene_aabb(const float* __restrict__ verts, uint32_t nfloat,
                                                    bool aligned, palz_aabb* __restrict__ out) {
    float lo[3] = {kFltMax, kFltMax, kFltMax};
    float hi[3] = {-kFltMax, -kFltMax, -kFltMax};
    const uint32_t stride = gridDim.x * blockDim.x;
    // 12 floats = 4 whole vertices per step, so the components line up.
    const uint32_t nchunk = aligned ? nfloat / 12u : 0u;
    const float4* v4 = reinterpret_cast<const float4*>(verts);
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nchunk; c += stride) {
        const float4 a = v4[3 * c + 0];
        const float4 b = v4[3 * c + 1];
        const float4 d = v4[3 * c + 2];
        const float f[12] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, d.x, d.y, d.z, d.w};
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            lo[k % 3] = cmin(lo[k % 3], f[k]);
            hi[k % 3] = cmax(hi[k % 3], f[k]);
        }
    }
    for (uint32_t k = nchunk * 12u + blockIdx.x * blockDim.x + threadIdx.x; k < nfloat; k += stride) {
        const float f = verts[k];
        const uint32_t c = k % 3u;
        lo[c] = cmin(lo[c], f);
        hi[c] = cmax(hi[c], f);
    }
    __shared__ float s_red[6][4];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float l = wave_cmin(lo[c]);
        const float h = wave_cmax(hi[c]);
        if (lane == 0) {
            s_red[c][wave] = l;
            s_red[3 + c][wave] = h;
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int c = threadIdx.x;
        float l = s_red[c][0], h = s_red[3 + c][0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
            l = cmin(l, s_red[c][w]);
            h = cmax(h, s_red[3 + c][w]);
        }
        atomic_min_float(&out->lower.x + c, l);
        atomic_max_float(&out->upper.x + c, h);
    }
}

// buildMortons (minimum_lbvh.cu:68-92) fused with the sort's digit count.
// Writes {key (bits 0-31), triangle index (bits 32-63)}.
__global__ void __launch_bounds__(kTriTile) k_morton_hist(const palz_triangle* __restrict__ tris,
                                                          uint32_t n, bool aligned,
                                                          const palz_aabb* __restrict__ scene,
                                                          uint32_t option, int bits,
                                                          uint64_t* __restrict__ words,
                                                          uint32_t* __restrict__ hist,
                                                          BuildCtl* __restrict__ ctl) {
    __shared__ float s_tri[kTriTile * 9];
    __shared__ uint32_t s_h[4 * 256];
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->epoch = ctl->epoch + 1u;  // the build's arrival epoch
    for (int i = threadIdx.x; i < 4 * 256; i += blockDim.x) s_h[i] = 0u;
    // Blocks loop over tiles so that each flushes its digit counts only once.
    const uint32_t ntiles = div_up(n, (uint32_t)kTriTile);
    for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const uint32_t first = tile * kTriTile;
    const uint32_t count = min((uint32_t)kTriTile, n - first);
    __syncthreads();  // the previous tile's LDS reads are done
    stage_triangles(tris, first, count, aligned, s_tri);  // syncs

    if (threadIdx.x < count) {
        const float* t = s_tri + threadIdx.x * 9;
        const v3 v0 = mk(t[0], t[1], t[2]);
        const v3 v1 = mk(t[3], t[4], t[5]);
        const v3 v2 = mk(t[6], t[7], t[8]);
        const v3 center = (v0 + v1 + v2) / 3.0f;
        const v3 lo = mk(scene->lower);
        const v3 hi = mk(scene->upper);
        uint32_t key;
        if (option & PALZ_BUILD_OPTION_USE_NORMAL) {
            // normalOf (minimum_lbvh.h:215-218) with normalize = v * (1/sqrt(v.v))
            const v3 ng = cross(v1 - v0, v2 - v0);
            const v3 nn = ng * (1.0f / sqrtf(dot(ng, ng)));
            const uint32_t nkey = morton_key(nn, mk(-1.f, -1.f, -1.f), mk(1.f, 1.f, 1.f), 32);
            const uint32_t pkey = morton_key(center, lo, hi, 32);
            key = (nkey & 0xFFFFFFC0u) | (pkey >> 26);  // minimum_lbvh.cu:85
        } else {
            key = morton_key(center, lo, hi, bits);
        }
        const uint32_t idx = first + threadIdx.x;
        words[idx] = (uint64_t)key | ((uint64_t)idx << 32);
#pragma unroll
        for (int p = 0; p < 4; ++p) atomicAdd(&s_h[p * 256 + ((key >> (8 * p)) & 0xFFu)], 1u);
    }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 4 * 256; i += blockDim.x) {
        const uint32_t c = s_h[i];
        if (c) atomicAdd(&hist[i], c);
    }
}

// Box hand-over between the two children of a node in the device-wide climb.
// On MI355X every XCD has its own L2, and an agent-scope release/acquire fence
// (the reference's __threadfence around atomicExch, minimum_lbvh.h:340-342)
// writes back and invalidates the whole L2 — once per climb step.  Instead the
// box travels in three coherent 64-bit stores (written through to the
// device-coherent level), a vmcnt wait orders them before the arrival
// exchange, and the second arriver reads the box back with coherent loads.
// Nothing else in the node is read during the build.
__device__ __forceinline__ uint64_t pack2(float a, float b) {
    return (uint64_t)__float_as_uint(a) | ((uint64_t)__float_as_uint(b) << 32);
}
__device__ __forceinline__ void store_box_coherent(palz_aabb* dst, v3 lo, v3 hi) {
    uint64_t* w = reinterpret_cast<uint64_t*>(dst);
    __hip_atomic_store(w + 0, pack2(lo.x, lo.y), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(w + 1, pack2(lo.z, hi.x), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(w + 2, pack2(hi.y, hi.z), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void load_box_coherent(palz_aabb* src, v3* lo, v3* hi) {
    uint64_t* w = reinterpret_cast<uint64_t*>(src);
    const uint64_t a = __hip_atomic_load(w + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t b = __hip_atomic_load(w + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t c = __hip_atomic_load(w + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    *lo = mk(__uint_as_float((uint32_t)a), __uint_as_float((uint32_t)(a >> 32)), __uint_as_float((uint32_t)b));
    *hi = mk(__uint_as_float((uint32_t)(b >> 32)), __uint_as_float((uint32_t)c), __uint_as_float((uint32_t)(c >> 32)));
}

// The Morton key of sorted leaf j: the low half of its word.
__device__ __forceinline__ uint32_t key_at(const uint64_t* __restrict__ words, uint32_t j) {
    return reinterpret_cast<const uint32_t*>(words)[2u * j];
}

// The device-wide climb (build_lbvh's loop, minimum_lbvh.h:293-370) from a
// finished subtree (node, leaf range, box) whose parent lies across chunks.
// Arrival words hold (epoch, far leaf end); a word from an earlier build reads
// as "no arrival".  kKeys: range-end deltas come from the sorted keys (inside
// the build kernel, where the chunk that writes them may not have run yet);
// otherwise from the delta array.
template <bool kKeys>
__device__ void climb_device(palz_node_index node, uint32_t lo, uint32_t hi, v3 blo, v3 bhi, uint32_t n,
                             const uint64_t* __restrict__ words, const uint8_t* __restrict__ deltas,
                             uint64_t* __restrict__ arrival, palz_internal_node* nodes,
                             palz_node_index* __restrict__ root, uint32_t ep) {
    const uint32_t n_int = n - 1u;
    auto split_delta = [&](uint32_t j) -> uint32_t {
        return kKeys ? delta32(key_at(words, j), key_at(words, j + 1u)) : (uint32_t)deltas[j];
    };
    bool finished_here = false;  // `node` was completed by this climb (not by its chunk)
    while (hi - lo < n_int) {
        const uint32_t dl = lo > 0u ? split_delta(lo - 1u) : 0xFFFFFFFFu;
        const uint32_t dr = hi < n_int ? split_delta(hi) : 0xFFFFFFFFu;
        const int go_left = dl < dr ? 1 : 0;
        const uint32_t parent = go_left ? lo - 1u : hi;
        palz_internal_node* pn = nodes + parent;
        pn->children[go_left] = node;
        store_box_coherent(&pn->aabbs[go_left], blo, bhi);
        if (finished_here) nodes[node].parent = parent;
        // The first of the two children to arrive stops; the second merges.
        // The range bound travels in the exchanged value itself.
        const uint64_t mine = ((uint64_t)ep << 32) | (uint64_t)(go_left ? hi : lo);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0);  // the coherent box stores have completed
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        const uint64_t old = __hip_atomic_exchange(arrival + parent, mine, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        if ((uint32_t)(old >> 32) != ep) return;
        const uint32_t other = (uint32_t)old;
        lo = min(lo, other);
        hi = max(hi, other);
        pn->context = hi < n_int ? hi : 0xFFFFFFFFu;
        node = parent;
        finished_here = true;
        v3 slo, shi;
        load_box_coherent(&pn->aabbs[go_left ^ 1], &slo, &shi);
        blo = cmin(blo, slo);
        bhi = cmax(bhi, shi);
    }
    *root = node;
    if (finished_here) nodes[node].parent = PALZ_NODE_INVALID;
}

// The count a chunk (or group) reports for its finished subtree roots:
// kDeviceChunk once they are more than its kChunkRoots slots.
__device__ __forceinline__ uint32_t staged_total(const uint32_t* s_up, uint32_t words) {
    uint32_t count = 0u;
    for (uint32_t w = 0; w < words; ++w) count += __popc(s_up[w]);
    return count > (uint32_t)kChunkRoots ? kDeviceChunk : count;
}

// Hands a finished subtree root (position `pos` of its chunk's bitmap of roots
// that go up) to the level above: staged in leaf order in the chunk's slots,
// or, when the chunk has more roots than slots, climbed device-wide at once —
// every root of such a chunk does, so that no node above mixes the LDS and the
// device-wide arrival protocols.
template <bool kKeys>
__device__ void hand_up(const uint32_t* s_up, uint32_t words_up, uint32_t pos, uint32_t slot,
                        palz_node_index node, uint32_t lo, uint32_t hi, v3 blo, v3 bhi, uint32_t n,
                        const uint64_t* __restrict__ words, const uint8_t* __restrict__ deltas,
                        uint64_t* __restrict__ arrival, palz_internal_node* nodes, palz_node_index* __restrict__ root,
                        uint32_t ep, SubtreeRoot* __restrict__ out) {
    if (staged_total(s_up, words_up) == kDeviceChunk) {
        climb_device<kKeys>(node, lo, hi, blo, bhi, n, words, deltas, arrival, nodes, root, ep);
        return;
    }
    const uint32_t word = pos >> 5;
    uint32_t rank = __popc(s_up[word] & ((1u << (pos & 31u)) - 1u));
    for (uint32_t w = 0; w < word; ++w) rank += __popc(s_up[w]);
    SubtreeRoot r;
    r.node = node;
    r.lo = lo;
    r.hi = hi;
    r.pad_ = 0u;
    r.box[0] = blo.x;
    r.box[1] = blo.y;
    r.box[2] = blo.z;
    r.box[3] = bhi.x;
    r.box[4] = bhi.y;
    r.box[5] = bhi.z;
    r.pad2_[0] = r.pad2_[1] = 0u;
    const float4* src = reinterpret_cast<const float4*>(&r);
    float4* dst = reinterpret_cast<float4*>(out + (size_t)slot * kChunkRoots + rank);
#pragma unroll
    for (int q = 0; q < 3; ++q) dst[q] = src[q];
}

// computeDeltas + build (minimum_lbvh.cu:94-121, minimum_lbvh.h:293-370), one
// workgroup per chunk of kBuildChunk sorted leaves.
//
// Node p (the split between leaves p and p+1) is "local" when both its
// children lie in the chunk, i.e. its leaf range [lo, hi] does; that range is
// fixed by the deltas alone (lo - 1 = nearest split on the left with delta >=
// delta[p], hi = nearest split on the right with delta > delta[p]), so prefix
// / suffix maxima of the chunk's deltas in LDS decide it before the climb.
// Local nodes hand their first child over in LDS and their records are written
// out whole at the end of the chunk; no local node needs device-wide
// coherence.  A climb that reaches a non-local node stops: its subtree root is
// staged (at most kChunkRoots per chunk, in leaf order) for k_build_level,
// whose workgroups run the same climb over the roots of kGroupChunks chunks,
// level by level, until one group holds them all.  A chunk or group with more
// roots than slots sends them up the device-wide climb instead (and so does
// every group above it).  For Morton-ordered input a chunk has a few roots;
// long runs of equal keys (right-leaning chains) make more of them.  Both
// children of a node always agree on how it is reached, and the boxes are
// exact min/max, so the tree is the bit-identical result of the
// one-thread-per-leaf climb.
struct ChunkShared {
    palz_internal_node node[kBuildChunk];  // records of the local nodes (read as float4)
    uint8_t delta[kBuildChunk + 2];        // splits c0-1 .. c_last at [0 .. cnt+1]
    uint8_t pre[kBuildChunk];              // max of the chunk's split deltas up to t
    uint8_t suf[kBuildChunk];              // max of the chunk's split deltas from t on
    uint8_t local[kBuildChunk];
    uint32_t key[kBuildChunk + 2];         // keys c0-1 .. c_last+1 at [0 .. cnt+2]
    uint32_t flag[kBuildChunk];            // arrival words of the local nodes
    uint32_t up[kBuildChunk / 32];         // leaves whose subtree roots go up
    uint32_t wave[2][kBuildChunk / 64];
};

__global__ void __launch_bounds__(kBuildChunk)
    k_build(const palz_triangle* __restrict__ tris, const uint64_t* __restrict__ words, uint32_t n,
            uint8_t* __restrict__ deltas, palz_internal_node* nodes, palz_node_index* __restrict__ root,
            uint64_t* __restrict__ arrival, const BuildCtl* __restrict__ ctl,
            SubtreeRoot* __restrict__ staged, uint32_t* __restrict__ staged_count) {
    __shared__ ChunkShared s;
    const uint32_t n_int = n - 1u;
    const uint32_t c0 = blockIdx.x * kBuildChunk;
    const uint32_t c_last = min(c0 + (uint32_t)kBuildChunk, n) - 1u;  // last leaf of the chunk
    const uint32_t cnt = c_last - c0;                                   // splits c0 .. c_last-1
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    const uint32_t it = c0 + t;
    const bool active = it <= c_last;

    uint32_t tri_index = 0u;
    if (active) {
        const uint64_t w = words[it];
        tri_index = (uint32_t)(w >> 32);
        s.key[t + 1u] = (uint32_t)w;
    }
    if (t == 0u && c0 > 0u) s.key[0] = key_at(words, c0 - 1u);
    if (t == 1u && c_last < n_int) s.key[cnt + 2u] = key_at(words, c_last + 1u);
    s.flag[t] = kNoArrival;
    if (t < kBuildChunk / 32) s.up[t] = 0u;
    __syncthreads();
    // the delta of split c0 + t (t <= cnt) and of split c0 - 1
    uint32_t d = 0u;
    if (active && it < n_int) {
        d = delta32(s.key[t + 1u], s.key[t + 2u]);
        deltas[it] = (uint8_t)d;
        s.delta[t + 1u] = (uint8_t)d;
    }
    if (t == 0u && c0 > 0u) s.delta[0] = (uint8_t)delta32(s.key[0], s.key[1]);
    palz_triangle tr;
    if (active) tr = tris[tri_index];
    // prefix and suffix maxima of the split deltas c0 .. c_last-1: within the
    // wave by shuffles, then across waves
    if (t >= cnt) d = 0u;
    uint32_t pm = d, sm = d;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t x = (uint32_t)__shfl_up((int)pm, o, 64);
        const uint32_t y = (uint32_t)__shfl_down((int)sm, o, 64);
        if (lane >= (uint32_t)o) pm = max(pm, x);
        if (lane + o < 64u) sm = max(sm, y);
    }
    if (lane == 63u) s.wave[0][wave] = pm;
    if (lane == 0u) s.wave[1][wave] = sm;
    __syncthreads();
    for (uint32_t w = 0; w < wave; ++w) pm = max(pm, s.wave[0][w]);
    for (uint32_t w = wave + 1; w < kBuildChunk / 64; ++w) sm = max(sm, s.wave[1][w]);
    s.pre[t] = (uint8_t)pm;
    s.suf[t] = (uint8_t)sm;
    __syncthreads();
    if (t < cnt) {
        const uint32_t dp = s.delta[t + 1u];
        const bool left = (t > 0u && s.pre[t - 1u] >= dp) || c0 == 0u || s.delta[0] >= dp;
        const bool right = (t + 1u < cnt && s.suf[t + 1u] > dp) || c_last == n_int || s.delta[cnt + 1u] > dp;
        s.local[t] = (left && right) ? 1 : 0;
    }
    __syncthreads();

    uint32_t lo = it, hi = it;  // leaf range
    palz_node_index node = tri_index | PALZ_NODE_LEAF_BIT;
    v3 blo = mk(0.f, 0.f, 0.f), bhi = blo;
    bool go_up = false;
    if (active) {
        blo = mk(tr.vs[0]);
        bhi = mk(tr.vs[0]);
        blo = cmin(blo, mk(tr.vs[1]));
        bhi = cmax(bhi, mk(tr.vs[1]));
        blo = cmin(blo, mk(tr.vs[2]));
        bhi = cmax(bhi, mk(tr.vs[2]));
        auto delta_at = [&](uint32_t j) -> uint32_t { return s.delta[j - c0 + 1u]; };  // j in [c0-1, c_last]
        while (hi - lo < n_int) {
            const uint32_t dl = lo > 0u ? delta_at(lo - 1u) : 0xFFFFFFFFu;
            const uint32_t dr = hi < n_int ? delta_at(hi) : 0xFFFFFFFFu;
            const int go_left = dl < dr ? 1 : 0;
            const uint32_t parent = go_left ? lo - 1u : hi;
            if (!(parent - c0 < cnt && s.local[parent - c0])) {
                go_up = true;
                break;
            }
            palz_internal_node& rec = s.node[parent - c0];
            rec.children[go_left] = node;
            rec.aabbs[go_left].lower = {blo.x, blo.y, blo.z};
            rec.aabbs[go_left].upper = {bhi.x, bhi.y, bhi.z};
            const uint32_t mine = go_left ? hi : lo;
            const uint32_t other = __hip_atomic_exchange(&s.flag[parent - c0], mine, __ATOMIC_ACQ_REL,
                                                         __HIP_MEMORY_SCOPE_WORKGROUP);
            if (other == kNoArrival) break;
            const palz_aabb sb = rec.aabbs[go_left ^ 1];
            lo = min(lo, other);
            hi = max(hi, other);
            uint32_t grand = PALZ_NODE_INVALID;
            if (hi - lo < n_int) {
                const uint32_t gl = lo > 0u ? delta_at(lo - 1u) : 0xFFFFFFFFu;
                const uint32_t gr = hi < n_int ? delta_at(hi) : 0xFFFFFFFFu;
                grand = gl < gr ? lo - 1u : hi;
            }
            rec.parent = grand;
            rec.context = hi < n_int ? hi : 0xFFFFFFFFu;
            node = parent;
            blo = cmin(blo, mk(sb.lower));
            bhi = cmax(bhi, mk(sb.upper));
        }
        if (go_up)
            atomicOr(&s.up[t >> 5], 1u << (t & 31u));
        else if (hi - lo == n_int)
            *root = node;
    }
    __syncthreads();
    if (go_up) hand_up<true>(s.up, kBuildChunk / 32, t, blockIdx.x, node, lo, hi, blo, bhi, n, words, nullptr,
                             arrival, nodes, root, ctl->epoch, staged);
    if (t == 0u && staged_count) staged_count[blockIdx.x] = staged_total(s.up, kBuildChunk / 32);
    // the chunk's local records, written out whole
    for (uint32_t i = t; i < cnt; i += blockDim.x) {
        if (!s.local[i]) continue;
        const float4* src = reinterpret_cast<const float4*>(&s.node[i]);
        float4* dst = reinterpret_cast<float4*>(nodes + c0 + i);
#pragma unroll
        for (int q = 0; q < 4; ++q) dst[q] = src[q];
    }
}

// The levels above the first: workgroup g takes the staged subtree roots of
// lower chunks [16g, 16g + 16) (at most kGroupItems, in leaf order) and runs the
// same LDS climb on them, with the split after each item at the item's last
// leaf.  A node is local to the group when its leaf range lies within the
// group's items.  A group with a lower chunk whose roots climbed device-wide
// sends all of its items device-wide too (the nodes above them have a
// device-wide child).  The last level has one group, which reaches the root.
struct GroupShared {
    palz_internal_node node[kGroupItems];  // records of the local nodes (read as float4)
    uint8_t delta[kGroupItems + 1];        // split before item 0, after items 0 .. M-1
    uint8_t pre[kGroupItems];
    uint8_t suf[kGroupItems];
    uint8_t local[kGroupItems];
    uint32_t flag[kGroupItems];
    uint32_t leaf_hi[kGroupItems];         // leaf-level split after each item
    uint32_t up[kGroupItems / 32];
    uint32_t wave[2][kGroupItems / 64];
    uint32_t offset[kGroupChunks + 1];
    uint32_t any_device, first_lo, last_hi;
};

__global__ void __launch_bounds__(kGroupItems)
    k_build_level(const SubtreeRoot* __restrict__ in, const uint32_t* __restrict__ in_count, uint32_t in_chunks,
                  uint32_t n, const uint8_t* __restrict__ deltas, palz_internal_node* nodes,
                  palz_node_index* __restrict__ root, uint64_t* __restrict__ arrival,
                  const BuildCtl* __restrict__ ctl, SubtreeRoot* __restrict__ out, uint32_t* __restrict__ out_count) {
    __shared__ GroupShared s;
    const uint32_t n_int = n - 1u;
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    const uint32_t g = blockIdx.x;
    const uint32_t k0 = g * kGroupChunks;
    const uint32_t nlow = min((uint32_t)kGroupChunks, in_chunks - k0);
    const uint32_t ep = ctl->epoch;
    if (t == 0u) {
        uint32_t acc = 0u, dev = 0u;
        for (uint32_t k = 0; k < nlow; ++k) {
            const uint32_t c = in_count[k0 + k];
            s.offset[k] = acc;
            if (c == kDeviceChunk)
                dev = 1u;
            else
                acc += c;
        }
        s.offset[nlow] = acc;
        s.any_device = dev;
    }
    s.flag[t] = kNoArrival;
    if (t < kGroupItems / 32) s.up[t] = 0u;
    __syncthreads();
    if (s.any_device) {
        const uint32_t k = t / (uint32_t)kChunkRoots, r = t % (uint32_t)kChunkRoots;
        if (k < nlow) {
            const uint32_t c = in_count[k0 + k];
            if (c != kDeviceChunk && r < c) {
                const SubtreeRoot x = in[(size_t)(k0 + k) * kChunkRoots + r];
                climb_device<false>(x.node, x.lo, x.hi, mk(x.box[0], x.box[1], x.box[2]),
                                    mk(x.box[3], x.box[4], x.box[5]), n, nullptr, deltas, arrival, nodes, root, ep);
            }
        }
        if (t == 0u && out_count) out_count[g] = kDeviceChunk;
        return;
    }
    const uint32_t M = s.offset[nlow];  // >= nlow: every lower chunk has a root
    const bool active = t < M;
    const uint32_t cnt = M - 1u;        // item splits 0 .. M-2
    uint32_t lo = 0u, hi = 0u;
    palz_node_index node = 0u;
    v3 blo = mk(0.f, 0.f, 0.f), bhi = blo;
    uint32_t d = 0u;
    if (active) {
        uint32_t k = 0u;
        while (s.offset[k + 1u] <= t) ++k;
        const SubtreeRoot x = in[(size_t)(k0 + k) * kChunkRoots + (t - s.offset[k])];
        node = x.node;
        lo = x.lo;
        hi = x.hi;
        blo = mk(x.box[0], x.box[1], x.box[2]);
        bhi = mk(x.box[3], x.box[4], x.box[5]);
        s.leaf_hi[t] = hi;
        if (hi < n_int) {
            d = deltas[hi];
            s.delta[t + 1u] = (uint8_t)d;
        }
        if (t == 0u) {
            s.first_lo = lo;
            if (lo > 0u) s.delta[0] = deltas[lo - 1u];
        }
        if (t == cnt) s.last_hi = hi;
    }
    // prefix and suffix maxima of the item split deltas 0 .. M-2
    if (t >= cnt) d = 0u;
    uint32_t pm = d, sm = d;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t x = (uint32_t)__shfl_up((int)pm, o, 64);
        const uint32_t y = (uint32_t)__shfl_down((int)sm, o, 64);
        if (lane >= (uint32_t)o) pm = max(pm, x);
        if (lane + o < 64u) sm = max(sm, y);
    }
    if (lane == 63u) s.wave[0][wave] = pm;
    if (lane == 0u) s.wave[1][wave] = sm;
    __syncthreads();
    for (uint32_t w = 0; w < wave; ++w) pm = max(pm, s.wave[0][w]);
    for (uint32_t w = wave + 1; w < kGroupItems / 64; ++w) sm = max(sm, s.wave[1][w]);
    s.pre[t] = (uint8_t)pm;
    s.suf[t] = (uint8_t)sm;
    __syncthreads();
    if (t < cnt) {
        const uint32_t dp = s.delta[t + 1u];
        const bool left = (t > 0u && s.pre[t - 1u] >= dp) || s.first_lo == 0u || s.delta[0] >= dp;
        const bool right = (t + 1u < cnt && s.suf[t + 1u] > dp) || s.last_hi == n_int || s.delta[M] > dp;
        s.local[t] = (left && right) ? 1 : 0;
    }
    __syncthreads();

    bool go_up = false;
    if (active) {
        uint32_t a = t, bb = t;  // item range, in the group's numbering
        while (hi - lo < n_int) {
            const uint32_t dl = lo > 0u ? s.delta[a] : 0xFFFFFFFFu;         // split a-1
            const uint32_t dr = hi < n_int ? s.delta[bb + 1u] : 0xFFFFFFFFu;  // split bb
            const int go_left = dl < dr ? 1 : 0;
            const uint32_t parent = go_left ? a - 1u : bb;  // item split (a - 1 wraps for a = 0)
            if (!(parent < cnt && s.local[parent])) {
                go_up = true;
                break;
            }
            palz_internal_node& rec = s.node[parent];
            rec.children[go_left] = node;
            rec.aabbs[go_left].lower = {blo.x, blo.y, blo.z};
            rec.aabbs[go_left].upper = {bhi.x, bhi.y, bhi.z};
            // The far leaf end travels in the field the other child leaves alone
            // (the second arriver overwrites both below).
            if (go_left)
                rec.context = hi;
            else
                rec.parent = lo;
            const uint32_t split_leaf = go_left ? lo - 1u : hi;
            const uint32_t mine = go_left ? bb : a;
            const uint32_t other = __hip_atomic_exchange(&s.flag[parent], mine, __ATOMIC_ACQ_REL,
                                                         __HIP_MEMORY_SCOPE_WORKGROUP);
            if (other == kNoArrival) break;
            const palz_aabb sb = rec.aabbs[go_left ^ 1];
            if (go_left) {
                a = other;
                lo = rec.parent;
            } else {
                bb = other;
                hi = rec.context;
            }
            uint32_t grand = PALZ_NODE_INVALID;
            if (hi - lo < n_int) {
                const uint32_t gl = lo > 0u ? s.delta[a] : 0xFFFFFFFFu;
                const uint32_t gr = hi < n_int ? s.delta[bb + 1u] : 0xFFFFFFFFu;
                grand = gl < gr ? lo - 1u : hi;
            }
            rec.parent = grand;
            rec.context = hi < n_int ? hi : 0xFFFFFFFFu;
            node = split_leaf;
            blo = cmin(blo, mk(sb.lower));
            bhi = cmax(bhi, mk(sb.upper));
        }
        if (go_up)
            atomicOr(&s.up[t >> 5], 1u << (t & 31u));
        else if (hi - lo == n_int)
            *root = node;
    }
    __syncthreads();
    if (out_count) {
        if (go_up)
            hand_up<false>(s.up, kGroupItems / 32, t, g, node, lo, hi, blo, bhi, n, nullptr, deltas, arrival, nodes,
                           root, ep, out);
        if (t == 0u) out_count[g] = staged_total(s.up, kGroupItems / 32);
    }
    // the group's local records, written out whole
    for (uint32_t i = t; i < cnt; i += blockDim.x) {
        if (!s.local[i]) continue;
        const float4* src = reinterpret_cast<const float4*>(&s.node[i]);
        float4* dst = reinterpret_cast<float4*>(nodes + s.leaf_hi[i]);
#pragma unroll
        for (int q = 0; q < 4; ++q) dst[q] = src[q];
    }
}

// palz_wide_node of every internal node: its record and its internal
// children's records, merged, the six boxes stored by plane.
__global__ void __launch_bounds__(256) k_widen(const palz_internal_node* __restrict__ nodes, uint32_t n_int,
                                               palz_wide_node* __restrict__ wide) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_int) return;
    const palz_internal_node nd = nodes[i];
    palz_wide_node w;
    w.pad_[0] = w.pad_[1] = 0u;
    palz_aabb box[6];
    const palz_aabb none = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        w.child[c] = nd.children[c];
        box[c] = nd.aabbs[c];
        if (nd.children[c] & PALZ_NODE_LEAF_BIT) {
            w.grand[2 * c] = w.grand[2 * c + 1] = PALZ_NODE_INVALID;
            box[2 + 2 * c] = box[3 + 2 * c] = none;
        } else {
            const palz_internal_node cn = nodes[nd.children[c]];
            w.grand[2 * c] = cn.children[0];
            w.grand[2 * c + 1] = cn.children[1];
            box[2 + 2 * c] = cn.aabbs[0];
            box[3 + 2 * c] = cn.aabbs[1];
        }
    }
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const palz_aabb b = box[s];
        const float lo[3] = {b.lower.x, b.lower.y, b.lower.z}, hi[3] = {b.upper.x, b.upper.y, b.upper.z};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            w.plane[a][0][s] = lo[a];
            w.plane[a][1][s] = hi[a];
            w.plane[a][2][s] = lo[a];
        }
    }
    wide[i] = w;
}

// InternalNormalBound pass (main.cpp:248-295): the counters start at zero.
__global__ void __launch_bounds__(256) k_normal_bounds_clear(palz_normal_bound* __restrict__ out, uint32_t n_int) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_int) out[i].counter = 0u;
}

// One thread per internal node climbs from each of its leaf children towards
// the root: it stores the subtree's normal box into the parent's slot for
// that child, and the second of the two arrivals at a node merges the
// sibling's box and goes on.  The box travels through memory written through
// to the device-coherent level (32-bit coherent stores: the 52-byte records
// are only 4-byte aligned), a vmcnt wait orders it before the arrival count,
// and the second arriver reads it back with coherent loads.
__device__ __forceinline__ void store_nbox(palz_aabb* dst, v3 lo, v3 hi) {
    uint32_t* w = reinterpret_cast<uint32_t*>(dst);
    const float f[6] = {lo.x, lo.y, lo.z, hi.x, hi.y, hi.z};
#pragma unroll
    for (int k = 0; k < 6; ++k)
        __hip_atomic_store(w + k, __float_as_uint(f[k]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void load_nbox(palz_aabb* src, v3* lo, v3* hi) {
    uint32_t* w = reinterpret_cast<uint32_t*>(src);
    float f[6];
#pragma unroll
    for (int k = 0; k < 6; ++k)
        f[k] = __uint_as_float(__hip_atomic_load(w + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    *lo = mk(f[0], f[1], f[2]);
    *hi = mk(f[3], f[4], f[5]);
}

__global__ void __launch_bounds__(256) k_normal_bounds(const palz_internal_node* __restrict__ nodes, uint32_t n_int,
                                                       const palz_triangle_attrib* __restrict__ attribs,
                                                       palz_normal_bound* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_int) return;
    for (int j = 0; j < 2; ++j) {
        palz_node_index me = nodes[i].children[j];
        if (!(me & PALZ_NODE_LEAF_BIT)) continue;
        const palz_triangle_attrib& at = attribs[me & ~PALZ_NODE_LEAF_BIT];
        v3 lo = mk(at.shading_normals[0]), hi = lo;
        lo = cmin(lo, mk(at.shading_normals[1]));
        hi = cmax(hi, mk(at.shading_normals[1]));
        lo = cmin(lo, mk(at.shading_normals[2]));
        hi = cmax(hi, mk(at.shading_normals[2]));
        uint32_t parent = i;
        while (parent != PALZ_NODE_INVALID) {
            const int c = nodes[parent].children[0] == me ? 0 : 1;
            store_nbox(&out[parent].normal_bounds[c], lo, hi);
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            __builtin_amdgcn_s_waitcnt(0);  // the coherent box stores have completed
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            const uint32_t v = __hip_atomic_fetch_add(&out[parent].counter, 1u, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT);
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            if (v == 0u) break;
            v3 slo, shi;
            load_nbox(&out[parent].normal_bounds[c ^ 1], &slo, &shi);
            lo = cmin(lo, slo);
            hi = cmax(hi, shi);
            me = parent;
            parent = nodes[me].parent;
        }
    }
}

}  // namespace

}  // namespace palz

struct palz_lbvh {
    palz_internal_node* nodes = nullptr;
    palz_node_index* root = nullptr;
    palz_aabb* scene = nullptr;
    uint64_t* words = nullptr;
    uint64_t* tmp = nullptr;
    uint8_t* deltas = nullptr;
    uint64_t* arrival = nullptr;  // per split: (epoch, far leaf end) of the device-wide climb
    palz::SubtreeRoot* staged[2] = {};   // per chunk / group, the subtree roots that go up (ping-pong)
    uint32_t* staged_count[2] = {};
    palz::BuildCtl* ctl = nullptr;
    palz_wide_node* wide = nullptr;  // BVH4 view, on request
    uint32_t wide_cap = 0;
    uint32_t capacity = 0;
    int n = 0;
    // Optional stage timing: events after each stage of the last build.
    static constexpr int kStages = 8;
    bool timing = false;
    hipEvent_t ev[kStages + 1] = {};

    int ensure(uint32_t want) {
        if (want <= capacity) return PALZ_OK;
        release_arrays();
        PALZ_HIP(hipMalloc(&nodes, sizeof(palz_internal_node) * std::max(1u, want - 1u)));
        PALZ_HIP(hipMalloc(&words, sizeof(uint64_t) * want));
        PALZ_HIP(hipMalloc(&tmp, sizeof(uint64_t) * want));
        PALZ_HIP(hipMalloc(&deltas, std::max(4u, want)));
        // arrival words start at epoch 0, which no build uses
        PALZ_HIP(hipMalloc(&arrival, sizeof(uint64_t) * std::max(1u, want)));
        PALZ_HIP(hipMemset(arrival, 0, sizeof(uint64_t) * std::max(1u, want)));
        const uint32_t chunks = palz::div_up(want, (uint32_t)palz::kBuildChunk);
        const uint32_t groups = palz::div_up(chunks, (uint32_t)palz::kGroupChunks);
        PALZ_HIP(hipMalloc(&staged[0], sizeof(palz::SubtreeRoot) * palz::kChunkRoots * chunks));
        PALZ_HIP(hipMalloc(&staged[1], sizeof(palz::SubtreeRoot) * palz::kChunkRoots * groups));
        PALZ_HIP(hipMalloc(&staged_count[0], sizeof(uint32_t) * chunks));
        PALZ_HIP(hipMalloc(&staged_count[1], sizeof(uint32_t) * groups));
        capacity = want;
        return PALZ_OK;
    }
    void release_arrays() {
        for (void* p : {(void*)nodes, (void*)words, (void*)tmp, (void*)deltas, (void*)arrival, (void*)staged[0],
                        (void*)staged[1], (void*)staged_count[0], (void*)staged_count[1]})
            if (p) (void)hipFree(p);
        staged[0] = staged[1] = nullptr;
        staged_count[0] = staged_count[1] = nullptr;
        nodes = nullptr;
        words = tmp = nullptr;
        deltas = nullptr;
        arrival = nullptr;
        capacity = 0;
    }
    ~palz_lbvh() {
        for (hipEvent_t& e : ev)
            if (e) (void)hipEventDestroy(e);
        release_arrays();
        if (root) (void)hipFree(root);
        if (scene) (void)hipFree(scene);
        if (ctl) (void)hipFree(ctl);
        if (wide) (void)hipFree(wide);
    }
};

extern "C" {

int palz_lbvh_create(palz_lbvh** out) {
    if (!out) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_lbvh_create: out is NULL");
    palz_lbvh* b = new palz_lbvh();
    hipError_t e = hipMalloc(&b->root, sizeof(palz_node_index));
    if (e == hipSuccess) e = hipMalloc(&b->scene, sizeof(palz_aabb));
    if (e == hipSuccess) e = hipMalloc(&b->ctl, sizeof(palz::BuildCtl));
    if (e == hipSuccess) e = hipMemset(b->ctl, 0, sizeof(palz::BuildCtl));
    if (e != hipSuccess) {
        delete b;
        return palz::hip_fail(e, "hipMalloc", __FILE__, __LINE__);
    }
    *out = b;
    return PALZ_OK;
}

int palz_lbvh_destroy(palz_lbvh* b) {
    delete b;
    return PALZ_OK;
}

int palz_lbvh_build(palz_lbvh* b, const palz_triangle* d_tris, int n, uint32_t build_option,
                    int morton_bits, palz_sorter* sorter, void* stream) {
    using namespace palz;
    if (!b || !sorter) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_lbvh_build: NULL handle");
    if (n < 1 || !d_tris) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_lbvh_build: need n >= 1 triangles");
    if ((uint32_t)n > kSortMaxN) return fail(PALZ_ERR_CAPACITY, "palz_lbvh_build: n must be < 2^30");
    if (morton_bits != 32 && morton_bits != 30)
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_lbvh_build: morton_bits must be 32 or 30");
    if ((build_option & PALZ_BUILD_OPTION_USE_NORMAL) && morton_bits != 32)
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_lbvh_build: USE_NORMAL needs 32-bit keys");
    if ((reinterpret_cast<uintptr_t>(d_tris) & 3u) != 0u)
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_lbvh_build: triangles must be 4-byte aligned");

    const uint32_t un = (uint32_t)n;
    int rc = b->ensure(un);
    if (rc != PALZ_OK) return rc;
    rc = sorter->reserve(un);
    if (rc != PALZ_OK) return rc;
    b->n = n;

    hipStream_t st = (hipStream_t)stream;
    const bool aligned = (reinterpret_cast<uintptr_t>(d_tris) & 15u) == 0u;
    const SortScratch& s = sorter->scratch;

    hipEvent_t* ev = b->timing ? b->ev : nullptr;
    if (ev) PALZ_HIP(hipEventRecord(ev[0], st));
    launch_sort_prepare(s, b->scene, st);
    {
        const uint32_t nfloat = un * 9u;
        const uint32_t blocks = std::max(1u, std::min(div_up(nfloat / 12u + 1u, 256u), 2048u));
        k_scene_aabb<<<blocks, 256, 0, st>>>(reinterpret_cast<const float*>(d_tris), nfloat, aligned,
                                             b->scene);
    }
    if (ev) PALZ_HIP(hipEventRecord(ev[1], st));
    k_morton_hist<<<std::min(div_up(un, (uint32_t)kTriTile), 2048u), kTriTile, 0, st>>>(
        d_tris, un, aligned, b->scene, build_option, morton_bits, b->words, s.hist, b->ctl);
    if (ev) PALZ_HIP(hipEventRecord(ev[2], st));
    const int npass = launch_sort_passes(s, b->words, b->tmp, nullptr, nullptr, un, 0, 32, st,
                                         ev ? ev + 3 : nullptr);
    (void)npass;  // 4 passes: the sorted words are back in b->words
    // level 1 over chunks of leaves, then one launch per level of groups of
    // kGroupChunks lower chunks, until one group holds everything
    uint32_t chunks = div_up(un, (uint32_t)kBuildChunk);
    k_build<<<chunks, kBuildChunk, 0, st>>>(d_tris, b->words, un, b->deltas, b->nodes, b->root, b->arrival, b->ctl,
                                            b->staged[0], chunks > 1u ? b->staged_count[0] : nullptr);
    if (ev) PALZ_HIP(hipEventRecord(ev[7], st));
    for (int cur = 0; chunks > 1u; cur ^= 1) {
        const uint32_t groups = div_up(chunks, (uint32_t)kGroupChunks);
        k_build_level<<<groups, kGroupItems, 0, st>>>(b->staged[cur], b->staged_count[cur], chunks, un, b->deltas,
                                                      b->nodes, b->root, b->arrival, b->ctl, b->staged[cur ^ 1],
                                                      groups > 1u ? b->staged_count[cur ^ 1] : nullptr);
        chunks = groups;
    }
    PALZ_LAUNCH_CHECK();
    if (ev) PALZ_HIP(hipEventRecord(ev[8], st));
    PALZ_LAUNCH_CHECK();
    return PALZ_OK;
}

int palz_lbvh_result(palz_lbvh* b, palz_internal_node** d_nodes, palz_node_index** d_root,
                     palz_aabb** d_scene_aabb, int* n_triangles) {
    if (!b) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_lbvh_result: NULL handle");
    if (d_nodes) *d_nodes = b->nodes;
    if (d_root) *d_root = b->root;
    if (d_scene_aabb) *d_scene_aabb = b->scene;
    if (n_triangles) *n_triangles = b->n;
    return PALZ_OK;
}

int palz_lbvh_widen(palz_lbvh* b, palz_wide_node** d_wide, void* stream) {
    using namespace palz;
    if (!b || !d_wide) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_lbvh_widen: NULL");
    if (b->n < 1) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_lbvh_widen: nothing built yet");
    const uint32_t n_int = (uint32_t)b->n - 1u;
    const uint32_t want = std::max(1u, n_int);
    if (want > b->wide_cap) {
        if (b->wide) PALZ_HIP(hipFree(b->wide));
        b->wide = nullptr;
        PALZ_HIP(hipMalloc(&b->wide, sizeof(palz_wide_node) * want));
        b->wide_cap = want;
    }
    if (n_int > 0u) {
        k_widen<<<div_up(n_int, 256u), 256, 0, (hipStream_t)stream>>>(b->nodes, n_int, b->wide);
        PALZ_LAUNCH_CHECK();
    }
    *d_wide = b->wide;
    return PALZ_OK;
}

int palz_lbvh_normal_bounds(palz_lbvh* b, const palz_triangle_attrib* d_attribs, palz_normal_bound* d_out,
                            void* stream) {
    using namespace palz;
    if (!b || !d_attribs || !d_out) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_lbvh_normal_bounds: NULL");
    if (b->n < 1) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_lbvh_normal_bounds: nothing built yet");
    const uint32_t n_int = (uint32_t)b->n - 1u;
    if (n_int == 0u) return PALZ_OK;
    hipStream_t st = (hipStream_t)stream;
    k_normal_bounds_clear<<<div_up(n_int, 256u), 256, 0, st>>>(d_out, n_int);
    k_normal_bounds<<<div_up(n_int, 256u), 256, 0, st>>>(b->nodes, n_int, d_attribs, d_out);
    PALZ_LAUNCH_CHECK();
    return PALZ_OK;
}

int palz_lbvh_set_timing(palz_lbvh* b, int enable) {
    if (!b) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_lbvh_set_timing: NULL handle");
    if (enable && !b->ev[0])
        for (hipEvent_t& e : b->ev) PALZ_HIP(hipEventCreate(&e));
    b->timing = enable != 0;
    return PALZ_OK;
}

int palz_lbvh_timings(palz_lbvh* b, float* ms, int max_stages, int* n_stages) {
    if (!b || !ms) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_lbvh_timings: NULL");
    if (!b->timing) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_lbvh_timings: timing is off");
    const int k = std::min(max_stages, palz_lbvh::kStages);
    PALZ_HIP(hipEventSynchronize(b->ev[palz_lbvh::kStages]));
    for (int i = 0; i < k; ++i) PALZ_HIP(hipEventElapsedTime(ms + i, b->ev[i], b->ev[i + 1]));
    if (n_stages) *n_stages = k;
    return PALZ_OK;
}

int palz_lbvh_sorted(palz_lbvh* b, uint64_t** d_words, uint8_t** d_deltas) {
    if (!b) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_lbvh_sorted: NULL handle");
    if (d_words) *d_words = b->words;
    if (d_deltas) *d_deltas = b->deltas;
    return PALZ_OK;
}

}  // extern "C"
