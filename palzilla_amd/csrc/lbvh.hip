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
                                                          uint32_t* __restrict__ hist) {
    __shared__ float s_tri[kTriTile * 9];
    __shared__ uint32_t s_h[4 * 256];
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

// computeDeltas (minimum_lbvh.cu:94-102), four deltas per thread, plus the
// arrival words of the build.
__global__ void __launch_bounds__(256) k_deltas(const uint64_t* __restrict__ words, uint32_t n,
                                                uint8_t* __restrict__ deltas,
                                                uint32_t* __restrict__ arrival) {
    const uint32_t m = n - 1u;  // number of deltas / internal nodes
    const uint32_t i0 = (blockIdx.x * blockDim.x + threadIdx.x) * 4u;
    if (i0 >= m) return;
    if (i0 + 4u <= m) {
        uint32_t k[5];
#pragma unroll
        for (int j = 0; j < 5; ++j) k[j] = (uint32_t)words[i0 + j];
        uint32_t packed = 0u;
#pragma unroll
        for (int j = 0; j < 4; ++j) packed |= delta32(k[j], k[j + 1]) << (8 * j);
        if ((i0 & 3u) == 0u && (reinterpret_cast<uintptr_t>(deltas) & 3u) == 0u) {
            *reinterpret_cast<uint32_t*>(deltas + i0) = packed;
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) deltas[i0 + j] = (uint8_t)(packed >> (8 * j));
        }
        *reinterpret_cast<uint4*>(arrival + i0) = make_uint4(kNoArrival, kNoArrival, kNoArrival, kNoArrival);
    } else {
        for (uint32_t i = i0; i < m; ++i) {
            deltas[i] = (uint8_t)delta32((uint32_t)words[i], (uint32_t)words[i + 1]);
            arrival[i] = kNoArrival;
        }
    }
}

// Box hand-over between the two children of a node.  On MI355X every XCD has
// its own L2, and an agent-scope release/acquire fence (the reference's
// __threadfence around atomicExch, minimum_lbvh.h:340-342) writes back and
// invalidates the whole L2 (buffer_wbl2 / buffer_inv) — once per climb step.
// Instead the box travels in three coherent 64-bit stores (sc1: written
// through to the device-coherent level), a vmcnt wait orders them before the
// arrival exchange, and the second arriver reads the box back with coherent
// loads.  Nothing else in the node is read during the build.
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

// build + build_lbvh (minimum_lbvh.cu:104-121, minimum_lbvh.h:293-370), in
// levels of workgroup-local climbs.
//
// Level 1 climbs from the sorted leaves, one workgroup per chunk of
// kBuildChunk consecutive leaves.  Node p (the split between leaves p and
// p+1) is "local" when both its children lie in the chunk, i.e. its leaf range
// [lo, hi] does; that range is fixed by the deltas alone (lo - 1 = nearest
// split on the left with delta >= delta[p], hi = nearest split on the right
// with delta > delta[p]), so prefix / suffix maxima of the chunk's deltas in
// LDS decide it before the climb.  Local nodes hand their first child over in
// LDS, and their records are written out whole at the end of the chunk.  A
// climb that reaches a non-local node stops there: its subtree root becomes an
// item of the next level.  Those items, in leaf order, are the leaves of the
// tree above (the non-local nodes split exactly at the boundaries between
// them, with the same deltas), so the next level runs the same climb on them,
// until one chunk holds every item.  No level needs device-wide coherence,
// and both children of a node always agree on its level; the boxes are exact
// min/max, so the tree is the bit-identical result of the one-thread-per-leaf
// climb.
constexpr int kBuildChunk = 512;

// A subtree root handed to the next level: its node, leaf range and box.
struct BuildItem {
    uint32_t node, lo, hi, pad_;
    float box[6];
};

template <bool kLeaves>
__global__ void __launch_bounds__(kBuildChunk)
    k_build(const palz_triangle* __restrict__ tris, const uint64_t* __restrict__ words,
            const BuildItem* __restrict__ items, uint32_t m, uint32_t n, const uint8_t* __restrict__ deltas,
            palz_internal_node* __restrict__ nodes, palz_node_index* __restrict__ root,
            BuildItem* __restrict__ staging, uint32_t* __restrict__ chunk_count,
            uint32_t* __restrict__ chunk_bits) {
    __shared__ uint8_t s_delta[kBuildChunk];  // delta of the split after each of the chunk's items
    __shared__ uint8_t s_pre[kBuildChunk];    // max of s_delta up to t
    __shared__ uint8_t s_suf[kBuildChunk];    // max of s_delta from t on
    __shared__ uint8_t s_local[kBuildChunk];
    __shared__ uint32_t s_flag[kBuildChunk];
    __shared__ uint32_t s_up[kBuildChunk / 32];             // chunk items that go up a level
    __shared__ uint32_t s_wave[2][kBuildChunk / 64];        // per-wave prefix / suffix maxima
    __shared__ palz_internal_node s_node[kBuildChunk];      // records of the local nodes
    const uint32_t n_int = n - 1u;
    const uint32_t m_int = m - 1u;  // item-level splits
    const uint32_t c0 = blockIdx.x * kBuildChunk;
    const uint32_t c_last = min(c0 + (uint32_t)kBuildChunk, m) - 1u;  // last item of the chunk
    const uint32_t cnt = c_last - c0;                                   // item splits c0 .. c_last-1
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    auto item_hi = [&](uint32_t j) -> uint32_t { return kLeaves ? j : items[j].hi; };
    auto split_delta = [&](uint32_t j) -> uint32_t { return (uint32_t)deltas[item_hi(j)]; };
    const uint32_t d = t < cnt ? split_delta(c0 + t) : 0u;
    s_delta[t] = (uint8_t)d;
    s_flag[t] = kNoArrival;
    if (t < kBuildChunk / 32) s_up[t] = 0u;
    // prefix and suffix maxima: within the wave by shuffles, then across waves
    uint32_t pm = d, sm = d;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t x = (uint32_t)__shfl_up((int)pm, o, 64);
        const uint32_t y = (uint32_t)__shfl_down((int)sm, o, 64);
        if (lane >= (uint32_t)o) pm = max(pm, x);
        if (lane + o < 64u) sm = max(sm, y);
    }
    if (lane == 63u) s_wave[0][wave] = pm;
    if (lane == 0u) s_wave[1][wave] = sm;
    __syncthreads();
    for (uint32_t w = 0; w < wave; ++w) pm = max(pm, s_wave[0][w]);
    for (uint32_t w = wave + 1; w < kBuildChunk / 64; ++w) sm = max(sm, s_wave[1][w]);
    s_pre[t] = (uint8_t)pm;
    s_suf[t] = (uint8_t)sm;
    __syncthreads();
    if (t < cnt) {
        const uint32_t dp = s_delta[t];
        const bool left = (t > 0u && s_pre[t - 1u] >= dp) || c0 == 0u || split_delta(c0 - 1u) >= dp;
        const bool right = (t + 1u < cnt && s_suf[t + 1u] > dp) || c_last == m_int || split_delta(c_last) > dp;
        s_local[t] = (left && right) ? 1 : 0;
    }
    __syncthreads();

    const uint32_t it = c0 + t;
    if (it < m) {
        uint32_t a = it, bb = it;  // item range
        uint32_t lo, hi;           // leaf range
        palz_node_index node;
        v3 blo, bhi;
        if (kLeaves) {
            lo = hi = it;
            const uint32_t tri_index = (uint32_t)(words[it] >> 32);
            node = tri_index | PALZ_NODE_LEAF_BIT;
            const palz_triangle tr = tris[tri_index];
            blo = mk(tr.vs[0]);
            bhi = mk(tr.vs[0]);
            blo = cmin(blo, mk(tr.vs[1]));
            bhi = cmax(bhi, mk(tr.vs[1]));
            blo = cmin(blo, mk(tr.vs[2]));
            bhi = cmax(bhi, mk(tr.vs[2]));
        } else {
            const BuildItem in = items[it];
            lo = in.lo;
            hi = in.hi;
            node = in.node;
            blo = mk(in.box[0], in.box[1], in.box[2]);
            bhi = mk(in.box[3], in.box[4], in.box[5]);
        }
        auto delta_at = [&](uint32_t j) -> uint32_t {  // item split j
            return j - c0 < cnt ? s_delta[j - c0] : split_delta(j);
        };
        bool go_up = false;
        while (bb - a < m_int) {
            const uint32_t dl = a > 0u ? delta_at(a - 1u) : 0xFFFFFFFFu;
            const uint32_t dr = bb < m_int ? delta_at(bb) : 0xFFFFFFFFu;
            const int go_left = dl < dr ? 1 : 0;
            const uint32_t parent = go_left ? a - 1u : bb;  // item split
            if (!(parent - c0 < cnt && s_local[parent - c0])) {
                go_up = true;
                break;
            }
            palz_internal_node& rec = s_node[parent - c0];
            rec.children[go_left] = node;
            rec.aabbs[go_left].lower = {blo.x, blo.y, blo.z};
            rec.aabbs[go_left].upper = {bhi.x, bhi.y, bhi.z};
            const uint32_t mine = go_left ? bb : a;
            const uint32_t other = __hip_atomic_exchange(&s_flag[parent - c0], mine, __ATOMIC_ACQ_REL,
                                                         __HIP_MEMORY_SCOPE_WORKGROUP);
            if (other == kNoArrival) break;
            const palz_aabb sb = rec.aabbs[go_left ^ 1];
            if (go_left) {
                a = other;
                lo = kLeaves ? other : items[other].lo;
            } else {
                bb = other;
                hi = kLeaves ? other : items[other].hi;
            }
            uint32_t grand = PALZ_NODE_INVALID;
            if (hi - lo < n_int) {
                const uint32_t gl = lo > 0u ? (uint32_t)deltas[lo - 1u] : 0xFFFFFFFFu;
                const uint32_t gr = hi < n_int ? (uint32_t)deltas[hi] : 0xFFFFFFFFu;
                grand = gl < gr ? lo - 1u : hi;
            }
            rec.parent = grand;
            rec.context = hi < n_int ? hi : 0xFFFFFFFFu;
            node = kLeaves ? parent : item_hi(parent);  // the split's leaf-level index
            blo = cmin(blo, mk(sb.lower));
            bhi = cmax(bhi, mk(sb.upper));
        }
        if (bb - a == m_int) {
            *root = node;
        } else if (go_up) {
            // this subtree root is an item of the next level, staged at its first item
            BuildItem up;
            up.node = node;
            up.lo = lo;
            up.hi = hi;
            up.box[0] = blo.x;
            up.box[1] = blo.y;
            up.box[2] = blo.z;
            up.box[3] = bhi.x;
            up.box[4] = bhi.y;
            up.box[5] = bhi.z;
            up.pad_ = 0u;
            staging[a] = up;
            atomicOr(&s_up[(a - c0) >> 5], 1u << ((a - c0) & 31u));
        }
    }
    __syncthreads();
    // the chunk's local records, written out whole
    for (uint32_t i = t; i < cnt; i += blockDim.x) {
        if (!s_local[i]) continue;
        const float4* src = reinterpret_cast<const float4*>(&s_node[i]);
        float4* dst = reinterpret_cast<float4*>(nodes + item_hi(c0 + i));
#pragma unroll
        for (int q = 0; q < 4; ++q) dst[q] = src[q];
    }
    // which of the chunk's positions hold an item of the next level
    if (staging) {
        if (t < kBuildChunk / 32) chunk_bits[blockIdx.x * (kBuildChunk / 32) + t] = s_up[t];
        if (t == 0) {
            uint32_t total = 0;
            for (uint32_t w = 0; w < kBuildChunk / 32; ++w) total += __popc(s_up[w]);
            chunk_count[blockIdx.x] = total;
        }
    }
}

// Compacts the staged items of each chunk (staging[c * kBuildChunk + j] for
// the set bits j of its bitmap) to out[offset[c] + rank]: the next level's
// items in leaf order.  One workgroup of kBuildChunk threads per chunk.
__global__ void __launch_bounds__(kBuildChunk) k_build_gather(const BuildItem* __restrict__ staging,
                                                              const uint32_t* __restrict__ bits,
                                                              const uint32_t* __restrict__ offset,
                                                              uint32_t n_chunks, BuildItem* __restrict__ out) {
    __shared__ uint32_t s_bits[kBuildChunk / 32];
    for (uint32_t c = blockIdx.x; c < n_chunks; c += gridDim.x) {
        __syncthreads();
        if (threadIdx.x < kBuildChunk / 32) s_bits[threadIdx.x] = bits[c * (kBuildChunk / 32) + threadIdx.x];
        __syncthreads();
        const uint32_t t = threadIdx.x, word = t >> 5, bit = 1u << (t & 31u);
        if (s_bits[word] & bit) {
            uint32_t rank = __popc(s_bits[word] & (bit - 1u));
            for (uint32_t w = 0; w < word; ++w) rank += __popc(s_bits[w]);
            out[offset[c] + rank] = staging[(size_t)c * kBuildChunk + t];
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
    uint32_t* arrival = nullptr;
    palz::BuildItem* staging = nullptr;  // per chunk, the items that go up a level
    palz::BuildItem* items[2] = {};      // the items of a level (ping-pong)
    uint32_t items_cap = 0;
    uint32_t* chunk_count = nullptr;     // [chunks] counts, [chunks] offsets, scan sums, total
    uint32_t* chunk_bits = nullptr;      // per chunk, the positions staged for the next level
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
        PALZ_HIP(hipMalloc(&arrival, sizeof(uint32_t) * std::max(4u, want)));
        const uint32_t chunks = palz::div_up(want, (uint32_t)palz::kBuildChunk);
        PALZ_HIP(hipMalloc(&staging, sizeof(palz::BuildItem) * (size_t)chunks * palz::kBuildChunk));
        PALZ_HIP(hipMalloc(&chunk_count, sizeof(uint32_t) * (2u * chunks + palz::scan_scratch_entries(chunks) + 4u)));
        PALZ_HIP(hipMalloc(&chunk_bits, sizeof(uint32_t) * chunks * (palz::kBuildChunk / 32)));
        capacity = want;
        return PALZ_OK;
    }
    void release_arrays() {
        if (nodes) (void)hipFree(nodes);
        if (words) (void)hipFree(words);
        if (tmp) (void)hipFree(tmp);
        if (deltas) (void)hipFree(deltas);
        if (arrival) (void)hipFree(arrival);
        if (staging) (void)hipFree(staging);
        if (chunk_count) (void)hipFree(chunk_count);
        if (chunk_bits) (void)hipFree(chunk_bits);
        chunk_bits = nullptr;
        for (auto& p : items)
            if (p) (void)hipFree(p);
        staging = nullptr;
        chunk_count = nullptr;
        items[0] = items[1] = nullptr;
        items_cap = 0;
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
    }
};

extern "C" {

int palz_lbvh_create(palz_lbvh** out) {
    if (!out) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_lbvh_create: out is NULL");
    palz_lbvh* b = new palz_lbvh();
    hipError_t e1 = hipMalloc(&b->root, sizeof(palz_node_index));
    hipError_t e2 = hipMalloc(&b->scene, sizeof(palz_aabb));
    if (e1 != hipSuccess || e2 != hipSuccess) {
        delete b;
        return palz::hip_fail(e1 != hipSuccess ? e1 : e2, "hipMalloc", __FILE__, __LINE__);
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
        d_tris, un, aligned, b->scene, build_option, morton_bits, b->words, s.hist);
    if (ev) PALZ_HIP(hipEventRecord(ev[2], st));
    const int npass = launch_sort_passes(s, b->words, b->tmp, nullptr, nullptr, un, 0, 32, st,
                                         ev ? ev + 3 : nullptr);
    (void)npass;  // 4 passes: the sorted words are back in b->words
    if (un > 1u) {
        const uint32_t quads = div_up(un - 1u, 4u);
        k_deltas<<<div_up(quads, 256u), 256, 0, st>>>(b->words, un, b->deltas, b->arrival);
    }
    if (ev) PALZ_HIP(hipEventRecord(ev[7], st));
    // level by level, until one chunk holds every item
    {
        uint32_t m = un;
        int cur = -1;  // level 1 reads the leaves
        for (int level = 0;; ++level) {
            const uint32_t chunks = div_up(m, (uint32_t)kBuildChunk);
            BuildItem* stage = chunks > 1u ? b->staging : nullptr;
            if (cur < 0)
                k_build<true><<<chunks, kBuildChunk, 0, st>>>(d_tris, b->words, nullptr, m, un, b->deltas, b->nodes,
                                                              b->root, stage, b->chunk_count, b->chunk_bits);
            else
                k_build<false><<<chunks, kBuildChunk, 0, st>>>(nullptr, nullptr, b->items[cur], m, un, b->deltas,
                                                               b->nodes, b->root, stage, b->chunk_count,
                                                               b->chunk_bits);
            PALZ_LAUNCH_CHECK();
            if (chunks == 1u) break;
            if (level > 64) return fail(PALZ_ERR_HIP, "palz_lbvh_build: the level-wise build does not converge");
            uint32_t* counts = b->chunk_count;
            uint32_t* offsets = counts + chunks;
            uint32_t* sums = offsets + chunks;
            uint32_t* total = sums + scan_scratch_entries(chunks);
            launch_exclusive_scan(counts, offsets, chunks, nullptr, total, sums, st);
            uint32_t next = 0;
            PALZ_HIP(hipMemcpyAsync(&next, total, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
            PALZ_HIP(hipStreamSynchronize(st));
            if (next > b->items_cap) {
                for (auto& p : b->items)
                    if (p) PALZ_HIP(hipFree(p));
                const uint32_t cap = std::max(next, b->items_cap * 2u);
                PALZ_HIP(hipMalloc(&b->items[0], sizeof(BuildItem) * cap));
                PALZ_HIP(hipMalloc(&b->items[1], sizeof(BuildItem) * cap));
                b->items_cap = cap;
            }
            const int out = cur < 0 ? 0 : cur ^ 1;
            k_build_gather<<<std::min(chunks, 8192u), kBuildChunk, 0, st>>>(b->staging, b->chunk_bits, offsets,
                                                                             chunks, b->items[out]);
            cur = out;
            m = next;
        }
    }
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
