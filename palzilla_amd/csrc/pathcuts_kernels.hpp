// Node-level path cuts: traverseAdmissibleNodes (pk.h:1760-2212) as a GPU
// enumeration of the admissible K-triangle tuples between the light and each
// receiver point.
//
// The traversal is the reference's: the query starts with every slot at the
// root; each step cuts the slot whose internal node has the largest surface
// area (first slot on ties, pk.h:1806-1826), tests both children with the
// other slots' boxes (a cut slot takes the child's box and normal bound, an
// uncut internal slot its node's box and both normal bounds, a leaf its
// triangle's vertex and shading-normal hulls, pk.h:2105-2140) and pushes the
// admissible ones, child 0 first, on a LIFO stack.  A tuple of leaves is
// emitted unless two consecutive slots name one triangle or the light or the
// receiver point lies behind the first or last triangle, winding-corrected by
// its first shading normal (pk.h:1830-1876).
//
// The per-tuple test is the conservative residual enclosure of interval.hpp
// in place of the reference's reflection / refraction_norm_free chain: it is
// inclusion-monotone, so a tuple of leaves is reached exactly when its own
// leaf boxes pass, and no tuple holding a solution the solver accepts is cut.
//
// Work is spread over the device in two stages.  A breadth-first expansion
// over all points replaces each pending tuple, in place in a list, by its
// admissible children in the order the walk would visit them, until the list
// holds enough tuples to fill the device (a tuple's traversal order is the
// concatenation of its children's, so the list keeps the reference's order).
// Then one thread per listed tuple walks its subtree on a LIFO stack in global
// memory (K node indices per entry, laid out slot-major across threads so that
// threads at equal depth touch adjacent words).  Every stage counts, scans and
// writes, so the output is deterministic and in traversal order.
//
// This header holds the device code; it is included only by the per-K
// translation units.
#pragma once

#include <algorithm>

#include "pathcuts.hpp"

namespace palz {
namespace cuts {

__device__ __forceinline__ iv::I3 aabb_box(const palz_aabb& b) { return iv::box(mk(b.lower), mk(b.upper)); }

__device__ __forceinline__ iv::I3 union_box(const palz_aabb& a, const palz_aabb& b) {
    return iv::box(cmin(mk(a.lower), mk(b.lower)), cmax(mk(a.upper), mk(b.upper)));
}

// AABB::surfaceArea (minimum_lbvh.h:256-260) of nodeAABB (pk.h:1744-1749)
__device__ __forceinline__ float surface_area(const palz_aabb& a, const palz_aabb& b) {
    const v3 s = cmax(mk(a.upper), mk(b.upper)) - cmin(mk(a.lower), mk(b.lower));
    return (s.x * s.y + s.y * s.z + s.z * s.x) * 2.0f;
}

__device__ __forceinline__ void leaf_boxes(const palz_triangle* tris, const palz_triangle_attrib* attribs,
                                           uint32_t t, iv::I3& vb, iv::I3& nb) {
    const palz_triangle tr = tris[t];
    const palz_triangle_attrib at = attribs[t];
    vb = iv::box(cmin(cmin(mk(tr.vs[0]), mk(tr.vs[1])), mk(tr.vs[2])),
                 cmax(cmax(mk(tr.vs[0]), mk(tr.vs[1])), mk(tr.vs[2])));
    nb = iv::box(cmin(cmin(mk(at.shading_normals[0]), mk(at.shading_normals[1])), mk(at.shading_normals[2])),
                 cmax(cmax(mk(at.shading_normals[0]), mk(at.shading_normals[1])), mk(at.shading_normals[2])));
}

// the closed-mesh side test of pk.h:1850-1875: p may not lie behind the
// triangle's geometric normal turned towards its first shading normal
__device__ __forceinline__ bool in_front(const palz_triangle* tris, const palz_triangle_attrib* attribs, uint32_t t,
                                         v3 p) {
    const palz_triangle tr = tris[t];
    const v3 v0 = mk(tr.vs[0]);
    v3 ng = cross(mk(tr.vs[1]) - v0, mk(tr.vs[2]) - v0);
    ng = ng * dot(ng, mk(attribs[t].shading_normals[0]));
    return !(dot(ng, p) < dot(ng, v0));
}

// every vertex of the chain light -> boxes -> point passes (k = vertex index,
// unrolled at compile time so that the boxes stay in registers)
template <int K, int k = 0>
__device__ __forceinline__ bool tuple_admissible(const iv::I3 (&vb)[K], const iv::I3 (&nb)[K], v3 pb, v3 pe,
                                                 uint32_t event, iv::I eta) {
    if constexpr (k == K) {
        return true;
    } else {
        iv::I3 prev, next;
        if constexpr (k == 0) prev = iv::pt3(pb); else prev = vb[k - 1];
        if constexpr (k == K - 1) next = iv::pt3(pe); else next = vb[k + 1];
        if (!iv::vertex_admissible(prev - vb[k], next - vb[k], nb[k], ev_is_T(event, k), eta)) return false;
        return tuple_admissible<K, k + 1>(vb, nb, pb, pe, event, eta);
    }
}

// The traversal step at tuple `cur`: the slot to cut (-1 when every slot is a
// leaf), the children of its node and whether each passes (pk.h:1800-2212).
template <int K>
__device__ __forceinline__ int cut_step(const CutsArgs& a, const uint32_t (&cur)[K], v3 pb, v3 pe, iv::I eta,
                                        bool (&ok)[2], uint32_t (&child)[2]) {
    // the slot to cut: largest node surface area, first on ties
    int cut = -1;
    float best = -1.0f;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (cur[k] & PALZ_NODE_LEAF_BIT) continue;
        const palz_internal_node& nd = a.nodes[cur[k]];
        const float sa = surface_area(nd.aabbs[0], nd.aabbs[1]);
        if (best < sa) {
            best = sa;
            cut = k;
        }
    }
    if (cut < 0) return -1;
    iv::I3 vb[K], nb[K];
    uint32_t cut_node = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (k == cut) {
            cut_node = cur[k];
        } else if (cur[k] & PALZ_NODE_LEAF_BIT) {
            leaf_boxes(a.tris, a.attribs, cur[k] & ~PALZ_NODE_LEAF_BIT, vb[k], nb[k]);
        } else {
            const palz_internal_node& nd = a.nodes[cur[k]];
            const palz_normal_bound& nn = a.nb[cur[k]];
            vb[k] = union_box(nd.aabbs[0], nd.aabbs[1]);
            nb[k] = union_box(nn.normal_bounds[0], nn.normal_bounds[1]);
        }
    }
    const palz_internal_node& nd = a.nodes[cut_node];
    const palz_normal_bound& nn = a.nb[cut_node];
#pragma unroll 1
    for (int i = 0; i < 2; ++i) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (k == cut) {
                vb[k] = aabb_box(nd.aabbs[i]);
                nb[k] = aabb_box(nn.normal_bounds[i]);
            }
        }
        const bool pass = tuple_admissible<K>(vb, nb, pb, pe, a.event, eta);
        if (i == 0) ok[0] = pass; else ok[1] = pass;
    }
    child[0] = nd.children[0];
    child[1] = nd.children[1];
    return cut;
}

// A tuple of leaves is emitted unless two consecutive slots name one
// triangle or an end point lies behind its triangle (pk.h:1830-1876).
template <int K>
__device__ __forceinline__ bool leaf_tuple_valid(const CutsArgs& a, const uint32_t (&cur)[K], v3 pb, v3 pe) {
#pragma unroll
    for (int k = 0; k + 1 < K; ++k)
        if (cur[k] == cur[k + 1]) return false;
    return in_front(a.tris, a.attribs, cur[0] & ~PALZ_NODE_LEAF_BIT, pb) &&
           in_front(a.tris, a.attribs, cur[K - 1] & ~PALZ_NODE_LEAF_BIT, pe);
}

// receiver points without a diffuse hit (main_gpu.cu:385)
__device__ __forceinline__ bool has_hit(const palz_first_diffuse& pt) {
    return !(pt.R.x == 0.0f && pt.R.y == 0.0f && pt.R.z == 0.0f);
}

// One level of the breadth-first expansion: each pending tuple is replaced,
// in place in the list, by its admissible children in the order the walk
// visits them (child 1, then child 0: child 0 is pushed first); a tuple of
// leaves stays as it is.  Counting pass, scan, writing pass.
template <int K, bool kWrite>
__global__ void __launch_bounds__(256) k_expand(CutsArgs a, const CutEntry* __restrict__ in, uint32_t n_in,
                                                uint32_t* __restrict__ counts, CutEntry* __restrict__ out,
                                                uint32_t* __restrict__ expandable) {
    const iv::I eta = {a.eta_lo, a.eta_hi};
    const v3 pb = mk(a.lx, a.ly, a.lz);
    for (uint32_t e = blockIdx.x * 256 + threadIdx.x; e < n_in; e += gridDim.x * 256) {
        const CutEntry en = in[e];
        const palz_first_diffuse pt = a.points[en.q];
        uint32_t cnt = 0;
        if (has_hit(pt)) {
            uint32_t cur[K];
#pragma unroll
            for (int k = 0; k < K; ++k) cur[k] = en.n[k];
            bool ok[2];
            uint32_t child[2];
            const int cut = cut_step<K>(a, cur, pb, mk(pt.p), eta, ok, child);
            if (cut < 0) {
                cnt = 1;
                if (kWrite) out[counts[e]] = en;
            } else {
                cnt = (uint32_t)ok[0] + (uint32_t)ok[1];
                if (!kWrite) {
                    *expandable = 1u;
                } else {
                    uint32_t o = counts[e];
#pragma unroll
                    for (int i = 1; i >= 0; --i) {
                        if (!ok[i]) continue;
                        CutEntry c = en;
#pragma unroll
                        for (int k = 0; k < K; ++k)
                            if (k == cut) c.n[k] = child[i];
                        out[o++] = c;
                    }
                }
            }
        }
        if (!kWrite) counts[e] = cnt;
    }
}

// The walk of every listed tuple's subtree on a LIFO stack in global memory,
// persistent: a lane that finishes a tuple takes the next one from a queue
// (one atomic per wave), so lanes of a wave do not wait for the longest walk.
// Modes: kCount (per-tuple counts and their 64-bit sum), kWrite (after the
// scan: records at the tuples' offsets), kEmit (one pass: records in arrival
// order at positions taken from a 64-bit counter, one atomic per wave, each
// with the word tuple << 32 | position; a stable sort of the words by tuple
// then restores the traversal order, since one lane emits a tuple's records
// at increasing positions).

template <int K, int kMode>
__global__ void __launch_bounds__(kCutsThreads) k_walk(CutsArgs a, const CutEntry* __restrict__ entries,
                                                       uint32_t n_entries, uint32_t* __restrict__ queue) {
    // n_threads is a multiple of the block size and every lane stays in the
    // loop until the wave is done: the ballots and the final reduction see
    // whole waves
    const uint32_t tid = blockIdx.x * kCutsThreads + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t* const stack = a.stack + tid;
    const uint32_t stride = a.n_threads;  // words between consecutive entries of one thread
    const iv::I eta = {a.eta_lo, a.eta_hi};
    const v3 pb = mk(a.lx, a.ly, a.lz);
    if (kMode == kWrite && *a.total == ~0ull) return;
    unsigned long long sum = 0;
    constexpr uint32_t kNone = 0xFFFFFFFFu;
    uint32_t e = kNone;
    bool done = false;
    uint32_t cur[K];
#pragma unroll
    for (int k = 0; k < K; ++k) cur[k] = 0;
    uint32_t top = 0, found = 0, base = 0, q = 0;
    v3 pe = {0.0f, 0.0f, 0.0f};
    while (true) {
        // lanes without a tuple take the next ones
        const bool need = !done && e == kNone;
        const uint64_t m = __ballot(need);
        if (m) {
            const int leader = __ffsll((unsigned long long)m) - 1;
            uint32_t first = 0;
            if ((int)lane == leader) first = atomicAdd(queue, (uint32_t)__popcll(m));
            first = __shfl(first, leader);
            if (need) {
                const uint32_t idx = first + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
                if (idx < n_entries) {
                    const CutEntry en = entries[idx];
                    const palz_first_diffuse pt = a.points[en.q];
                    e = idx;
                    q = en.q;
                    pe = mk(pt.p);
#pragma unroll
                    for (int k = 0; k < K; ++k) cur[k] = en.n[k];
                    top = 0;
                    found = 0;
                    base = kMode == kWrite ? a.counts[idx] : 0u;
                    if (!has_hit(pt)) {
                        if (kMode == kCount) a.counts[idx] = 0u;
                        e = kNone;
                    }
                } else {
                    done = true;
                }
            }
        }
        if (__ballot(!done) == 0) break;
        if (e == kNone) continue;
        // one step of this lane's walk
        bool ok[2];
        uint32_t child[2];
        bool finished = false;
        const int cut = cut_step<K>(a, cur, pb, pe, eta, ok, child);
        const bool emit = cut < 0 && leaf_tuple_valid<K>(a, cur, pb, pe);
        uint64_t pos = base + found;
        if (kMode == kEmit) {
            // positions for this step's records, one atomic per wave
            const uint64_t em = __ballot(emit);
            if (em) {
                const int leader = __ffsll((unsigned long long)em) - 1;
                unsigned long long first = 0;
                if ((int)lane == leader) first = atomicAdd(a.total, (unsigned long long)__popcll(em));
                first = __shfl(first, leader);
                pos = first + (uint64_t)__popcll(em & ((1ull << lane) - 1ull));
            }
        }
        if (cut < 0) {
            if (emit) {
                if (kMode != kCount && pos < a.capacity) {
                    palz_path p;
                    p.scene = a.scene;
                    p.pixel = q;
                    p.event = a.event;
                    p.photon = 0xFFFFFFFFu;
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        p.tris[k] = k < K ? (int32_t)(cur[k < K ? k : 0] & ~PALZ_NODE_LEAF_BIT) : -1;
                        // the solver's commented-out start point (pk.h:470-473)
                        p.params[2 * k] = k < K ? 1.0f / 3.0f : 0.0f;
                        p.params[2 * k + 1] = k < K ? 1.0f / 3.0f : 0.0f;
                    }
                    a.out[pos] = p;
                    if (kMode == kEmit) a.words[pos] = (uint64_t)e << 32 | pos;
                }
                ++found;
            }
        } else if (top + (uint32_t)ok[0] + (uint32_t)ok[1] > kCutsStack) {
            // deeper than the stack: this tuple stops here and the call reports it
            if (a.status) atomicMax(a.status, (uint32_t)PALZ_ERR_TRAVERSAL_STACK);
            finished = true;
        } else {
            // child 0 first: child 1 is popped next (pk.h:2201-2208)
            auto push = [&](uint32_t c) {
#pragma unroll
                for (int k = 0; k < K; ++k) stack[(top * K + k) * stride] = k == cut ? c : cur[k];
                ++top;
            };
            if (ok[0]) push(child[0]);
            if (ok[1]) push(child[1]);
        }
        if (!finished && top > 0) {
            --top;
#pragma unroll
            for (int k = 0; k < K; ++k) cur[k] = stack[(top * K + k) * stride];
        } else {
            if (kMode == kCount) a.counts[e] = found;
            sum += found;
            e = kNone;
        }
    }
    if (kMode == kCount) {
        // one 64-bit atomic per wave
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
        if (lane == 0) atomicAdd(a.total, sum);
    }
}

template <int K>
void launch_expand(const CutsArgs& a, const CutEntry* in, uint32_t n_in, uint32_t* counts, CutEntry* out,
                   uint32_t* expandable, hipStream_t st, bool write) {
    const uint32_t blocks = std::min(div_up(n_in, 256u), 8192u);
    if (write)
        k_expand<K, true><<<blocks, 256, 0, st>>>(a, in, n_in, counts, out, expandable);
    else
        k_expand<K, false><<<blocks, 256, 0, st>>>(a, in, n_in, counts, out, expandable);
}

template <int K, int kMode>
void launch_walk(const CutsArgs& a, const CutEntry* entries, uint32_t n_entries, uint32_t* queue, hipStream_t st) {
    const uint32_t blocks = div_up(a.n_threads, (uint32_t)kCutsThreads);
    k_walk<K, kMode><<<blocks, kCutsThreads, 0, st>>>(a, entries, n_entries, queue);
}

}  // namespace cuts
}  // namespace palz
