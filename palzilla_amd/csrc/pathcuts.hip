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
#include <algorithm>

#include "interval.hpp"
#include "prims.hpp"

namespace palz {

namespace {

constexpr int kCutsThreads = 128;
constexpr uint32_t kCutsStack = 192;                      // stack entries per thread
constexpr uint32_t kCutsMaxThreads = 2048u * kCutsThreads;  // resident walkers
constexpr uint32_t kCutsExpandTarget = 1u << 20;  // listed tuples: work for every walker
constexpr uint32_t kCutsListCap = 1u << 24;       // receiver points per call
constexpr int kCutsMaxLevels = 64;
constexpr uint32_t kCutsOnePassCap = 1u << 26;   // records: larger capacities take the two-pass walk

struct CutsArgs {
    const palz_triangle* tris;
    const palz_triangle_attrib* attribs;
    const palz_internal_node* nodes;
    const palz_node_index* root;
    const palz_normal_bound* nb;
    const palz_first_diffuse* points;
    uint32_t n_points;
    uint32_t event;
    uint32_t scene;
    uint32_t n_threads;
    float eta_lo, eta_hi;
    float lx, ly, lz;
    uint32_t* stack;
    uint32_t* counts;  // counting pass: tuples per point; writing pass: their offsets
    palz_path* out;
    uint32_t capacity;
    uint32_t* status;
    unsigned long long* total;  // 64-bit sum of the counts; all ones once it does not fit 32 bits
    uint64_t* words;            // one-pass mode: (tuple << 32 | record position) per record
};

// A pending tuple of the traversal: receiver point and one node per slot.
struct alignas(16) CutEntry {
    uint32_t q;
    uint32_t n[3];
};

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

__global__ void __launch_bounds__(256) k_seed(CutEntry* __restrict__ out, uint32_t n_points,
                                              const palz_node_index* __restrict__ root) {
    const uint32_t q = blockIdx.x * 256 + threadIdx.x;
    if (q >= n_points) return;
    const palz_node_index r = *root;
    out[q] = {q, {r, r, r}};
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
enum WalkMode { kCount = 0, kWrite = 1, kEmit = 2 };

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

__global__ void k_clear_total(unsigned long long* total) { *total = 0ull; }

// a total past 32 bits: report it and skip the writing pass
__global__ void k_check_total(unsigned long long* total, uint32_t* count, uint32_t* status) {
    if (*total > 0xFFFFFFFFull) {
        *count = 0xFFFFFFFFu;
        *total = ~0ull;
        if (status) atomicMax(status, (uint32_t)PALZ_ERR_CAPACITY);
    }
}

__global__ void k_write_zero(uint32_t* p) { *p = 0u; }

template <int K>
void launch_expand(const CutsArgs& a, const CutEntry* in, uint32_t n_in, uint32_t* counts, CutEntry* out,
                   uint32_t* expandable, hipStream_t st, bool write) {
    const uint32_t blocks = std::min(div_up(n_in, 256u), 8192u);
    if (write)
        k_expand<K, true><<<blocks, 256, 0, st>>>(a, in, n_in, counts, out, expandable);
    else
        k_expand<K, false><<<blocks, 256, 0, st>>>(a, in, n_in, counts, out, expandable);
}

template <int K>
void launch_walk(const CutsArgs& a, const CutEntry* entries, uint32_t n_entries, uint32_t* queue, hipStream_t st,
                 int mode) {
    const uint32_t blocks = div_up(a.n_threads, (uint32_t)kCutsThreads);
    k_write_zero<<<1, 1, 0, st>>>(queue);
    if (mode == kCount) k_walk<K, kCount><<<blocks, kCutsThreads, 0, st>>>(a, entries, n_entries, queue);
    if (mode == kWrite) k_walk<K, kWrite><<<blocks, kCutsThreads, 0, st>>>(a, entries, n_entries, queue);
    if (mode == kEmit) k_walk<K, kEmit><<<blocks, kCutsThreads, 0, st>>>(a, entries, n_entries, queue);
}

// records in sorted-word order: out[i] = scratch[position of word i]
__global__ void __launch_bounds__(256) k_gather_records(const uint64_t* __restrict__ words,
                                                        const palz_path* __restrict__ scratch, uint32_t n,
                                                        palz_path* __restrict__ out) {
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
        out[i] = scratch[(uint32_t)words[i]];
}

__global__ void k_write_count(uint32_t* count, const unsigned long long* total) { *count = (uint32_t)*total; }

}  // namespace

}  // namespace palz

struct palz_path_cuts {
    uint32_t cap_entries = 0;
    palz::CutEntry* list[2] = {nullptr, nullptr};  // pending tuples, ping-pong
    uint32_t* counts = nullptr;
    uint32_t* scan_sums = nullptr;
    size_t cap_stack = 0;  // words
    uint32_t* stack = nullptr;
    unsigned long long* total = nullptr;
    uint32_t* scalars = nullptr;  // [0] expandable flag, [1] level total, [2] walk queue
    // one-pass mode: unsorted records, their words and the sort's tmp
    uint32_t cap_emit = 0;
    palz_path* emit_records = nullptr;
    uint64_t* emit_words = nullptr;
    uint64_t* emit_tmp = nullptr;
    palz_sorter* sorter = nullptr;

    int ensure_emit(uint32_t n) {
        if (!sorter && palz_sorter_create(&sorter) != PALZ_OK) return PALZ_ERR_OUT_OF_MEMORY;
        if (n <= cap_emit) return PALZ_OK;
        for (void* p : {(void*)emit_records, (void*)emit_words, (void*)emit_tmp})
            if (p) (void)hipFree(p);
        emit_records = nullptr;
        emit_words = emit_tmp = nullptr;
        cap_emit = 0;
        PALZ_HIP(hipMalloc(&emit_records, sizeof(palz_path) * n));
        PALZ_HIP(hipMalloc(&emit_words, sizeof(uint64_t) * n));
        PALZ_HIP(hipMalloc(&emit_tmp, sizeof(uint64_t) * n));
        cap_emit = n;
        return PALZ_OK;
    }

    int ensure_lists(uint32_t n_entries) {
        if (n_entries > cap_entries) {
            for (void* p : {(void*)list[0], (void*)list[1], (void*)counts, (void*)scan_sums})
                if (p) (void)hipFree(p);
            list[0] = list[1] = nullptr;
            counts = scan_sums = nullptr;
            cap_entries = 0;
            PALZ_HIP(hipMalloc(&list[0], sizeof(palz::CutEntry) * n_entries));
            PALZ_HIP(hipMalloc(&list[1], sizeof(palz::CutEntry) * n_entries));
            PALZ_HIP(hipMalloc(&counts, sizeof(uint32_t) * n_entries));
            PALZ_HIP(hipMalloc(&scan_sums, sizeof(uint32_t) * palz::scan_scratch_entries(n_entries)));
            cap_entries = n_entries;
        }
        if (!total) PALZ_HIP(hipMalloc(&total, sizeof(unsigned long long)));
        if (!scalars) PALZ_HIP(hipMalloc(&scalars, sizeof(uint32_t) * 4));
        return PALZ_OK;
    }
    int ensure_stack(size_t stack_words) {
        if (stack_words > cap_stack) {
            if (stack) (void)hipFree(stack);
            stack = nullptr;
            cap_stack = 0;
            PALZ_HIP(hipMalloc(&stack, sizeof(uint32_t) * stack_words));
            cap_stack = stack_words;
        }
        return PALZ_OK;
    }
    ~palz_path_cuts() {
        for (void* p : {(void*)list[0], (void*)list[1], (void*)counts, (void*)scan_sums, (void*)stack, (void*)total,
                        (void*)scalars, (void*)emit_records, (void*)emit_words, (void*)emit_tmp})
            if (p) (void)hipFree(p);
        if (sorter) palz_sorter_destroy(sorter);
    }
};

extern "C" {

int palz_path_cuts_create(palz_path_cuts** out) {
    if (!out) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_path_cuts_create: NULL");
    *out = new palz_path_cuts();
    return PALZ_OK;
}

int palz_path_cuts_destroy(palz_path_cuts* pc) {
    delete pc;
    return PALZ_OK;
}

int palz_path_cuts_enumerate(palz_path_cuts* pc, const palz_scene_desc* scene, const palz_normal_bound* d_normal_bounds,
                             uint32_t event, float eta_lo, float eta_hi, uint32_t scene_index, palz_path* d_paths,
                             uint32_t capacity, uint32_t* d_count, uint32_t* d_status, void* stream) {
    using namespace palz;
    if (!pc || !scene || !d_count) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_path_cuts_enumerate: NULL argument");
    const int K = ev_size(event);
    if (K < 1 || K > 3) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_path_cuts_enumerate: chain length must be 1..3");
    if (capacity && !d_paths) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_path_cuts_enumerate: NULL d_paths");
    if (!(eta_lo <= eta_hi)) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_path_cuts_enumerate: eta_lo > eta_hi");
    hipStream_t st = (hipStream_t)stream;
    if (scene->n_points == 0) {
        k_write_zero<<<1, 1, 0, st>>>(d_count);
        PALZ_LAUNCH_CHECK();
        return PALZ_OK;
    }
    if (!scene->tris || !scene->attribs || !scene->nodes || !scene->root || !scene->points || !d_normal_bounds)
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_path_cuts_enumerate: incomplete scene");
    if (scene->n_points > kCutsListCap)
        return fail(PALZ_ERR_CAPACITY, "palz_path_cuts_enumerate: more than 2^24 receiver points");
    const uint32_t cap = std::max(scene->n_points, kCutsListCap / 8);
    int rc = pc->ensure_lists(cap);
    if (rc != PALZ_OK) return rc;
    CutsArgs a;
    a.tris = scene->tris;
    a.attribs = scene->attribs;
    a.nodes = scene->nodes;
    a.root = scene->root;
    a.nb = d_normal_bounds;
    a.points = scene->points;
    a.n_points = scene->n_points;
    a.event = event;
    a.scene = scene_index;
    a.eta_lo = eta_lo;
    a.eta_hi = eta_hi;
    a.lx = scene->p_light.x;
    a.ly = scene->p_light.y;
    a.lz = scene->p_light.z;
    a.stack = nullptr;  // set once the walkers are counted
    a.counts = pc->counts;
    a.out = d_paths;
    a.capacity = capacity;
    a.status = d_status;
    a.total = pc->total;
    a.words = nullptr;

    // Breadth-first expansion until the list holds enough tuples to fill the
    // device: one host read of the level's total per level.
    int cur = 0;
    uint32_t n = scene->n_points;
    k_seed<<<div_up(n, 256u), 256, 0, st>>>(pc->list[0], n, scene->root);
    PALZ_LAUNCH_CHECK();
    for (int level = 0; level < kCutsMaxLevels && n > 0 && n < kCutsExpandTarget; ++level) {
        k_write_zero<<<1, 1, 0, st>>>(pc->scalars);
        if (K == 1) launch_expand<1>(a, pc->list[cur], n, pc->counts, nullptr, pc->scalars, st, false);
        if (K == 2) launch_expand<2>(a, pc->list[cur], n, pc->counts, nullptr, pc->scalars, st, false);
        if (K == 3) launch_expand<3>(a, pc->list[cur], n, pc->counts, nullptr, pc->scalars, st, false);
        PALZ_LAUNCH_CHECK();
        launch_exclusive_scan(pc->counts, pc->counts, n, nullptr, pc->scalars + 1, pc->scan_sums, st);
        PALZ_LAUNCH_CHECK();
        uint32_t h[2];
        PALZ_HIP(hipMemcpyAsync(h, pc->scalars, sizeof(h), hipMemcpyDeviceToHost, st));
        PALZ_HIP(hipStreamSynchronize(st));
        if (!h[0] || h[1] > cap) break;  // nothing left to cut, or the list would not fit
        if (K == 1) launch_expand<1>(a, pc->list[cur], n, pc->counts, pc->list[cur ^ 1], pc->scalars, st, true);
        if (K == 2) launch_expand<2>(a, pc->list[cur], n, pc->counts, pc->list[cur ^ 1], pc->scalars, st, true);
        if (K == 3) launch_expand<3>(a, pc->list[cur], n, pc->counts, pc->list[cur ^ 1], pc->scalars, st, true);
        PALZ_LAUNCH_CHECK();
        cur ^= 1;
        n = h[1];
    }
    if (n == 0) {
        k_write_zero<<<1, 1, 0, st>>>(d_count);
        PALZ_LAUNCH_CHECK();
        return PALZ_OK;
    }

    // walk every listed tuple's subtree: count, scan, write
    a.n_threads = std::min(div_up(n, (uint32_t)kCutsThreads) * kCutsThreads, kCutsMaxThreads);
    rc = pc->ensure_stack((size_t)kCutsStack * K * a.n_threads);
    if (rc != PALZ_OK) return rc;
    a.stack = pc->stack;
    auto walk = [&](int mode) {
        if (K == 1) launch_walk<1>(a, pc->list[cur], n, pc->scalars + 2, st, mode);
        if (K == 2) launch_walk<2>(a, pc->list[cur], n, pc->scalars + 2, st, mode);
        if (K == 3) launch_walk<3>(a, pc->list[cur], n, pc->scalars + 2, st, mode);
    };
    if (capacity && capacity <= kCutsOnePassCap) {
        // one walk into scratch, a stable sort of its words by listed tuple and a
        // gather; a total past `capacity` falls back to the two-pass walk below
        rc = pc->ensure_emit(capacity);
        if (rc != PALZ_OK) return rc;
        a.out = pc->emit_records;
        a.words = pc->emit_words;
        k_clear_total<<<1, 1, 0, st>>>(pc->total);
        walk(kEmit);
        PALZ_LAUNCH_CHECK();
        unsigned long long total = 0;
        PALZ_HIP(hipMemcpyAsync(&total, pc->total, sizeof(total), hipMemcpyDeviceToHost, st));
        PALZ_HIP(hipStreamSynchronize(st));
        a.out = d_paths;
        a.words = nullptr;
        if (total <= capacity) {
            const uint32_t m = (uint32_t)total;
            int bits = 1;
            while (bits < 32 && (n - 1u) >> bits) ++bits;
            if (m > 1) {
                rc = palz_sort_u64(pc->sorter, pc->emit_words, pc->emit_tmp, nullptr, nullptr, m, 32, 32 + bits, st);
                if (rc != PALZ_OK) return rc;
            }
            if (m) {
                k_gather_records<<<std::min(div_up(m, 256u), 65536u), 256, 0, st>>>(pc->emit_words, pc->emit_records,
                                                                                 m, d_paths);
                PALZ_LAUNCH_CHECK();
            }
            k_write_count<<<1, 1, 0, st>>>(d_count, pc->total);
            PALZ_LAUNCH_CHECK();
            return PALZ_OK;
        }
    }
    k_clear_total<<<1, 1, 0, st>>>(pc->total);
    walk(kCount);
    PALZ_LAUNCH_CHECK();
    launch_exclusive_scan(pc->counts, pc->counts, n, nullptr, d_count, pc->scan_sums, st);
    k_check_total<<<1, 1, 0, st>>>(pc->total, d_count, d_status);
    PALZ_LAUNCH_CHECK();
    if (capacity) {
        walk(kWrite);
        PALZ_LAUNCH_CHECK();
    }
    return PALZ_OK;
}

}  // extern "C"
