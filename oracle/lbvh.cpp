// Oracle: scene AABB, Morton keys, stable sort, deltas, Apetrei LBVH build,
// an independent top-down LBVH build and validate_lbvh.  TEST INFRASTRUCTURE.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>
#include <vector>

#include "oracle.hpp"

using namespace orc;

namespace {

// splat3 / encodeMortonCode (minimum_lbvh.h:173-195)
uint64_t splat3(uint32_t a) {
    uint64_t x = a & 0x1fffff;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}
uint64_t encode_morton(uint32_t x, uint32_t y, uint32_t z) {
    return splat3(x) | splat3(y) << 1 | splat3(z) << 2;
}

// int3 coord = clamp(int3((p - lower) / (upper - lower) * scale), 0, maxv)
// (minimum_lbvh.h:249-254).  The truncating conversion is spelled out so that
// NaN gives 0 and +inf gives maxv on every platform.
uint32_t grid_coord(float p, float lo, float hi, float scale, uint32_t maxv) {
    const float f = (p - lo) / (hi - lo) * scale;
    if (std::isnan(f) || f <= 0.0f) return 0u;
    if (f >= scale) return maxv;
    const uint32_t c = (uint32_t)f;
    return c > maxv ? maxv : c;
}

uint32_t key_of(float3 c, const float* box, int bits) {
    if (bits == 30) {
        const uint32_t x = grid_coord(c.x, box[0], box[3], 1024.0f, 1023u);
        const uint32_t y = grid_coord(c.y, box[1], box[4], 1024.0f, 1023u);
        const uint32_t z = grid_coord(c.z, box[2], box[5], 1024.0f, 1023u);
        return (uint32_t)encode_morton(x, y, z);
    }
    const uint32_t x = grid_coord(c.x, box[0], box[3], 2097152.0f, 0x1FFFFFu);
    const uint32_t y = grid_coord(c.y, box[1], box[4], 2097152.0f, 0x1FFFFFu);
    const uint32_t z = grid_coord(c.z, box[2], box[5], 2097152.0f, 0x1FFFFFu);
    return (uint32_t)(encode_morton(x, y, z) >> 31);  // minimum_lbvh.h:565
}

uint32_t delta_of(uint32_t a, uint32_t b) {
    const uint32_t x = a ^ b;
    return x == 0u ? 0u : 32u - (uint32_t)__builtin_clz(x);  // minimum_lbvh.h:140-143
}

palz_aabb tri_box(const palz_triangle& t) {
    // setEmpty() then extend() by each vertex (minimum_lbvh.h:309-316); the
    // empty box folds away against the first vertex.
    palz_aabb b{t.vs[0], t.vs[0]};
    for (int j = 1; j < 3; ++j) {
        b.lower = {box_min(b.lower.x, t.vs[j].x), box_min(b.lower.y, t.vs[j].y), box_min(b.lower.z, t.vs[j].z)};
        b.upper = {box_max(b.upper.x, t.vs[j].x), box_max(b.upper.y, t.vs[j].y), box_max(b.upper.z, t.vs[j].z)};
    }
    return b;
}
palz_aabb box_union(const palz_aabb& a, const palz_aabb& b) {
    palz_aabb r;
    r.lower = {box_min(a.lower.x, b.lower.x), box_min(a.lower.y, b.lower.y), box_min(a.lower.z, b.lower.z)};
    r.upper = {box_max(a.upper.x, b.upper.x), box_max(a.upper.y, b.upper.y), box_max(a.upper.z, b.upper.z)};
    return r;
}

constexpr uint32_t kLeaf = PALZ_NODE_LEAF_BIT;
constexpr uint32_t kInvalid = PALZ_NODE_INVALID;

// Top-down restatement: the node over sorted leaves [lo, hi] splits at the
// first position s in [lo, hi-1] holding the largest delta; internal node s
// owns that split.  This is the unique tree the bottom-up climb produces
// (goLeft = deltaL < deltaR, minimum_lbvh.h:324), derived independently.
uint32_t topdown(const std::vector<uint32_t>& idx, const uint8_t* d, const palz_triangle* tris,
                 uint32_t lo, uint32_t hi, uint32_t n_int, palz_internal_node* nodes, palz_aabb* box) {
    if (lo == hi) {
        *box = tri_box(tris[idx[lo]]);
        return idx[lo] | kLeaf;
    }
    uint32_t s = lo;
    for (uint32_t g = lo + 1; g < hi; ++g)
        if (d[g] > d[s]) s = g;
    palz_aabb lb, rb;
    const uint32_t l = topdown(idx, d, tris, lo, s, n_int, nodes, &lb);
    const uint32_t r = topdown(idx, d, tris, s + 1, hi, n_int, nodes, &rb);
    palz_internal_node& nd = nodes[s];
    nd.children[0] = l;
    nd.children[1] = r;
    nd.aabbs[0] = lb;
    nd.aabbs[1] = rb;
    nd.context = hi < n_int ? hi : 0xFFFFFFFFu;
    if (!(l & kLeaf)) nodes[l].parent = s;
    if (!(r & kLeaf)) nodes[r].parent = s;
    *box = box_union(lb, rb);
    return s;
}

bool validate(uint32_t node, const palz_internal_node* nodes, const uint8_t* d, uint32_t max_delta) {
    if (node & kLeaf) return true;  // validate_lbvh (minimum_lbvh.h:424-435)
    const uint32_t dv = d[node];
    if (dv > max_delta) return false;
    return validate(nodes[node].children[0], nodes, d, dv) && validate(nodes[node].children[1], nodes, d, dv);
}

}  // namespace

extern "C" {

// getSceneAABB / BVHCPUBuilder scene box (minimum_lbvh.h:531-541).
void orc_scene_aabb(const palz_triangle* tris, int n, palz_aabb* out) {
    palz_aabb b;
    b.lower = {3.402823466e+38F, 3.402823466e+38F, 3.402823466e+38F};
    b.upper = {-3.402823466e+38F, -3.402823466e+38F, -3.402823466e+38F};
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < 3; ++j) {
            const palz_float3& v = tris[i].vs[j];
            b.lower = {box_min(b.lower.x, v.x), box_min(b.lower.y, v.y), box_min(b.lower.z, v.z)};
            b.upper = {box_max(b.upper.x, v.x), box_max(b.upper.y, v.y), box_max(b.upper.z, v.z)};
        }
    *out = b;
}

// buildMortons (minimum_lbvh.cu:68-92) / BVHCPUBuilder keys (minimum_lbvh.h:543-568).
// USE_NORMAL uses the GPU packing (minimum_lbvh.cu:85).
void orc_morton_keys(const palz_triangle* tris, int n, const palz_aabb* scene, uint32_t option, int bits,
                     uint32_t* keys) {
    const float box[6] = {scene->lower.x, scene->lower.y, scene->lower.z,
                          scene->upper.x, scene->upper.y, scene->upper.z};
    const float unit[6] = {-1.f, -1.f, -1.f, 1.f, 1.f, 1.f};
    for (int i = 0; i < n; ++i) {
        const float3 v0 = load3(tris[i].vs[0]), v1 = load3(tris[i].vs[1]), v2 = load3(tris[i].vs[2]);
        const float3 c = (v0 + v1 + v2) / 3.0f;
        if (option & PALZ_BUILD_OPTION_USE_NORMAL) {
            const float3 ng = cross(v1 - v0, v2 - v0);
            const float3 nn = ng * (1.0f / std::sqrt(dot(ng, ng)));
            keys[i] = (key_of(nn, unit, 32) & 0xFFFFFFC0u) | (key_of(c, box, 32) >> 26);
        } else {
            keys[i] = key_of(c, box, bits);
        }
    }
}

// encodeMortonCode_Naive (minimum_lbvh.h:149-160), for cross-checking splat3.
uint64_t orc_morton_naive(uint32_t x, uint32_t y, uint32_t z) {
    uint64_t code = 0;
    for (uint64_t i = 0; i < 64 / 3; ++i)
        code |= ((uint64_t)(x & (1u << i)) << (2 * i + 0)) | ((uint64_t)(y & (1u << i)) << (2 * i + 1)) |
                ((uint64_t)(z & (1u << i)) << (2 * i + 2));
    return code;
}
uint64_t orc_morton_splat(uint32_t x, uint32_t y, uint32_t z) { return encode_morton(x, y, z); }

// Stable sort of 64-bit words by bits [start_bit, end_bit) (the semantics of
// OnesweepSort::sort, tinyhiponesweep.h:596-668); values move with the words.
void orc_stable_sort_u64(uint64_t* words, uint32_t* values, size_t n, int start_bit, int end_bit) {
    const int width = end_bit - start_bit;
    const uint64_t mask = width >= 64 ? ~0ull : ((1ull << width) - 1ull);
    std::vector<size_t> order(n);
    for (size_t i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        return ((words[a] >> start_bit) & mask) < ((words[b] >> start_bit) & mask);
    });
    std::vector<uint64_t> w(n);
    std::vector<uint32_t> v(values ? n : 0);
    for (size_t i = 0; i < n; ++i) {
        w[i] = words[order[i]];
        if (values) v[i] = values[order[i]];
    }
    std::memcpy(words, w.data(), n * sizeof(uint64_t));
    if (values) std::memcpy(values, v.data(), n * sizeof(uint32_t));
}

void orc_deltas(const uint64_t* sorted_words, int n, uint8_t* deltas) {
    for (int i = 0; i + 1 < n; ++i)
        deltas[i] = (uint8_t)delta_of((uint32_t)sorted_words[i], (uint32_t)sorted_words[i + 1]);
}

// build_lbvh (minimum_lbvh.h:293-370) run serially over the leaves, as
// BVHCPUBuilder::build does without CPU_PARALLEL (minimum_lbvh.h:592-606).
// Arrival bookkeeping lives in a separate array; the reference uses `context`
// (memset to 0xFFFFFFFF) for it and then overwrites it with the escape link.
void orc_build_lbvh(const palz_triangle* tris, int n, const uint64_t* sorted_words, const uint8_t* deltas,
                    palz_internal_node* nodes, uint32_t* root) {
    const uint32_t n_int = (uint32_t)n - 1u;
    std::vector<uint32_t> arrival(n_int, 0xFFFFFFFFu);
    for (uint32_t leaf = 0; leaf < (uint32_t)n; ++leaf) {
        uint32_t lo = leaf, hi = leaf;
        uint32_t node = (uint32_t)(sorted_words[leaf] >> 32) | kLeaf;
        palz_aabb box = tri_box(tris[node & ~kLeaf]);
        bool is_root = true;
        while (hi - lo < n_int) {
            const uint32_t dl = lo > 0 ? deltas[lo - 1] : 0xFFFFFFFFu;
            const uint32_t dr = hi < n_int ? deltas[hi] : 0xFFFFFFFFu;
            const int go_left = dl < dr ? 1 : 0;
            const uint32_t parent = go_left ? lo - 1 : hi;
            nodes[parent].children[go_left] = node;
            nodes[parent].aabbs[go_left] = box;
            if (!(node & kLeaf)) nodes[node].parent = parent;
            const uint32_t other = arrival[parent];
            arrival[parent] = go_left ? hi : lo;
            if (other == 0xFFFFFFFFu) {
                is_root = false;
                break;
            }
            lo = std::min(lo, other);
            hi = std::max(hi, other);
            nodes[parent].context = hi < n_int ? hi : 0xFFFFFFFFu;
            node = parent;
            box = box_union(box, nodes[parent].aabbs[go_left ^ 1]);
        }
        if (is_root) {
            *root = node;
            if (!(node & kLeaf)) nodes[node].parent = kInvalid;
        }
    }
}

// Whole BVHCPUBuilder::build pipeline (minimum_lbvh.h:521-607) with the
// (key, index) stable order of the GPU sort.
void orc_lbvh(const palz_triangle* tris, int n, uint32_t option, int bits, palz_internal_node* nodes,
              uint32_t* root, uint64_t* words_out, uint8_t* deltas_out, palz_aabb* scene_out) {
    palz_aabb scene;
    orc_scene_aabb(tris, n, &scene);
    std::vector<uint32_t> keys(n);
    orc_morton_keys(tris, n, &scene, option, bits, keys.data());
    std::vector<uint64_t> words(n);
    for (int i = 0; i < n; ++i) words[i] = (uint64_t)keys[i] | ((uint64_t)(uint32_t)i << 32);
    orc_stable_sort_u64(words.data(), nullptr, (size_t)n, 0, 32);
    std::vector<uint8_t> d(n > 1 ? n - 1 : 1);
    orc_deltas(words.data(), n, d.data());
    orc_build_lbvh(tris, n, words.data(), d.data(), nodes, root);
    if (words_out) std::memcpy(words_out, words.data(), sizeof(uint64_t) * n);
    if (deltas_out && n > 1) std::memcpy(deltas_out, d.data(), (size_t)n - 1);
    if (scene_out) *scene_out = scene;
}

// The same pipeline on `threads` host threads, the BVHCPUBuilder::build shape
// with CPU_PARALLEL (minimum_lbvh.h:592-606): scene box, keys and deltas in
// slices, a stable LSD radix sort of the (key, index) words with 8-bit digits
// (per-thread digit counts, then a stable scatter of each slice), and the
// bottom-up climb with one thread per leaf slice, where an atomic exchange on
// the parent's arrival word (acquire-release) decides which child completes
// it, as the reference's InterlockedExchange does (minimum_lbvh.h:340-344).
// The result equals orc_lbvh's bit for bit: the tree is fixed by the sorted
// keys, and boxes are exact min/max.
void orc_lbvh_parallel(const palz_triangle* tris, int n, uint32_t option, int bits, palz_internal_node* nodes,
                       uint32_t* root, int threads) {
    threads = std::max(1, threads);
    auto par = [&](auto&& body) {
        std::vector<std::thread> pool;
        for (int t = 0; t < threads; ++t)
            pool.emplace_back(body, (size_t)n * t / threads, (size_t)n * (t + 1) / threads, t);
        for (auto& th : pool) th.join();
    };
    // scene box
    std::vector<palz_aabb> part(threads);
    par([&](size_t b, size_t e, int t) { orc_scene_aabb(tris + b, (int)(e - b), &part[t]); });
    palz_aabb scene = part[0];
    for (int t = 1; t < threads; ++t)
        if (part[t].lower.x <= part[t].upper.x) scene = box_union(scene, part[t]);
    // keys as (key, index) words
    std::vector<uint64_t> words(n), tmp(n);
    par([&](size_t b, size_t e, int) {
        std::vector<uint32_t> k(e - b);
        orc_morton_keys(tris + b, (int)(e - b), &scene, option, bits, k.data());
        for (size_t i = b; i < e; ++i) words[i] = (uint64_t)k[i - b] | ((uint64_t)i << 32);
    });
    // stable LSD radix sort on bits [0, 32)
    std::vector<uint32_t> counts((size_t)threads * 256);
    for (int pass = 0; pass < 4; ++pass) {
        const int shift = 8 * pass;
        par([&](size_t b, size_t e, int t) {
            uint32_t* c = &counts[(size_t)t * 256];
            std::fill(c, c + 256, 0u);
            for (size_t i = b; i < e; ++i) ++c[(words[i] >> shift) & 0xFFu];
        });
        // exclusive offsets, digit-major then thread order (stability)
        uint32_t acc = 0;
        for (int d = 0; d < 256; ++d)
            for (int t = 0; t < threads; ++t) {
                const uint32_t c = counts[(size_t)t * 256 + d];
                counts[(size_t)t * 256 + d] = acc;
                acc += c;
            }
        par([&](size_t b, size_t e, int t) {
            uint32_t* off = &counts[(size_t)t * 256];
            for (size_t i = b; i < e; ++i) tmp[off[(words[i] >> shift) & 0xFFu]++] = words[i];
        });
        words.swap(tmp);
    }
    std::vector<uint8_t> d(n > 1 ? n - 1 : 1);
    par([&](size_t b, size_t e, int) {
        for (size_t i = b; i < e && i + 1 < (size_t)n; ++i)
            d[i] = (uint8_t)delta_of((uint32_t)words[i], (uint32_t)words[i + 1]);
    });
    // bottom-up climb
    const uint32_t n_int = (uint32_t)n - 1u;
    std::vector<std::atomic<uint32_t>> arrival(n_int > 0 ? n_int : 1);
    for (auto& a : arrival) a.store(0xFFFFFFFFu, std::memory_order_relaxed);
    std::atomic<uint32_t> root_node{0xFFFFFFFFu};
    if (n == 1) root_node = (uint32_t)(words[0] >> 32) | kLeaf;
    par([&](size_t b, size_t e, int) {
        for (uint32_t leaf = (uint32_t)b; leaf < (uint32_t)e && n > 1; ++leaf) {
            uint32_t lo = leaf, hi = leaf;
            uint32_t node = (uint32_t)(words[leaf] >> 32) | kLeaf;
            palz_aabb box = tri_box(tris[node & ~kLeaf]);
            for (;;) {
                const uint32_t dl = lo > 0 ? d[lo - 1] : 0xFFFFFFFFu;
                const uint32_t dr = hi < n_int ? d[hi] : 0xFFFFFFFFu;
                const int go_left = dl < dr ? 1 : 0;
                const uint32_t parent = go_left ? lo - 1 : hi;
                nodes[parent].children[go_left] = node;
                nodes[parent].aabbs[go_left] = box;
                if (!(node & kLeaf)) nodes[node].parent = parent;
                const uint32_t other = arrival[parent].exchange(go_left ? hi : lo, std::memory_order_acq_rel);
                if (other == 0xFFFFFFFFu) break;
                lo = std::min(lo, other);
                hi = std::max(hi, other);
                nodes[parent].context = hi < n_int ? hi : 0xFFFFFFFFu;
                node = parent;
                box = box_union(box, nodes[parent].aabbs[go_left ^ 1]);
                if (hi - lo == n_int) {
                    nodes[node].parent = kInvalid;
                    root_node.store(node);
                    break;
                }
            }
        }
    });
    *root = root_node.load();
}

void orc_lbvh_topdown(const palz_triangle* tris, int n, const uint64_t* sorted_words, const uint8_t* deltas,
                      palz_internal_node* nodes, uint32_t* root) {
    std::vector<uint32_t> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = (uint32_t)(sorted_words[i] >> 32);
    palz_aabb box;
    const uint32_t r = topdown(idx, deltas, tris, 0, (uint32_t)n - 1u, (uint32_t)n - 1u, nodes, &box);
    *root = r;
    if (!(r & kLeaf)) nodes[r].parent = kInvalid;
}

// The InternalNormalBound pass (main.cpp:248-295), serially: for each internal
// node's leaf children, climb by parent links, storing the subtree's box of
// shading normals into the parent's slot; the second arrival merges and goes
// on.
void orc_normal_bounds(const palz_internal_node* nodes, int n_int, const palz_triangle_attrib* attribs,
                       palz_normal_bound* out) {
    for (int i = 0; i < n_int; ++i) out[i].counter = 0u;
    for (int i = 0; i < n_int; ++i) {
        for (int j = 0; j < 2; ++j) {
            uint32_t me = nodes[i].children[j];
            if (!(me & kLeaf)) continue;
            const palz_triangle_attrib& at = attribs[me & ~kLeaf];
            float lo[3] = {+3.40282347e+38f, +3.40282347e+38f, +3.40282347e+38f};
            float hi[3] = {-3.40282347e+38f, -3.40282347e+38f, -3.40282347e+38f};
            for (int k = 0; k < 3; ++k) {
                const float p[3] = {at.shading_normals[k].x, at.shading_normals[k].y, at.shading_normals[k].z};
                for (int a = 0; a < 3; ++a) {
                    lo[a] = box_min(lo[a], p[a]);
                    hi[a] = box_max(hi[a], p[a]);
                }
            }
            uint32_t parent = (uint32_t)i;
            while (parent != kInvalid) {
                const int c = nodes[parent].children[0] == me ? 0 : 1;
                palz_aabb& dst = out[parent].normal_bounds[c];
                dst.lower = {lo[0], lo[1], lo[2]};
                dst.upper = {hi[0], hi[1], hi[2]};
                if (out[parent].counter++ == 0u) break;
                const palz_aabb& sb = out[parent].normal_bounds[c ^ 1];
                const float slo[3] = {sb.lower.x, sb.lower.y, sb.lower.z};
                const float shi[3] = {sb.upper.x, sb.upper.y, sb.upper.z};
                for (int a = 0; a < 3; ++a) {
                    lo[a] = box_min(lo[a], slo[a]);
                    hi[a] = box_max(hi[a], shi[a]);
                }
                me = parent;
                parent = nodes[me].parent;
            }
        }
    }
}

// 1 when deltas never increase going down the tree (validate_lbvh).
int orc_validate_lbvh(const palz_internal_node* nodes, uint32_t root, const uint8_t* deltas) {
    return validate(root, nodes, deltas, 255u) ? 1 : 0;
}

}  // extern "C"
