// Photon tracing, the deterministic tuple table and the receiver lookup.
//
//   photonTrace ........... main_gpu.cu:56-211
//   PathCache::store ...... pk.h:291-349 (made deterministic, see below)
//   lookupFlatten ......... main_gpu.cu:373-416
//   sort + reorder ........ pk.h:1526-1541, main_gpu.cu:515-525
//
// The reference's PathCache is an open-addressing hash table filled by racing
// photons: it dedupes on the 32-bit path hash alone and keeps the parameters
// of whichever photon lands first.  Here the same set semantics are built with
// sorts, so the result depends on nothing but the photons:
//   1. photon tuples are sorted by a 64-bit content hash (stable, so equal
//      tuples stay in photon order) and numbered; equal hashes with different
//      content are detected and reported;
//   2. each photon emits its 2x2x2 nearest cells (pk.h:302-313) as keys
//      (cell x,y,z: 14 bits each, tuple id: 22 bits);
//   3. a stable sort and a first-of-run filter keep one entry per
//      (cell, tuple), seeded by the smallest photon id;
//   4. receiver points binary-search their own cell (pk.h:352-373).
#include <algorithm>

#include "optics.hpp"
#include "prims.hpp"
#include "sort.hpp"

namespace palz {

namespace {

constexpr int kTraceThreads = 128;
constexpr int kTraceStack = 64;      // traversal stack entries per photon
constexpr int kTraceStackLds = 16;   // of which in LDS; the rest in private memory
constexpr int kCellBits = 14;
constexpr int kTupleBits = 22;
constexpr uint32_t kCellMax = (1u << kCellBits) - 1u;
constexpr uint32_t kTupleMax = (1u << kTupleBits) - 1u;

enum : uint32_t { kStatusCellRange = 1u, kStatusTupleCount = 2u, kStatusHashCollision = 4u };

__global__ void __launch_bounds__(kTraceThreads)
    k_photon_trace(palz_scene_view sc, palz_photon_params pp, palz_photon_record* __restrict__ out,
                   uint32_t* __restrict__ status) {
    __shared__ uint32_t s_stack[kTraceStackLds * kTraceThreads];
    SplitStack<kTraceStackLds, kTraceStack - kTraceStackLds> stack;
    stack.near = StridedStack{s_stack + threadIdx.x, kTraceThreads};
    const uint64_t g = (uint64_t)blockIdx.x * kTraceThreads + threadIdx.x;
    const uint32_t P = pp.photons_per_triangle;
    if (g >= (uint64_t)sc.n_tris * P) return;
    const uint32_t i_tri = (uint32_t)(g / P);
    const uint32_t j = (uint32_t)(g % P);

    palz_photon_record rec;
    rec.photon = (uint32_t)g;
    rec.event = 0u;
#pragma unroll
    for (int k = 0; k < 4; ++k) rec.tris[k] = -1;
#pragma unroll
    for (int k = 0; k < 8; ++k) rec.params[k] = 0.0f;
    rec.p_final = {0.0f, 0.0f, 0.0f};
    rec.throughput = 0.0f;

    const palz_triangle_attrib at = sc.attribs[i_tri];
    if (at.material != PALZ_MATERIAL_DIFFUSE) {
        const palz_triangle tri = sc.tris[i_tri];
        const v3 v0 = mk(tri.vs[0]), v1 = mk(tri.vs[1]), v2 = mk(tri.vs[2]);
        const v3 p_light = mk(pp.p_light);
        // back-face skip with the winding aligned to the shading normal (main_gpu.cu:68-75)
        v3 ng = cross(v1 - v0, v2 - v0);
        ng = ng * dot(ng, mk(at.shading_normals[0]));
        if (!(dot(ng, p_light - v0) < 0.0f)) {
            const uint32_t root = *sc.root;
            float px, py;
            sobol2d(&px, &py, j, (uint32_t)pp.iteration, i_tri, 789u);
            square_to_triangle(&px, &py);
            float ex, ey;
            sobol2d(&ex, &ey, j, (uint32_t)pp.iteration, i_tri, 178u);
            const float eta = pp.eta_min + ex * (pp.eta_max - pp.eta_min);  // lerp
            const v3 p = v0 + (v1 - v0) * px + (v2 - v0) * py;
            v3 ro = p_light;
            v3 rd = p - p_light;

            uint32_t events = 0u;
            int size = 0;
            bool admissible = false;
            int tris_k[4];
            float prm[8];
            v3 p_final = mk(0.0f, 0.0f, 0.0f);
            float throughput = 1.0f;
            for (int d = 0; d < 4 + 1; ++d) {
                Hit hit;
                hit.t = kFltMax;
                hit.u = hit.v = 0.0f;
                hit.ng = mk(0.0f, 0.0f, 0.0f);
                hit.tri = 0xFFFFFFFFu;
                if (!closest_hit_on(&hit, sc.nodes, sc.tris, root, ro, rd, inv_rd(rd), stack, kTraceStack)) {
                    if (status) atomicOr(status, 1u << 8);
                    admissible = false;
                    break;
                }
                if (hit.t == kFltMax) break;
                const palz_triangle_attrib ha = sc.attribs[hit.tri];
                const v3 p_hit = ro + rd * hit.t;
                const v3 s0 = mk(ha.shading_normals[0]);
                const v3 ns = s0 + (mk(ha.shading_normals[1]) - s0) * hit.u + (mk(ha.shading_normals[2]) - s0) * hit.v;
                if (ha.material == PALZ_MATERIAL_DIFFUSE) {
                    if (d == 0) break;
                    admissible = true;
                    p_final = p_hit;
                    break;
                }
                if (d == 4) break;
                const v3 wi = -rd;
                tris_k[d] = (int)hit.tri;
                prm[2 * d] = hit.u;
                prm[2 * d + 1] = hit.v;
                float rx, ry;
                sobol2d(&rx, &ry, j, (uint32_t)d, i_tri, (uint32_t)pp.iteration);
                const float refl = fresnel_exact_norm_free(wi, ns, eta);
                bool e_T = false;
                if (ha.material == PALZ_MATERIAL_DIELECTRIC) e_T = !(rx < refl);
                if (e_T) events |= 1u << size;
                ++size;
                float sign = 1.0f;
                v3 wo;
                if (!e_T) {
                    wo = reflection(wi, ns);
                    if (ha.material == PALZ_MATERIAL_DIELECTRIC) throughput *= refl;
                } else {
                    if (!refraction_norm_free(&wo, wi, ns, eta)) break;
                    sign = -1.0f;
                    throughput *= 1.0f - refl;
                }
                if (0.0f < dot(hit.ng, wi) * dot(hit.ng, wo) * sign) {  // geometrically admissible
                    const v3 offset_dir = normalize(hit.ng * dot(hit.ng, wo));
                    ro = p_hit + offset_dir * ray_offset_scale(p_hit);
                    rd = wo;
                    continue;
                }
                break;
            }
            if (admissible && pp.min_throughput < throughput) {
                rec.event = ((uint32_t)size << 24) | events;
                for (int k = 0; k < size; ++k) {
                    rec.tris[k] = tris_k[k];
                    rec.params[2 * k] = prm[2 * k];
                    rec.params[2 * k + 1] = prm[2 * k + 1];
                }
                rec.p_final = {p_final.x, p_final.y, p_final.z};
                rec.throughput = throughput;
            }
        }
    }
    out[g] = rec;
}

// ---------------------------------------------------------------------------
// tuple table
__device__ __forceinline__ uint64_t mix64(uint64_t z) {  // splitmix64 finaliser
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ uint64_t tuple_hash(const palz_photon_record& r) {
    uint64_t h = mix64(0x9E3779B97F4A7C15ull ^ (uint64_t)r.event);
    const int size = ev_size(r.event);
    for (int k = 0; k < size; ++k) h = mix64(h ^ ((uint64_t)(uint32_t)r.tris[k] + 0x632BE59BD9B4E019ull * (k + 1)));
    return h == ~0ull ? h - 1ull : h;  // ~0 marks "no tuple"
}
__device__ __forceinline__ bool same_tuple(const palz_photon_record& a, const palz_photon_record& b) {
    if (a.event != b.event) return false;
    const int size = ev_size(a.event);
    for (int k = 0; k < size; ++k)
        if (a.tris[k] != b.tris[k]) return false;
    return true;
}

__global__ void __launch_bounds__(256) k_tuple_hash(const palz_photon_record* __restrict__ rec, uint32_t n,
                                                    uint64_t* __restrict__ words, uint32_t* __restrict__ vals,
                                                    uint32_t* __restrict__ n_valid) {
    // grid-stride, valid photons counted per block: one atomic per block
    __shared__ uint32_t s_count;
    if (threadIdx.x == 0) s_count = 0u;
    __syncthreads();
    uint32_t count = 0u;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const palz_photon_record r = rec[i];
        const bool valid = r.event != 0u;
        words[i] = valid ? tuple_hash(r) : ~0ull;
        vals[i] = i;
        count += valid ? 1u : 0u;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) count += (uint32_t)__shfl_xor((int)count, o, 64);
    if ((threadIdx.x & 63) == 0 && count) atomicAdd(&s_count, count);
    __syncthreads();
    if (threadIdx.x == 0 && s_count) atomicAdd(n_valid, s_count);
}

// The smallest cell of the valid photons' end points per axis (the key
// origin): a grid-stride loop reduced per block, one atomic per block and axis
// (atomics from every wave on the same three words serialise at the L2).
__global__ void __launch_bounds__(256) k_cell_min(const palz_photon_record* __restrict__ rec, uint32_t n,
                                                  float step, int* __restrict__ cell_min) {
    __shared__ int s_min[3][256 / 64];
    int m[3] = {0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (rec[i].event != 0u) {
            const palz_float3 p = rec[i].p_final;
            m[0] = min(m[0], (int)floorf(p.x / step));
            m[1] = min(m[1], (int)floorf(p.y / step));
            m[2] = min(m[2], (int)floorf(p.z / step));
        }
    }
    const uint32_t wave = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        int v = m[c];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
        if ((threadIdx.x & 63) == 0) s_min[c][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        int v = 0x7FFFFFFF;
        for (uint32_t w = 0; w < blockDim.x / 64u; ++w) v = min(v, s_min[threadIdx.x][w]);
        if (v != 0x7FFFFFFF) atomicMin(cell_min + threadIdx.x, v);
    }
}

__global__ void __launch_bounds__(256) k_tuple_flags(const uint64_t* __restrict__ words, const uint32_t* __restrict__ vals,
                                                     const palz_photon_record* __restrict__ rec,
                                                     const uint32_t* __restrict__ n_valid, uint32_t* __restrict__ flags,
                                                     uint32_t* __restrict__ status) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *n_valid) return;
    uint32_t f = 1u;
    if (i > 0 && words[i] == words[i - 1]) {
        f = 0u;
        if (!same_tuple(rec[vals[i]], rec[vals[i - 1]])) atomicOr(status, kStatusHashCollision);
    }
    flags[i] = f;
}

// Entries: 8 per valid photon, key = cell(42 bits) << 22 | tuple id.
__global__ void __launch_bounds__(256) k_cell_entries(const uint32_t* __restrict__ vals, const uint32_t* __restrict__ tid_excl,
                                                      const uint32_t* __restrict__ flags,
                                                      const palz_photon_record* __restrict__ rec,
                                                      const uint32_t* __restrict__ n_valid, const int* __restrict__ cell_min,
                                                      float step, uint64_t* __restrict__ keys,
                                                      uint32_t* __restrict__ photons, uint32_t* __restrict__ status) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *n_valid) return;
    const uint32_t slot = vals[i];
    const uint32_t tuple = tid_excl[i] + flags[i] - 1u;
    if (tuple > kTupleMax) atomicOr(status, kStatusTupleCount);
    const palz_float3 p = rec[slot].p_final;
    // PathCache::store cell choice (pk.h:302-313)
    const float fx = p.x / step, fy = p.y / step, fz = p.z / step;
    const int x = (int)floorf(fx), y = (int)floorf(fy), z = (int)floorf(fz);
    const int dx = fx - (float)x < 0.5f ? -1 : 1;
    const int dy = fy - (float)y < 0.5f ? -1 : 1;
    const int dz = fz - (float)z < 0.5f ? -1 : 1;
    const int bx = cell_min[0] - 1, by = cell_min[1] - 1, bz = cell_min[2] - 1;
    for (int iz = 0; iz < 2; ++iz)
        for (int iy = 0; iy < 2; ++iy)
            for (int ix = 0; ix < 2; ++ix) {
                const int cx = x + ix * dx - bx, cy = y + iy * dy - by, cz = z + iz * dz - bz;
                uint64_t key = ~0ull;
                if (cx < 0 || cy < 0 || cz < 0 || (uint32_t)cx > kCellMax || (uint32_t)cy > kCellMax ||
                    (uint32_t)cz > kCellMax) {
                    atomicOr(status, kStatusCellRange);
                } else {
                    key = ((uint64_t)cx << 50) | ((uint64_t)cy << 36) | ((uint64_t)cz << 22) |
                          (uint64_t)(tuple & kTupleMax);
                }
                const uint32_t e = i * 8u + (uint32_t)(iz * 4 + iy * 2 + ix);
                keys[e] = key;
                photons[e] = slot;
            }
}

__global__ void __launch_bounds__(256) k_first_of_run(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ n_dev,
                                                      int shift, uint32_t* __restrict__ flags) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = *n_dev;
    if (i >= n) return;
    const uint64_t k = keys[i] >> shift;
    flags[i] = (keys[i] != ~0ull) && (i == 0 || k != (keys[i - 1] >> shift)) ? 1u : 0u;
}

__global__ void __launch_bounds__(256) k_compact_entries(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ photons,
                                                         const uint32_t* __restrict__ flags, const uint32_t* __restrict__ pos,
                                                         const uint32_t* __restrict__ n_dev, uint64_t* __restrict__ ukeys,
                                                         uint32_t* __restrict__ uphotons) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *n_dev) return;
    if (flags[i]) {
        ukeys[pos[i]] = keys[i];
        uphotons[pos[i]] = photons[i];
    }
}

__global__ void __launch_bounds__(256) k_compact_cells(const uint64_t* __restrict__ ukeys, const uint32_t* __restrict__ flags,
                                                       const uint32_t* __restrict__ pos, const uint32_t* __restrict__ n_entries,
                                                       const uint32_t* __restrict__ n_cells, uint64_t* __restrict__ cell_keys,
                                                       uint32_t* __restrict__ cell_begin) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = *n_entries;
    if (i == 0) cell_begin[*n_cells] = n;
    if (i >= n) return;
    if (flags[i]) {
        cell_keys[pos[i]] = ukeys[i] >> kTupleBits;
        cell_begin[pos[i]] = i;
    }
}

// ---------------------------------------------------------------------------
// lookup
__device__ __forceinline__ bool cell_range(const palz_first_diffuse& q, float step, const int* cell_min,
                                           const uint64_t* cell_keys, const uint32_t* cell_begin, uint32_t n_cells,
                                           uint32_t* b, uint32_t* e) {
    *b = *e = 0u;
    if (q.R.x == 0.0f && q.R.y == 0.0f && q.R.z == 0.0f) return false;  // invalid (main_gpu.cu:385)
    // spacial_hash cell (pk.h:219-225)
    const int cx = (int)floorf(q.p.x / step) - (cell_min[0] - 1);
    const int cy = (int)floorf(q.p.y / step) - (cell_min[1] - 1);
    const int cz = (int)floorf(q.p.z / step) - (cell_min[2] - 1);
    if (cx < 0 || cy < 0 || cz < 0 || (uint32_t)cx > kCellMax || (uint32_t)cy > kCellMax || (uint32_t)cz > kCellMax)
        return false;
    const uint64_t key = ((uint64_t)cx << 28) | ((uint64_t)cy << 14) | (uint64_t)cz;
    uint32_t lo = 0, hi = n_cells;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (cell_keys[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    if (lo >= n_cells || cell_keys[lo] != key) return false;
    *b = cell_begin[lo];
    *e = cell_begin[lo + 1];
    return true;
}

__global__ void __launch_bounds__(256) k_lookup_count(const palz_first_diffuse* __restrict__ pts, uint32_t n_pts, float step,
                                                      const int* __restrict__ cell_min, const uint64_t* __restrict__ cell_keys,
                                                      const uint32_t* __restrict__ cell_begin, const uint32_t* __restrict__ n_cells,
                                                      const uint32_t* __restrict__ uphotons,
                                                      const palz_photon_record* __restrict__ rec, uint32_t event,
                                                      uint32_t* __restrict__ counts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pts) return;
    uint32_t b, e, c = 0;
    if (cell_range(pts[i], step, cell_min, cell_keys, cell_begin, *n_cells, &b, &e))
        for (uint32_t k = b; k < e; ++k)
            if (event == 0u || rec[uphotons[k]].event == event) ++c;
    counts[i] = c;
}

__global__ void __launch_bounds__(256) k_lookup_write(const palz_first_diffuse* __restrict__ pts, uint32_t n_pts, float step,
                                                      const int* __restrict__ cell_min, const uint64_t* __restrict__ cell_keys,
                                                      const uint32_t* __restrict__ cell_begin, const uint32_t* __restrict__ n_cells,
                                                      const uint64_t* __restrict__ ukeys, const uint32_t* __restrict__ uphotons,
                                                      const palz_photon_record* __restrict__ rec, uint32_t event,
                                                      const uint32_t* __restrict__ offsets, uint32_t scene, uint32_t capacity,
                                                      palz_path* __restrict__ paths, uint64_t* __restrict__ sort_words) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pts) return;
    uint32_t b, e;
    if (!cell_range(pts[i], step, cell_min, cell_keys, cell_begin, *n_cells, &b, &e)) return;
    uint32_t o = offsets[i];
    for (uint32_t k = b; k < e; ++k) {
        const palz_photon_record& r = rec[uphotons[k]];
        if (event != 0u && r.event != event) continue;
        if (o < capacity) {
            palz_path pa;
            pa.scene = scene;
            pa.pixel = i;
            pa.event = r.event;
            pa.photon = r.photon;
#pragma unroll
            for (int t = 0; t < 4; ++t) pa.tris[t] = r.tris[t];
#pragma unroll
            for (int t = 0; t < 8; ++t) pa.params[t] = r.params[t];
            paths[o] = pa;
            // coherence order: chain length, then tuple (pk.h:1526 sorts by path hash)
            const uint32_t tuple = (uint32_t)(ukeys[k] & kTupleMax);
            sort_words[o] = ((uint64_t)((uint32_t)ev_size(r.event) << kTupleBits | tuple)) | ((uint64_t)o << 32);
        }
        ++o;
    }
}

__global__ void __launch_bounds__(256) k_gather_paths(const uint64_t* __restrict__ words, const palz_path* __restrict__ src,
                                                      palz_path* __restrict__ dst, const uint32_t* __restrict__ n_dev,
                                                      uint32_t capacity) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= min(*n_dev, capacity)) return;
    dst[i] = src[(uint32_t)(words[i] >> 32)];
}

__global__ void k_init_table(uint32_t* scalars, int* cell_min) {
    if (threadIdx.x < 8) scalars[threadIdx.x] = 0u;
    if (threadIdx.x < 3) cell_min[threadIdx.x] = 0x7FFFFFFF;
}

__global__ void k_scale(const uint32_t* in, uint32_t mul, uint32_t* out) { *out = *in * mul; }

}  // namespace

}  // namespace palz

struct palz_tuple_table {
    uint32_t cap_records = 0;
    float step = 0.01f;
    uint32_t n_records = 0;
    // device scalars: [0] n_valid, [1] n_tuples, [2] n_entries, [3] n_cells, [4] status, [5] path count tmp
    uint32_t* scalars = nullptr;
    int* cell_min = nullptr;
    uint64_t* words = nullptr;  // hashes, then entry keys (8 per record)
    uint64_t* words_tmp = nullptr;
    uint32_t* vals = nullptr;
    uint32_t* vals_tmp = nullptr;
    uint32_t* flags = nullptr;
    uint32_t* pos = nullptr;
    uint32_t* scan_scratch = nullptr;
    uint64_t* ukeys = nullptr;
    uint32_t* uphotons = nullptr;
    uint64_t* cell_keys = nullptr;
    uint32_t* cell_begin = nullptr;
    const palz_photon_record* records = nullptr;
    // lookup scratch
    uint32_t cap_points = 0;
    uint32_t* counts = nullptr;
    uint32_t* pts_scan = nullptr;
    uint32_t cap_paths = 0;
    palz_path* paths_tmp = nullptr;
    uint64_t* path_words = nullptr;
    uint64_t* path_words_tmp = nullptr;

    void free_all() {
        for (void* p : {(void*)words, (void*)words_tmp, (void*)vals, (void*)vals_tmp, (void*)flags, (void*)pos,
                        (void*)scan_scratch, (void*)ukeys, (void*)uphotons, (void*)cell_keys, (void*)cell_begin})
            if (p) (void)hipFree(p);
        words = words_tmp = ukeys = cell_keys = nullptr;
        vals = vals_tmp = flags = pos = scan_scratch = uphotons = cell_begin = nullptr;
        cap_records = 0;
    }
    int ensure(uint32_t n) {
        if (n <= cap_records) return PALZ_OK;
        free_all();
        const size_t e = (size_t)n * 8u;
        PALZ_HIP(hipMalloc(&words, sizeof(uint64_t) * e));
        PALZ_HIP(hipMalloc(&words_tmp, sizeof(uint64_t) * e));
        PALZ_HIP(hipMalloc(&vals, sizeof(uint32_t) * e));
        PALZ_HIP(hipMalloc(&vals_tmp, sizeof(uint32_t) * e));
        PALZ_HIP(hipMalloc(&flags, sizeof(uint32_t) * e));
        PALZ_HIP(hipMalloc(&pos, sizeof(uint32_t) * e));
        PALZ_HIP(hipMalloc(&scan_scratch, sizeof(uint32_t) * palz::scan_scratch_entries((uint32_t)e)));
        PALZ_HIP(hipMalloc(&ukeys, sizeof(uint64_t) * e));
        PALZ_HIP(hipMalloc(&uphotons, sizeof(uint32_t) * e));
        PALZ_HIP(hipMalloc(&cell_keys, sizeof(uint64_t) * e));
        PALZ_HIP(hipMalloc(&cell_begin, sizeof(uint32_t) * (e + 1)));
        cap_records = n;
        return PALZ_OK;
    }
    ~palz_tuple_table() {
        free_all();
        if (scalars) (void)hipFree(scalars);
        if (cell_min) (void)hipFree(cell_min);
        if (counts) (void)hipFree(counts);
        if (pts_scan) (void)hipFree(pts_scan);
        if (paths_tmp) (void)hipFree(paths_tmp);
        if (path_words) (void)hipFree(path_words);
        if (path_words_tmp) (void)hipFree(path_words_tmp);
    }
};

extern "C" {

int palz_photon_trace(const palz_scene_view* scene, const palz_photon_params* params, palz_photon_record* d_out,
                      uint32_t* d_status, void* stream) {
    using namespace palz;
    if (!scene || !params || !d_out) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_photon_trace: NULL argument");
    if (!scene->tris || !scene->attribs || !scene->nodes || !scene->root)
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_photon_trace: incomplete scene view");
    if (params->photons_per_triangle == 0 || scene->n_tris < 2)
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_photon_trace: need >= 2 triangles and >= 1 photon each");
    const uint64_t total = (uint64_t)scene->n_tris * params->photons_per_triangle;
    if (total >= (1ull << 32)) return fail(PALZ_ERR_CAPACITY, "palz_photon_trace: more than 2^32 photons");
    const uint32_t blocks = (uint32_t)((total + kTraceThreads - 1) / kTraceThreads);
    k_photon_trace<<<blocks, kTraceThreads, 0, (hipStream_t)stream>>>(*scene, *params, d_out, d_status);
    PALZ_LAUNCH_CHECK();
    return PALZ_OK;
}

int palz_tuple_table_create(palz_tuple_table** out) {
    if (!out) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_tuple_table_create: NULL");
    palz_tuple_table* t = new palz_tuple_table();
    if (hipMalloc(&t->scalars, sizeof(uint32_t) * 8) != hipSuccess || hipMalloc(&t->cell_min, sizeof(int) * 4) != hipSuccess) {
        delete t;
        return palz::fail(PALZ_ERR_OUT_OF_MEMORY, "palz_tuple_table_create: hipMalloc failed");
    }
    *out = t;
    return PALZ_OK;
}

int palz_tuple_table_destroy(palz_tuple_table* t) {
    delete t;
    return PALZ_OK;
}

int palz_tuple_table_build(palz_tuple_table* t, const palz_photon_record* d_records, uint32_t n_records, float cell_step,
                           palz_sorter* sorter, void* stream) {
    using namespace palz;
    if (!t || !sorter || (!d_records && n_records)) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_tuple_table_build: NULL");
    if (!(cell_step > 0.0f)) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_tuple_table_build: cell_step must be > 0");
    if ((uint64_t)n_records * 8u >= (uint64_t)kSortMaxN)
        return fail(PALZ_ERR_CAPACITY, "palz_tuple_table_build: too many records (8 entries each must stay < 2^30)");
    const uint32_t cap = std::max(n_records, 1u);
    int rc = t->ensure(cap);
    if (rc != PALZ_OK) return rc;
    rc = sorter->reserve(cap * 8u);
    if (rc != PALZ_OK) return rc;
    t->step = cell_step;
    t->n_records = n_records;
    t->records = d_records;
    hipStream_t st = (hipStream_t)stream;
    uint32_t* n_valid = t->scalars + 0;
    uint32_t* n_tuples = t->scalars + 1;
    uint32_t* n_entries = t->scalars + 2;
    uint32_t* n_cells = t->scalars + 3;
    uint32_t* status = t->scalars + 4;
    uint32_t* n_raw = t->scalars + 5;  // 8 * n_valid
    k_init_table<<<1, 64, 0, st>>>(t->scalars, t->cell_min);
    const uint32_t nb = div_up(cap, 256u);
    k_tuple_hash<<<std::min(nb, 1024u), 256, 0, st>>>(d_records, n_records, t->words, t->vals, n_valid);
    k_cell_min<<<std::min(nb, 1024u), 256, 0, st>>>(d_records, n_records, cell_step, t->cell_min);
    k_scale<<<1, 1, 0, st>>>(n_valid, 8u, n_raw);
    const SortScratch& s = sorter->scratch;
    // 1. tuples by content hash; stable, so equal tuples keep photon order.
    //    Every slot takes part: slots without a tuple carry ~0 and sort last.
    launch_sort_prepare(s, nullptr, st);
    launch_sort_count(s, t->words, n_records, 0, 64, st);
    launch_sort_passes(s, t->words, t->words_tmp, t->vals, t->vals_tmp, n_records, 0, 64, st);
    // 2. dense tuple ids
    k_tuple_flags<<<nb, 256, 0, st>>>(t->words, t->vals, d_records, n_valid, t->flags, status);
    launch_exclusive_scan(t->flags, t->pos, cap, n_valid, n_tuples, t->scan_scratch, st);
    // 3. 8 cell entries per valid photon, in photon order within each tuple
    k_cell_entries<<<nb, 256, 0, st>>>(t->vals, t->pos, t->flags, d_records, n_valid, t->cell_min, cell_step,
                                       t->words_tmp, t->vals_tmp, status);
    const uint32_t ecap = cap * 8u;
    const uint32_t nbe = div_up(ecap, 256u);
    launch_sort_prepare(s, nullptr, st);
    launch_sort_count(s, t->words_tmp, ecap, 0, 64, st, n_raw);
    launch_sort_passes(s, t->words_tmp, t->words, t->vals_tmp, t->vals, ecap, 0, 64, st, nullptr, n_raw);
    // 4. one entry per (cell, tuple): the first of each run has the smallest photon id
    k_first_of_run<<<nbe, 256, 0, st>>>(t->words_tmp, n_raw, 0, t->flags);
    launch_exclusive_scan(t->flags, t->pos, ecap, n_raw, n_entries, t->scan_scratch, st);
    k_compact_entries<<<nbe, 256, 0, st>>>(t->words_tmp, t->vals_tmp, t->flags, t->pos, n_raw, t->ukeys, t->uphotons);
    // 5. cell list
    k_first_of_run<<<nbe, 256, 0, st>>>(t->ukeys, n_entries, kTupleBits, t->flags);
    launch_exclusive_scan(t->flags, t->pos, ecap, n_entries, n_cells, t->scan_scratch, st);
    k_compact_cells<<<nbe, 256, 0, st>>>(t->ukeys, t->flags, t->pos, n_entries, n_cells, t->cell_keys, t->cell_begin);
    PALZ_LAUNCH_CHECK();
    return PALZ_OK;
}

int palz_tuple_table_info(palz_tuple_table* t, uint32_t* n_valid, uint32_t* n_tuples, uint32_t* n_entries,
                          uint32_t* n_cells) {
    using namespace palz;
    if (!t) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_tuple_table_info: NULL");
    uint32_t h[8] = {};
    PALZ_HIP(hipDeviceSynchronize());
    PALZ_HIP(hipMemcpy(h, t->scalars, sizeof(h), hipMemcpyDeviceToHost));
    if (n_valid) *n_valid = h[0];
    if (n_tuples) *n_tuples = h[1];
    if (n_entries) *n_entries = h[2];
    if (n_cells) *n_cells = h[3];
    if (h[4] & kStatusHashCollision) return fail(PALZ_ERR_CAPACITY, "tuple table: 64-bit tuple hash collision");
    if (h[4] & kStatusTupleCount) return fail(PALZ_ERR_CAPACITY, "tuple table: more than 2^22 distinct tuples");
    if (h[4] & kStatusCellRange) return fail(PALZ_ERR_CAPACITY, "tuple table: landing points span more than 2^14 cells per axis");
    return PALZ_OK;
}

int palz_lookup_flatten(palz_tuple_table* t, uint32_t event, const palz_first_diffuse* d_points, uint32_t n_points,
                        uint32_t scene, palz_path* d_paths, uint32_t capacity, uint32_t* d_count, palz_sorter* sorter,
                        void* stream) {
    using namespace palz;
    if (!t || !sorter || !d_count || (!d_points && n_points) || (!d_paths && capacity))
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_lookup_flatten: NULL argument");
    if (!t->records && t->n_records) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_lookup_flatten: table not built");
    if (capacity > kSortMaxN) return fail(PALZ_ERR_CAPACITY, "palz_lookup_flatten: capacity must be < 2^30");
    hipStream_t st = (hipStream_t)stream;
    const uint32_t np = std::max(n_points, 1u);
    if (np > t->cap_points) {
        if (t->counts) PALZ_HIP(hipFree(t->counts));
        if (t->pts_scan) PALZ_HIP(hipFree(t->pts_scan));
        PALZ_HIP(hipMalloc(&t->counts, sizeof(uint32_t) * np));
        PALZ_HIP(hipMalloc(&t->pts_scan, sizeof(uint32_t) * scan_scratch_entries(np)));
        t->cap_points = np;
    }
    const uint32_t pc = std::max(capacity, 1u);
    if (pc > t->cap_paths) {
        if (t->paths_tmp) PALZ_HIP(hipFree(t->paths_tmp));
        if (t->path_words) PALZ_HIP(hipFree(t->path_words));
        if (t->path_words_tmp) PALZ_HIP(hipFree(t->path_words_tmp));
        PALZ_HIP(hipMalloc(&t->paths_tmp, sizeof(palz_path) * pc));
        PALZ_HIP(hipMalloc(&t->path_words, sizeof(uint64_t) * pc));
        PALZ_HIP(hipMalloc(&t->path_words_tmp, sizeof(uint64_t) * pc));
        t->cap_paths = pc;
    }
    int rc = sorter->reserve(pc);
    if (rc != PALZ_OK) return rc;
    const uint32_t nbp = div_up(np, 256u);
    uint32_t* n_cells = t->scalars + 3;
    k_lookup_count<<<nbp, 256, 0, st>>>(d_points, n_points, t->step, t->cell_min, t->cell_keys, t->cell_begin, n_cells,
                                        t->uphotons, t->records, event, t->counts);
    launch_exclusive_scan(t->counts, t->counts, np, nullptr, d_count, t->pts_scan, st);
    k_lookup_write<<<nbp, 256, 0, st>>>(d_points, n_points, t->step, t->cell_min, t->cell_keys, t->cell_begin, n_cells,
                                        t->ukeys, t->uphotons, t->records, event, t->counts, scene, capacity,
                                        t->paths_tmp, t->path_words);
    const SortScratch& s = sorter->scratch;
    launch_sort_prepare(s, nullptr, st);
    launch_sort_count(s, t->path_words, pc, 0, 3 + kTupleBits, st, d_count);
    launch_sort_passes(s, t->path_words, t->path_words_tmp, nullptr, nullptr, pc, 0, 3 + kTupleBits, st, nullptr,
                       d_count);  // 4 passes: result back in path_words
    k_gather_paths<<<div_up(pc, 256u), 256, 0, st>>>(t->path_words, t->paths_tmp, d_paths, d_count, capacity);
    PALZ_LAUNCH_CHECK();
    return PALZ_OK;
}

}  // extern "C"
