// A PKRenderer-shaped host driver (pk.h:1363-1640) over this library's entry
// points, in the reference's order: loadFrame builds the BVH of a scene;
// step runs solvePrimary (camera rays to the first diffuse hits, direct light
// into the float4 accumulators), photonTrace, the tuple table, lookupFlatten
// of every chain of length 1-3 and the batched solve (the reference's
// per-descriptor lookups and launches, pk.h:1480-1586), then the T,T,T,T
// chains (the reference's K = 4 launch) with the plain solver; resolve packs
// the accumulators (pack, main_gpu.cu:499-513).
//
// Spectral handling: the reference draws one wavelength per pixel and frame
// and weights each contribution by the CIE fit at that wavelength.  Here
// each frame solves a fixed list of wavelengths into per-wavelength images
// (32.32 fixed-point sums, so a frame does not depend on the order in which
// the solvers' items finish),
// which the spectral resolve weights by the CIE fit, the caller's quadrature
// weights and the frame's own receiver reflectance before they are added to
// the accumulators and cleared.
#include <cstring>
#include <vector>

#include "common.hpp"
#include "spectral.hpp"

namespace palz {
namespace {

__global__ void __launch_bounds__(256) k_add_frame(float4* __restrict__ accum, const float4* __restrict__ frame,
                                                   uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float4 a = accum[i];
    const float4 f = frame[i];
    a.x += f.x;
    a.y += f.y;
    a.z += f.z;
    accum[i] = a;
}

__global__ void __launch_bounds__(256) k_fill_zero(uint32_t* __restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = 0u;
}

// pack (main_gpu.cu:499-513): packRGBA(srgb_oetf(acc.xyz / acc.w))
__global__ void __launch_bounds__(256) k_pack(const float4* __restrict__ accum, uint32_t* __restrict__ pixels,
                                              uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float4 a = accum[i];
    pixels[i] = pack_rgba(srgb_oetf(a.x / a.w), srgb_oetf(a.y / a.w), srgb_oetf(a.z / a.w));
}

void fill_zero(void* p, size_t bytes, hipStream_t st) {
    const size_t words = bytes / 4;
    if (!words) return;
    const uint32_t blocks = (uint32_t)std::min<size_t>((words + 255) / 256, 8192);
    k_fill_zero<<<blocks, 256, 0, st>>>((uint32_t*)p, words);
}

template <class T>
int grow(T*& p, size_t& cap, size_t n) {
    if (n <= cap) return PALZ_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    PALZ_HIP(hipMalloc(&p, sizeof(T) * n));
    cap = n;
    return PALZ_OK;
}

constexpr uint32_t kEventTTTT = (4u << 24) | 0xFu;

}  // namespace
}  // namespace palz

struct palz_renderer {
    uint32_t width = 0, height = 0, n_lambdas = 0;
    palz_renderer_params params{};
    palz_camera camera{};
    bool has_camera = false;
    int32_t iteration = 0;
    // handles
    palz_sorter* sorter = nullptr;
    palz_lbvh* lbvh = nullptr;
    palz_tuple_table* table = nullptr;
    palz_batched* batched = nullptr;
    // scene (loadFrame)
    palz_triangle* d_tris = nullptr;
    palz_triangle_attrib* d_attribs = nullptr;
    size_t cap_tris = 0, cap_attribs = 0;
    uint32_t n_tris = 0;
    palz_internal_node* d_nodes = nullptr;
    palz_node_index* d_root = nullptr;
    palz_wide_node* d_wide = nullptr;
    // per pixel (allocate)
    palz_first_diffuse* d_points = nullptr;
    palz_float4* d_accum = nullptr;
    palz_float4* d_frame = nullptr;
    uint32_t* d_pixels = nullptr;
    float* d_images = nullptr;
    uint64_t* d_fixed = nullptr;  // the solvers' 32.32 fixed-point sums (order-independent)
    float* d_lambdas = nullptr;
    float* d_weights = nullptr;
    palz_scene_desc* d_desc = nullptr;
    // per frame
    palz_photon_record* d_records = nullptr;
    size_t cap_records = 0;
    palz_path* d_paths = nullptr;
    size_t cap_paths = 0;
    uint32_t* d_scalars = nullptr;  // [0] photon status, [1] path count, [2] T,T,T,T path count
    float lambda_min = 0.0f, lambda_max = 0.0f;  // nm, for the photons' eta range

    palz_scene_desc desc() const {
        palz_scene_desc d;
        std::memset(&d, 0, sizeof(d));
        d.tris = d_tris;
        d.attribs = d_attribs;
        d.nodes = d_nodes;
        d.root = d_root;
        d.points = d_points;
        d.lambdas = d_lambdas;
        d.image = d_images;
        d.n_points = width * height;
        d.n_lambdas = n_lambdas;
        d.p_light = params.p_light;
        d.light_intensity = params.light_intensity;
        d.radiance_clamp = params.radiance_clamp;
        d.cauchy_a = params.cauchy_a;
        d.cauchy_b = params.cauchy_b;
        d.wide = d_wide;
        d.image_fixed = d_fixed;
        return d;
    }
    palz_scene_view view() const {
        palz_scene_view v;
        std::memset(&v, 0, sizeof(v));
        v.tris = d_tris;
        v.attribs = d_attribs;
        v.nodes = d_nodes;
        v.root = d_root;
        v.n_tris = n_tris;
        return v;
    }
    ~palz_renderer() {
        for (void* p : {(void*)d_tris, (void*)d_attribs, (void*)d_points, (void*)d_accum, (void*)d_frame,
                        (void*)d_pixels, (void*)d_images, (void*)d_fixed, (void*)d_lambdas, (void*)d_weights,
                        (void*)d_desc,
                        (void*)d_records, (void*)d_paths, (void*)d_scalars})
            if (p) (void)hipFree(p);
        if (batched) palz_batched_destroy(batched);
        if (table) palz_tuple_table_destroy(table);
        if (lbvh) palz_lbvh_destroy(lbvh);
        if (sorter) palz_sorter_destroy(sorter);
    }
};

extern "C" {

int palz_renderer_create(uint32_t width, uint32_t height, const float* lambdas, const float* weights,
                         uint32_t n_lambdas, palz_renderer** out) {
    using namespace palz;
    if (!out || !lambdas || !weights) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_renderer_create: NULL argument");
    if (width == 0 || height == 0 || (uint64_t)width * height >= (1ull << 26))
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_renderer_create: image must have 1 .. 2^26 pixels");
    if (n_lambdas < 1 || n_lambdas > 256)
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_renderer_create: need 1..256 wavelengths");
    palz_renderer* r = new palz_renderer();
    r->width = width;
    r->height = height;
    r->n_lambdas = n_lambdas;
    palz_renderer_params& p = r->params;  // the reference's defaults (pk.h:1683-1694)
    p.p_light = {-0.804876f, 0.121239f, -1.58616f};
    p.light_intensity = 16.0f;
    p.radiance_clamp = 12.0f;
    p.min_throughput = 0.05f;
    p.cauchy_a = 1.5f;
    p.cauchy_b = 4200.0f;
    p.photons_per_triangle = 256;
    p.use_lens = 1;
    p.lens_distance = 0.6f;
    p.lens_thickness = 0.05f;
    p.lens_R = 3.0f;
    p.albedo = {0.75f, 0.75f, 0.75f};
    p.path_capacity = 1u << 24;
    p.refine = PALZ_REFINE_NONE;
    p.solve_tttt = 1;
    const size_t np = (size_t)width * height;
    int rc = PALZ_OK;
    auto ok = [&](int c) {
        if (rc == PALZ_OK && c != PALZ_OK) rc = c;
        return rc == PALZ_OK;
    };
    auto alloc = [&](void** ptr, size_t bytes) {
        if (rc == PALZ_OK && hipMalloc(ptr, bytes) != hipSuccess)
            rc = fail(PALZ_ERR_OUT_OF_MEMORY, "palz_renderer_create: hipMalloc failed");
        return rc == PALZ_OK;
    };
    ok(palz_sorter_create(&r->sorter)) && ok(palz_lbvh_create(&r->lbvh)) && ok(palz_tuple_table_create(&r->table)) &&
        ok(palz_batched_create(&r->batched)) &&
        ok(palz_batched_set_fused(r->batched, 2)) &&  // one occlusion kernel over every chain length
        alloc((void**)&r->d_points, sizeof(palz_first_diffuse) * np) &&
        alloc((void**)&r->d_accum, sizeof(palz_float4) * np) && alloc((void**)&r->d_frame, sizeof(palz_float4) * np) &&
        alloc((void**)&r->d_pixels, sizeof(uint32_t) * np) &&
        alloc((void**)&r->d_images, sizeof(float) * np * n_lambdas) &&
        alloc((void**)&r->d_fixed, sizeof(uint64_t) * np * n_lambdas) &&
        alloc((void**)&r->d_lambdas, sizeof(float) * n_lambdas) &&
        alloc((void**)&r->d_weights, sizeof(float) * n_lambdas) &&
        alloc((void**)&r->d_desc, sizeof(palz_scene_desc)) && alloc((void**)&r->d_scalars, sizeof(uint32_t) * 4);
    if (rc == PALZ_OK) {
        if (hipMemcpy(r->d_lambdas, lambdas, sizeof(float) * n_lambdas, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(r->d_weights, weights, sizeof(float) * n_lambdas, hipMemcpyHostToDevice) != hipSuccess)
            rc = fail(PALZ_ERR_HIP, "palz_renderer_create: upload failed");
    }
    if (rc == PALZ_OK) {
        hipStream_t st = nullptr;
        fill_zero(r->d_accum, sizeof(palz_float4) * np, st);
        fill_zero(r->d_fixed, sizeof(uint64_t) * np * n_lambdas, st);
        fill_zero(r->d_scalars, sizeof(uint32_t) * 4, st);
        if (hipStreamSynchronize(st) != hipSuccess) rc = fail(PALZ_ERR_HIP, "palz_renderer_create: clear failed");
    }
    if (rc != PALZ_OK) {
        delete r;
        return rc;
    }
    float lo = lambdas[0], hi = lambdas[0];
    for (uint32_t l = 1; l < n_lambdas; ++l) {
        lo = std::min(lo, lambdas[l]);
        hi = std::max(hi, lambdas[l]);
    }
    r->lambda_min = lo;
    r->lambda_max = hi;
    *out = r;
    return PALZ_OK;
}

int palz_renderer_destroy(palz_renderer* r) {
    delete r;
    return PALZ_OK;
}

int palz_renderer_set_params(palz_renderer* r, const palz_renderer_params* p) {
    using namespace palz;
    if (!r || !p) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_renderer_set_params: NULL argument");
    if (p->photons_per_triangle == 0 || p->path_capacity == 0 || p->refine < PALZ_REFINE_NONE ||
        p->refine > PALZ_REFINE_ALL)
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_renderer_set_params: bad photon budget, capacity or refine");
    if (!(p->radiance_clamp > 0.0f && p->radiance_clamp < 2147483648.0f))
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_renderer_set_params: radiance_clamp must be in (0, 2^31)");
    r->params = *p;
    return PALZ_OK;
}

int palz_renderer_get_params(const palz_renderer* r, palz_renderer_params* p) {
    if (!r || !p) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_renderer_get_params: NULL argument");
    *p = r->params;
    return PALZ_OK;
}

int palz_renderer_set_camera(palz_renderer* r, const palz_camera* c) {
    if (!r || !c) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_renderer_set_camera: NULL argument");
    r->camera = *c;
    r->has_camera = true;
    return PALZ_OK;
}

// loadFrame (pk.h:1246-1362): the scene's triangles and attributes (host
// arrays) to the device and the BVH over them.
int palz_renderer_load_scene(palz_renderer* r, const palz_triangle* tris, const palz_triangle_attrib* attribs,
                             uint32_t n, void* stream) {
    using namespace palz;
    if (!r || !tris || !attribs) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_renderer_load_scene: NULL argument");
    if (n < 2) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_renderer_load_scene: need at least 2 triangles");
    hipStream_t st = (hipStream_t)stream;
    int rc = grow(r->d_tris, r->cap_tris, n);
    if (rc == PALZ_OK) rc = grow(r->d_attribs, r->cap_attribs, n);
    if (rc != PALZ_OK) return rc;
    PALZ_HIP(hipMemcpyAsync(r->d_tris, tris, sizeof(palz_triangle) * n, hipMemcpyHostToDevice, st));
    PALZ_HIP(hipMemcpyAsync(r->d_attribs, attribs, sizeof(palz_triangle_attrib) * n, hipMemcpyHostToDevice, st));
    r->n_tris = n;
    rc = palz_lbvh_build(r->lbvh, r->d_tris, (int)n, 0, 32, r->sorter, stream);
    if (rc != PALZ_OK) return rc;
    palz_aabb* aabb = nullptr;
    int built = 0;
    rc = palz_lbvh_result(r->lbvh, &r->d_nodes, &r->d_root, &aabb, &built);
    if (rc != PALZ_OK) return rc;
    rc = palz_lbvh_widen(r->lbvh, &r->d_wide, stream);
    if (rc != PALZ_OK) return rc;
    PALZ_HIP(hipStreamSynchronize(st));
    return PALZ_OK;
}

int palz_renderer_load_obj(palz_renderer* r, const char* path, int default_material, void* stream) {
    using namespace palz;
    if (!r || !path) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_renderer_load_obj: NULL argument");
    palz_mesh* m = nullptr;
    int rc = palz_mesh_load_obj(path, default_material, &m);
    if (rc != PALZ_OK) return rc;
    uint32_t n = 0;
    palz_mesh_triangle_count(m, &n);
    std::vector<palz_triangle> t(n);
    std::vector<palz_triangle_attrib> a(n);
    palz_mesh_copy(m, t.data(), a.data());
    palz_mesh_destroy(m);
    return palz_renderer_load_scene(r, t.data(), a.data(), n, stream);
}

// step (pk.h:1393-1627)
int palz_renderer_step(palz_renderer* r, void* stream) {
    using namespace palz;
    if (!r) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_renderer_step: NULL renderer");
    if (!r->n_tris || !r->has_camera) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_renderer_step: no scene or camera");
    hipStream_t st = (hipStream_t)stream;
    const palz_renderer_params& p = r->params;
    const uint32_t np = r->width * r->height;
    const palz_scene_view view = r->view();
    const palz_scene_desc desc = r->desc();
    PALZ_HIP(hipMemcpyAsync(r->d_desc, &desc, sizeof(desc), hipMemcpyHostToDevice, st));

    // solvePrimary: first diffuse hits and direct light
    palz_primary_params pp;
    std::memset(&pp, 0, sizeof(pp));
    pp.p_light = p.p_light;
    pp.light_intensity = p.light_intensity;
    pp.radiance_clamp = p.radiance_clamp;
    pp.cauchy_a = p.cauchy_a;
    pp.cauchy_b = p.cauchy_b;
    pp.iteration = r->iteration;
    pp.width = r->width;
    pp.height = r->height;
    pp.use_lens = p.use_lens;
    pp.lens_distance = p.lens_distance;
    pp.lens_thickness = p.lens_thickness;
    pp.lens_R = p.lens_R;
    pp.albedo = p.albedo;
    int rc = palz_solve_primary(&view, &r->camera, &pp, r->d_points, r->d_accum, stream);
    if (rc != PALZ_OK) return rc;

    // photonTrace over the eta range of the frame's wavelengths (pk.h:1448-1449)
    const uint64_t n_rec = (uint64_t)r->n_tris * p.photons_per_triangle;
    if (n_rec >= (1ull << 27)) return fail(PALZ_ERR_CAPACITY, "palz_renderer_step: more than 2^27 photons");
    rc = grow(r->d_records, r->cap_records, (size_t)n_rec);
    if (rc == PALZ_OK) rc = grow(r->d_paths, r->cap_paths, p.path_capacity);
    if (rc != PALZ_OK) return rc;
    palz_photon_params php;
    std::memset(&php, 0, sizeof(php));
    php.p_light = p.p_light;
    php.eta_min = p.cauchy_a + p.cauchy_b / (r->lambda_min * r->lambda_min);  // n at the shortest wavelength
    php.eta_max = p.cauchy_a + p.cauchy_b / (r->lambda_max * r->lambda_max);
    php.iteration = r->iteration;
    php.min_throughput = p.min_throughput;
    php.photons_per_triangle = p.photons_per_triangle;
    rc = palz_photon_trace(&view, &php, r->d_records, r->d_scalars, stream);
    if (rc != PALZ_OK) return rc;
    rc = palz_tuple_table_build(r->table, r->d_records, (uint32_t)n_rec, 0.01f, r->sorter, stream);
    if (rc != PALZ_OK) return rc;

    // every chain of length 1-3 in one batch (the reference: one lookup and
    // launch per descriptor)
    rc = palz_lookup_flatten(r->table, 0, r->d_points, np, 0, r->d_paths, p.path_capacity, r->d_scalars + 1,
                             r->sorter, stream);
    if (rc != PALZ_OK) return rc;
    rc = palz_batched_prepare(r->batched, r->d_paths, p.path_capacity, r->d_scalars + 1, r->sorter, stream);
    if (rc != PALZ_OK) return rc;
    rc = palz_batched_solve(r->batched, r->d_desc, 1, r->n_lambdas, p.refine, nullptr, nullptr, stream);
    if (rc != PALZ_OK) return rc;
    if (p.solve_tttt) {  // solveSpecular(4, {T, T, T, T}) (pk.h:1563)
        rc = palz_lookup_flatten(r->table, kEventTTTT, r->d_points, np, 0, r->d_paths, p.path_capacity,
                                 r->d_scalars + 2, r->sorter, stream);
        if (rc != PALZ_OK) return rc;
        rc = palz_solve_specular(4, kEventTTTT, &desc, r->d_paths, p.path_capacity, r->d_scalars + 2, nullptr,
                                 stream);
        if (rc != PALZ_OK) return rc;
    }

    // the frame's spectral images into the accumulators, then cleared
    rc = palz_image_fixed_to_float(r->d_fixed, r->d_images, (size_t)np * r->n_lambdas, stream);
    if (rc != PALZ_OK) return rc;
    rc = palz_spectral_resolve(r->d_images, r->d_lambdas, r->d_weights, r->d_points, np, r->n_lambdas, 1.0f,
                               r->d_frame, nullptr, stream);
    if (rc != PALZ_OK) return rc;
    k_add_frame<<<div_up(np, 256u), 256, 0, st>>>(reinterpret_cast<float4*>(r->d_accum),
                                                  reinterpret_cast<const float4*>(r->d_frame), np);
    fill_zero(r->d_fixed, sizeof(uint64_t) * np * r->n_lambdas, st);
    PALZ_LAUNCH_CHECK();
    ++r->iteration;
    return PALZ_OK;
}

// clear (pk.h:1629-1633)
int palz_renderer_clear(palz_renderer* r, void* stream) {
    using namespace palz;
    if (!r) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_renderer_clear: NULL renderer");
    fill_zero(r->d_accum, sizeof(palz_float4) * r->width * r->height, (hipStream_t)stream);
    PALZ_LAUNCH_CHECK();
    r->iteration = 0;
    return PALZ_OK;
}

// resolve (pk.h:1635-1650): pixels = pack(accumulators) into host memory
int palz_renderer_resolve(palz_renderer* r, uint32_t* rgba8, void* stream) {
    using namespace palz;
    if (!r || !rgba8) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_renderer_resolve: NULL argument");
    if (r->iteration == 0) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_renderer_resolve: no frame rendered");
    hipStream_t st = (hipStream_t)stream;
    const uint32_t np = r->width * r->height;
    k_pack<<<div_up(np, 256u), 256, 0, st>>>(reinterpret_cast<const float4*>(r->d_accum), r->d_pixels, np);
    PALZ_LAUNCH_CHECK();
    PALZ_HIP(hipMemcpyAsync(rgba8, r->d_pixels, sizeof(uint32_t) * np, hipMemcpyDeviceToHost, st));
    PALZ_HIP(hipStreamSynchronize(st));
    return PALZ_OK;
}

int palz_renderer_buffers(const palz_renderer* r, const palz_first_diffuse** d_points, const palz_float4** d_accum,
                          const uint32_t** d_counts) {
    if (!r) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_renderer_buffers: NULL renderer");
    if (d_points) *d_points = r->d_points;
    if (d_accum) *d_accum = r->d_accum;
    if (d_counts) *d_counts = r->d_scalars;
    return PALZ_OK;
}

}  // extern "C"
