// solvePrimary (main_gpu.cu:213-371): camera rays through the lens to their
// first diffuse hit, the receiver points of the caustic solve.
#include "spectral.hpp"

namespace palz {
namespace {

constexpr int kPrimaryThreads = 128;
constexpr int kPrimaryStack = 64;     // traversal stack entries per pixel
constexpr int kPrimaryStackLds = 16;  // of which in LDS; the rest in private memory
constexpr int kMaxBounces = 32;

// sphIntersect (pk.h:52-65)
__device__ __forceinline__ float sph_intersect_near(v3 ro, v3 rd, v3 ce, float ra) {
    const float len2rd = dot(rd, rd);
    const v3 oc = ro - ce;
    const float b = dot(oc, rd) / len2rd;
    const float c = dot(oc, oc) - ra * ra;
    float h = b * b - c / len2rd;
    if (h < 0.0f) return -1.0f;
    h = sqrtf(h);
    return -b - h;
}

// solveLens (pk.h:1042-1073); its result is not used by solvePrimary
__device__ void solve_lens(v3* ro, v3* rd, v3 cam_origin, v3 cam_forward, float distance, float thickness, float R,
                           float eta) {
    const v3 lens_origin = cam_origin + cam_forward * distance;
    const v3 center1 = lens_origin + cam_forward * (R - thickness * 0.5f);
    const v3 center2 = lens_origin - cam_forward * (R - thickness * 0.5f);
    const float t1 = sph_intersect_near(*ro, *rd, center1, R);
    if (t1 < 0.0f) return;
    const v3 hit1 = *ro + *rd * t1;
    v3 wo;
    refraction_norm_free(&wo, -*rd, hit1 - center1, eta);
    *rd = wo;
    *ro = hit1;
    const float t2 = sph_intersect_near(*ro, *rd, center2, R);
    const v3 hit2 = *ro + *rd * t2;
    if (!refraction_norm_free(&wo, -*rd, hit2 - center2, eta)) return;
    *rd = wo;
    *ro = hit2;
}

// Texture8RGBX::samplePoint (pk.h:1081-1094), powf(c / 255, 2.2) as
// exp(2.2 log(c / 255)) with this library's exp / log
__device__ v3 sample_point(const uint8_t* tex, uint32_t w, uint32_t h, float u, float v) {
    u = u - floorf(u);
    v = v - floorf(v);
    const int ix = (int)hmin(u * (float)w, (float)w - 1.0f);
    const int iy = (int)hmin(v * (float)h, (float)h - 1.0f);
    const uint32_t idx = (uint32_t)ix + w * (uint32_t)iy;
    float c[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float x = (float)tex[idx * 4u + k] / 255.0f;
        c[k] = x == 0.0f ? 0.0f : exp_f32(2.2f * log_f32(x));
    }
    return mk(c[0], c[1], c[2]);
}

__global__ void __launch_bounds__(kPrimaryThreads)
    k_solve_primary(palz_scene_view sc, palz_camera cam, palz_primary_params pp,
                    palz_first_diffuse* __restrict__ out, float4* __restrict__ accum) {
    __shared__ uint32_t s_stack[kPrimaryStackLds * kPrimaryThreads];
    SplitStack<kPrimaryStackLds, kPrimaryStack - kPrimaryStackLds> stack;
    stack.near = StridedStack{s_stack + threadIdx.x, kPrimaryThreads};
    const uint32_t pixel = blockIdx.x * kPrimaryThreads + threadIdx.x;
    if (pixel >= pp.width * pp.height) return;
    const uint32_t xi = pixel % pp.width, yi = pixel / pp.width;
    const uint32_t it = (uint32_t)pp.iteration;
    uint32_t dim = 0u;

    float jx, jy;
    sobol2d(&jx, &jy, it, xi, yi, dim++);
    // RayGenerator::shoot (camera.h:24-31), lerp(a, b, t) = a + t (b - a)
    const v3 origin = mk(cam.origin), right = mk(cam.right), up = mk(cam.up);
    const v3 forward = normalize(cross(up, right));
    const float u = ((float)xi + jx) / (float)pp.width, v = ((float)yi + jy) / (float)pp.height;
    const v3 to = origin + forward + (-right + (right - -right) * u) + (up + (-up - up) * v);
    v3 ro = origin, rd = to - origin;

    float lx, ly;
    sobol2d(&lx, &ly, it, xi, yi, dim++);
    const float lambda = cmf_y_sample(lx);
    const float p_lambda = cmf_y_pdf(lambda);
    const float eta = pp.cauchy_a + pp.cauchy_b / (lambda * lambda);  // CauchyDispersion (pk.h:935-949)
    if (pp.use_lens)
        solve_lens(&ro, &rd, origin, forward, pp.lens_distance, pp.lens_thickness, pp.lens_R, eta);

    const uint32_t root = *sc.root;
    bool diffuse_hit = false;
    Hit hit;
    for (int d = 0; d < kMaxBounces; ++d) {
        hit.t = kFltMax;
        hit.u = hit.v = 0.0f;
        hit.ng = mk(0.0f, 0.0f, 0.0f);
        hit.tri = 0xFFFFFFFFu;
        if (!closest_hit_on(&hit, sc.nodes, sc.tris, root, ro, rd, inv_rd(rd), stack, kPrimaryStack))
            break;  // the stack would overflow: no hit is reported
        if (hit.t == kFltMax) break;
        const palz_triangle_attrib at = sc.attribs[hit.tri];
        float rx, ry;
        sobol2d(&rx, &ry, it, xi, yi, dim++);
        if (at.material == PALZ_MATERIAL_DIFFUSE) {
            diffuse_hit = true;
            break;
        }
        const v3 p_hit = ro + rd * hit.t;
        const v3 s0 = mk(at.shading_normals[0]);
        const v3 ns = s0 + (mk(at.shading_normals[1]) - s0) * hit.u + (mk(at.shading_normals[2]) - s0) * hit.v;
        const v3 wi = -rd;
        bool e_T = false;
        if (at.material == PALZ_MATERIAL_DIELECTRIC) e_T = !(rx < fresnel_exact_norm_free(wi, ns, eta));
        float sign = 1.0f;
        v3 wo;
        if (!e_T) {
            wo = reflection(wi, ns);
        } else {
            if (!refraction_norm_free(&wo, wi, ns, eta)) break;
            sign = -1.0f;
        }
        if (0.0f < dot(hit.ng, wi) * dot(hit.ng, wo) * sign) {  // geometrically admissible
            const v3 offset_dir = normalize(hit.ng * dot(hit.ng, wo));
            ro = p_hit + offset_dir * ray_offset_scale(p_hit);
            rd = wo;
            continue;
        }
        break;
    }

    float4 acc = accum ? accum[pixel] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (!diffuse_hit) {
        out[pixel].R = {0.0f, 0.0f, 0.0f};
        if (accum) accum[pixel] = make_float4(acc.x, acc.y, acc.z, acc.w + 1.0f);
        return;
    }
    const v3 p = ro + rd * hit.t;
    v3 n = normalize(hit.ng);
    if (0.0f < dot(n, rd)) n = -n;
    const v3 p_light = mk(pp.p_light);
    const v3 to_light = p_light - p;
    const float d2 = dot(to_light, to_light);
    const v3 refl = pp.texture ? sample_point(pp.texture, pp.texture_width, pp.texture_height, p.x, p.z)
                               : mk(pp.albedo);
    v3 L = mk(0.0f, 0.0f, 0.0f);
    if (!occluded(sc.nodes, sc.tris, root, p, n, p_light, mk(0.0f, 0.0f, 0.0f))) {
        float contrib = pp.light_intensity / d2 * hmax(dot(normalize(to_light), n), 0.0f);
        contrib = hmin(contrib, pp.radiance_clamp);
        const v3 srgb = xyz_to_srgb_linear(
            mk(cmf_x(lambda) / kIntegralOfCmfY, cmf_y(lambda) / kIntegralOfCmfY, cmf_z(lambda) / kIntegralOfCmfY));
        // (reflectance / PI) * srgblinear * contrib / p_lambda, in that order
        const v3 brdf = refl / 3.14159265358979323846f;
        L = mk(brdf.x * srgb.x, brdf.y * srgb.y, brdf.z * srgb.z) * contrib / p_lambda;
    }
    palz_first_diffuse fd;
    fd.p = {p.x, p.y, p.z};
    fd.ng = {n.x, n.y, n.z};
    fd.R = {refl.x, refl.y, refl.z};
    fd.lambda = lambda;
    out[pixel] = fd;
    if (accum) accum[pixel] = make_float4(acc.x + L.x, acc.y + L.y, acc.z + L.z, acc.w + 1.0f);
}

}  // namespace
}  // namespace palz

extern "C" int palz_solve_primary(const palz_scene_view* scene, const palz_camera* camera,
                                  const palz_primary_params* pp, palz_first_diffuse* d_first_diffuse,
                                  palz_float4* d_accum, void* stream) {
    using namespace palz;
    if (!scene || !camera || !pp || !d_first_diffuse)
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_solve_primary: NULL argument");
    if (!scene->tris || !scene->attribs || !scene->nodes || !scene->root || scene->n_tris == 0u)
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_solve_primary: incomplete scene");
    if (pp->texture && (pp->texture_width == 0u || pp->texture_height == 0u))
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_solve_primary: empty texture");
    const uint64_t n = (uint64_t)pp->width * pp->height;
    if (n == 0u) return PALZ_OK;
    if (n > 0xFFFFFFFFull) return fail(PALZ_ERR_CAPACITY, "palz_solve_primary: image too large");
    k_solve_primary<<<div_up((uint32_t)n, (uint32_t)kPrimaryThreads), kPrimaryThreads, 0, (hipStream_t)stream>>>(
        *scene, *camera, *pp, d_first_diffuse, reinterpret_cast<float4*>(d_accum));
    PALZ_LAUNCH_CHECK();
    return PALZ_OK;
}
