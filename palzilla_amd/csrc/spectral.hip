// Spectral resolve of the receiver irradiance: the CIE 2015 10-degree colour
// matching fit, XYZ -> linear sRGB, the sRGB transfer function and the RGBA8
// pack (pk.h:960-1031, main_gpu.cu:9-14, :473-483, :499-513).
//
// The reference weights each converged chain's contribution by
// (R / pi) * xyz2srgblinear(cmf(lambda) / INTEGRAL_OF_CMF_Y_IN_NM) / p(lambda)
// as it adds it to the pixel's float4 accumulator.  That weight depends on
// the wavelength and the receiver point only, so here it is applied once per
// (point, wavelength) to the solver's per-wavelength irradiance images:
// accumulator = (sum_l (R / pi) * srgb(l) * E_l[p] * w_l, n_frames), then
// pixel = packRGBA(srgb_oetf(acc.xyz / acc.w)).  w_l is the caller's sample
// weight of wavelength l (1 / p(lambda) for the reference's sampled
// wavelengths, a quadrature weight for a fixed set).
#include "spectral.hpp"

#include <algorithm>
#include <cmath>

namespace palz {
namespace {

constexpr int kMaxLambdas = 256;

// One thread per receiver point; the per-wavelength linear-sRGB weights are
// computed once per workgroup into LDS.
__global__ void __launch_bounds__(256)
    k_spectral_resolve(const float* __restrict__ images, const float* __restrict__ lambdas,
                       const float* __restrict__ weights, const palz_first_diffuse* __restrict__ points,
                       uint32_t n_points, uint32_t n_lambdas, float n_frames, float4* __restrict__ accum,
                       uint32_t* __restrict__ pixels) {
    __shared__ float s_w[kMaxLambdas][3];
    for (uint32_t l = threadIdx.x; l < n_lambdas; l += blockDim.x) {
        const float lam = lambdas[l];
        const float x = cmf_x(lam) / kIntegralOfCmfY, y = cmf_y(lam) / kIntegralOfCmfY,
                    z = cmf_z(lam) / kIntegralOfCmfY;
        const float w = weights[l];
        const v3 c = xyz_to_srgb_linear(mk(x, y, z));
        s_w[l][0] = c.x * w;
        s_w[l][1] = c.y * w;
        s_w[l][2] = c.z * w;
    }
    __syncthreads();
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_points) return;
    float r = 0.0f, g = 0.0f, b = 0.0f;
    for (uint32_t l = 0; l < n_lambdas; ++l) {
        const float e = images[(size_t)l * n_points + p];
        r += s_w[l][0] * e;
        g += s_w[l][1] * e;
        b += s_w[l][2] * e;
    }
    const palz_first_diffuse q = points[p];
    const float k = 1.0f / 3.14159265358979323846f;
    r *= q.R.x * k;
    g *= q.R.y * k;
    b *= q.R.z * k;
    if (accum) accum[p] = make_float4(r, g, b, n_frames);
    if (pixels) pixels[p] = pack_rgba(srgb_oetf(r / n_frames), srgb_oetf(g / n_frames), srgb_oetf(b / n_frames));
}

__global__ void __launch_bounds__(256) k_fixed_to_float(const uint64_t* __restrict__ fixed,
                                                        float* __restrict__ image, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        image[i] = (float)((double)fixed[i] * 0x1p-32);
}

}  // namespace
}  // namespace palz

extern "C" int palz_image_fixed_to_float(const uint64_t* d_fixed, float* d_image, size_t n, void* stream) {
    using namespace palz;
    if (n && (!d_fixed || !d_image)) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_image_fixed_to_float: NULL buffer");
    if (n == 0) return PALZ_OK;
    const uint32_t blocks = (uint32_t)std::min<size_t>((n + 255) / 256, 16384);
    k_fixed_to_float<<<blocks, 256, 0, (hipStream_t)stream>>>(d_fixed, d_image, n);
    PALZ_LAUNCH_CHECK();
    return PALZ_OK;
}

extern "C" int palz_spectral_resolve(const float* d_images, const float* d_lambdas, const float* d_weights,
                                     const palz_first_diffuse* d_points, uint32_t n_points, uint32_t n_lambdas,
                                     float n_frames, palz_float4* d_accum, uint32_t* d_pixels, void* stream) {
    using namespace palz;
    if (!d_images || !d_lambdas || !d_weights || !d_points)
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_spectral_resolve: NULL input");
    if (!d_accum && !d_pixels) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_spectral_resolve: no output");
    if (n_lambdas < 1 || n_lambdas > (uint32_t)kMaxLambdas)
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_spectral_resolve: need 1..256 wavelengths");
    if (!(n_frames > 0.0f)) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_spectral_resolve: n_frames must be > 0");
    if (n_points == 0u) return PALZ_OK;
    k_spectral_resolve<<<div_up(n_points, 256u), 256, 0, (hipStream_t)stream>>>(
        d_images, d_lambdas, d_weights, d_points, n_points, n_lambdas, n_frames,
        reinterpret_cast<float4*>(d_accum), d_pixels);
    PALZ_LAUNCH_CHECK();
    return PALZ_OK;
}
