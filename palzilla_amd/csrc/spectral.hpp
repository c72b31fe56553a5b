// CIE 2015 10-degree colour matching fit, its y-weighted wavelength sampling,
// XYZ -> linear sRGB and this library's FP32 exp / log (pk.h:960-1033).
#pragma once

#include "optics.hpp"

namespace palz {

// FP32 exp and log: the reference calls the hardware's fast forms
// (NATIVE_EXPF / NATIVE_LOGF); these are plain FP32 operations in a fixed
// order, so host and device give the same bits.
//   exp: x = k ln2 + r (ln 2 in two parts), e^r by its Taylor polynomial to
//        degree 7 (|r| <= ln2 / 2), times 2^k in two factors; 0 below -87
//        (no denormal results).
//   log: x = m 2^e with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s) with
//        s = (m - 1) / (m + 1), the odd series to s^9.
PALZ_HD float pow2i(int e) { return __builtin_bit_cast(float, (uint32_t)(e + 127) << 23); }  // e in [-126, 127]
PALZ_HD float exp_f32(float x) {
    if (!(x == x)) return x;
    if (x > 88.72283f) return __builtin_inff();
    if (x < -87.0f) return 0.0f;
    const float kf = floorf(x * 1.44269504f + 0.5f);
    const float r = (x - kf * 0.693145751953125f) - kf * 1.428606765330187e-06f;
    float p = 1.98412698e-4f;
    p = p * r + 1.38888889e-3f;
    p = p * r + 8.33333333e-3f;
    p = p * r + 4.16666667e-2f;
    p = p * r + 1.66666667e-1f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    const int k = (int)kf, k1 = k / 2;
    return p * pow2i(k1) * pow2i(k - k1);
}
PALZ_HD float log_f32(float x) {
    if (!(x > 0.0f)) return x == 0.0f ? -__builtin_inff() : __builtin_nanf("");
    if (x == __builtin_inff()) return x;
    uint32_t b = __builtin_bit_cast(uint32_t, x);
    int e = (int)(b >> 23) - 127;
    if ((b >> 23) == 0u) {  // denormal
        b = __builtin_bit_cast(uint32_t, x * 8388608.0f);
        e = (int)(b >> 23) - 127 - 23;
    }
    float m = __builtin_bit_cast(float, (b & 0x7FFFFFu) | 0x3F800000u);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e += 1;
    }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    float q = 1.11111111e-1f;
    q = q * s2 + 1.42857143e-1f;
    q = q * s2 + 2.0e-1f;
    q = q * s2 + 3.33333333e-1f;
    q = q * s2 + 1.0f;
    const float ef = (float)e;
    return ef * 0.693145751953125f + (2.0f * s * q + ef * 1.428606765330187e-06f);
}

constexpr float kIntegralOfCmfY = 118.51810464018001f;  // INTEGRAL_OF_CMF_Y_IN_NM (pk.h:1033)

PALZ_HD float asymmetric_gaussian(float x, float mean, float sigma, float a) {  // pk.h:962-968
    float denom = sigma + a * (x - mean);
    if (denom < 1.0e-15f) denom = 1.0e-15f;
    if (sigma * 2.0f < denom) denom = sigma * 2.0f;
    const float k = (x - mean) / denom;
    return exp_f32(-k * k);
}
PALZ_HD float cmf_x(float x) {  // pk.h:969-973
    const float a = 0.42f * asymmetric_gaussian(x, 445.5849609375f, 31.146467208862305f, 0.06435633450746536f);
    const float b = 1.16f * asymmetric_gaussian(x, 594.5570068359375f, 48.602108001708984f, -0.04772702232003212f);
    return (a + b - a * b * 42.559776306152344f) * 1.0008530432426956f;
}
PALZ_HD float cmf_y(float x) {  // pk.h:974-976
    return 1.0f * asymmetric_gaussian(x, 556.8383178710938f, 66.54190826416016f, -0.026492968201637268f) *
           1.004315086937574f;
}
PALZ_HD float cmf_z(float x) {  // pk.h:977-979
    return 2.146832f * asymmetric_gaussian(x, 445.9251708984375f, 30.91781997680664f, 0.08379141241312027f) *
           0.9975112815948937f;
}
// logistic_pdf (pk.h:981-985): its 1.0 literals make the denominator a double
PALZ_HD float logistic_pdf(float x, float s) {
    const float k = exp_f32(-fabsf(x) / s);
    return (float)(k / ((1.0 + k) * (1.0 + k) * s));
}
PALZ_HD float logistic_cdf(float x, float s) { return 1.0f / (1.0f + exp_f32(-x / s)); }  // pk.h:986-989
PALZ_HD float inverse_logistic_cdf(float u, float s) {                                  // pk.h:990-995
    if (0.99999994f < u) u = 0.99999994f;
    if (u < 1.175494351e-38f) u = 1.175494351e-38f;
    return -s * log_f32(1.0f / u - 1.0f);
}
PALZ_HD float cmf_y_pdf(float x) {  // pk.h:996-1002
    const float sx = x - 554.270751953125f;
    const float s = 26.879621505737305f, a = -164.270751953125f, b = 275.729248046875f;
    return logistic_pdf(sx, s) / (logistic_cdf(b, s) - logistic_cdf(a, s));
}
PALZ_HD float cmf_y_sample(float u) {  // pk.h:1004-1010
    const float s = 26.879621505737305f, a = -164.270751953125f, b = 275.729248046875f;
    const float pa = logistic_cdf(a, s), pb = logistic_cdf(b, s);
    return inverse_logistic_cdf(pa + (pb - pa) * u, s) + 554.270751953125f;
}
// xyz2srgblinear (pk.h:1014-1022): the dot products in the reference's order
PALZ_HD v3 xyz_to_srgb_linear(v3 c) {
    return mk(3.2406f * c.x + -1.5372f * c.y + -0.4986f * c.z, -0.9689f * c.x + 1.8758f * c.y + 0.0415f * c.z,
              0.0557f * c.x + -0.204f * c.y + 1.057f * c.z);
}

// The sRGB transfer function and the RGBA8 pack of the reference's `pack`
// kernel (main_gpu.cu:9-14, :499-513).
PALZ_HD float srgb_oetf(float r) {
    r = fmaxf(r, 0.0f);
    if (r <= 0.003130f) return 12.92f * r;
    return 1.055f * powf(r, 1.0f / 2.4f) - 0.055f;
}
PALZ_HD uint32_t pack_rgba(float r, float g, float b) {
    auto q = [](float c) -> uint32_t {
        const int i = (int)(c * 255.0f + 0.5f);  // make_int4: truncation toward zero
        return (uint32_t)(i < 0 ? 0 : i > 255 ? 255 : i);
    };
    return 0xFF000000u | (q(b) << 16) | (q(g) << 8) | q(r);
}

}  // namespace palz
