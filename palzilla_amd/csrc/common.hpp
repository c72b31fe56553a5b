// Shared device/host helpers for the gfx950 caustic hot path.
//
// Arithmetic conventions (they make CPU and GPU results bit-identical):
//   * every translation unit is compiled with -ffp-contract=off, so a*b+c is
//     two correctly rounded operations on both sides, never an FMA unless a
//     kernel calls fmaf explicitly;
//   * FP32 division and sqrtf are correctly rounded (HIP's default);
//   * min/max of floats use ordered selects (pmin/pmax) or the canonical
//     signed-zero-aware forms (cmin/cmax) — never the hardware min/max whose
//     treatment of -0/+0 and NaN is left open.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "palz.h"

#define PALZ_HD __host__ __device__ __forceinline__

namespace palz {

static_assert(sizeof(palz_float3) == 12, "float3 layout");
static_assert(sizeof(palz_triangle) == 36, "Triangle layout (minimum_lbvh.h:207)");
static_assert(sizeof(palz_hit) == 28, "Hit fields (minimum_lbvh.h:905-911)");
static_assert(sizeof(palz_aabb) == 24, "AABB layout (minimum_lbvh.h:220)");
static_assert(sizeof(palz_internal_node) == 64, "InternalNode layout (minimum_lbvh.h:285)");
static_assert(offsetof(palz_internal_node, aabbs) == 8, "InternalNode layout");
static_assert(offsetof(palz_internal_node, parent) == 56, "InternalNode layout");
static_assert(offsetof(palz_internal_node, context) == 60, "InternalNode layout");
static_assert(sizeof(palz_triangle_attrib) == 40, "TriangleAttrib layout (pk.h:74)");
static_assert(sizeof(palz_wide_node) == 248, "wide node layout");
static_assert(sizeof(palz_normal_bound) == 52, "InternalNormalBound layout (pk.h:1705)");

constexpr float kFltMax = 3.402823466e+38F;

// ---------------------------------------------------------------------------
// 3-vectors
struct v3 {
    float x, y, z;
};
PALZ_HD v3 mk(float x, float y, float z) { return {x, y, z}; }
PALZ_HD v3 mk(const palz_float3& p) { return {p.x, p.y, p.z}; }
PALZ_HD v3 operator+(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
PALZ_HD v3 operator-(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
PALZ_HD v3 operator-(v3 a) { return {-a.x, -a.y, -a.z}; }
PALZ_HD v3 operator*(v3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
PALZ_HD v3 operator*(v3 a, v3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
PALZ_HD v3 operator/(v3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
PALZ_HD v3 operator/(v3 a, v3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
PALZ_HD float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PALZ_HD v3 cross(v3 a, v3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Ordered selects, the host semantics of the reference's fminf/fmaxf
// (helper_math.h:134-142): a<b?a:b.
PALZ_HD float pmin(float a, float b) { return a < b ? a : b; }
PALZ_HD float pmax(float a, float b) { return a < b ? b : a; }

// Canonical min/max for bounding boxes: -0 orders below +0, so the result does
// not depend on the order in which values are folded in (NaN-free inputs).
PALZ_HD float cmin(float a, float b) {
    return (a < b || (a == b && __builtin_signbit(a))) ? a : b;
}
PALZ_HD float cmax(float a, float b) {
    return (b < a || (a == b && !__builtin_signbit(a))) ? a : b;
}
PALZ_HD v3 cmin(v3 a, v3 b) { return {cmin(a.x, b.x), cmin(a.y, b.y), cmin(a.z, b.z)}; }
PALZ_HD v3 cmax(v3 a, v3 b) { return {cmax(a.x, b.x), cmax(a.y, b.y), cmax(a.z, b.z)}; }

// ---------------------------------------------------------------------------
// Morton encoding (minimum_lbvh.h:173-195, :249-254)
PALZ_HD uint64_t splat3_21(uint32_t a) {
    uint64_t x = a & 0x1fffff;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}
PALZ_HD uint64_t morton63(uint32_t x, uint32_t y, uint32_t z) {
    return splat3_21(x) | splat3_21(y) << 1 | splat3_21(z) << 2;
}

// Quantise a normalised coordinate f = (c - lo) / (hi - lo) * 2^bits to
// [0, 2^bits - 1].  Equals clamp(int(f), 0, max) of minimum_lbvh.h:251-252 for
// every finite f; NaN maps to 0 and +inf to max (the reference leaves both to
// the float->int conversion of the platform).
PALZ_HD uint32_t quantise(float f, float fmaxv, uint32_t maxv) {
    if (!(f > 0.0f)) return 0u;
    if (f >= fmaxv) return maxv;
    return (uint32_t)(int)f;
}

// Bit-exact 32-bit key of a centroid.  bits == 32: top 32 bits of the 63-bit
// code (minimum_lbvh.cu:89).  bits == 30: 10 bits per axis interleaved.
PALZ_HD uint32_t morton_key(v3 p, v3 lo, v3 hi, int bits) {
    v3 f = (p - lo) / (hi - lo);
    if (bits == 30) {
        uint32_t x = quantise(f.x * 1024.0f, 1023.0f, 1023u);
        uint32_t y = quantise(f.y * 1024.0f, 1023.0f, 1023u);
        uint32_t z = quantise(f.z * 1024.0f, 1023.0f, 1023u);
        return (uint32_t)morton63(x, y, z);
    }
    uint32_t x = quantise(f.x * 2097152.0f, 2097151.0f, 2097151u);
    uint32_t y = quantise(f.y * 2097152.0f, 2097151.0f, 2097151u);
    uint32_t z = quantise(f.z * 2097152.0f, 2097151.0f, 2097151u);
    return (uint32_t)(morton63(x, y, z) >> 31);
}

PALZ_HD int clz32(uint32_t x) { return x == 0 ? 32 : __builtin_clz(x); }
// minimum_lbvh.h:140-143
PALZ_HD uint32_t delta32(uint32_t a, uint32_t b) { return 32u - (uint32_t)clz32(a ^ b); }

PALZ_HD uint32_t div_up(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

// ---------------------------------------------------------------------------
// host-side error state
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what, const char* file, int line);

}  // namespace palz

#define PALZ_HIP(call)                                                          \
    do {                                                                        \
        hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess) return ::palz::hip_fail(e_, #call, __FILE__, __LINE__); \
    } while (0)

#define PALZ_LAUNCH_CHECK()                                                     \
    do {                                                                        \
        hipError_t e_ = hipGetLastError();                                      \
        if (e_ != hipSuccess) return ::palz::hip_fail(e_, "kernel launch", __FILE__, __LINE__); \
    } while (0)
