// Interval arithmetic with directed rounding, and the per-tuple admissibility
// test of path cuts (interval.h:34-713, pk.h:2101-2207) as a conservative
// filter in front of the Newton solve.
//
// Directed rounding without changing the FP rounding mode: every operation is
// computed round-to-nearest and its exact error is recovered with an
// error-free transformation (Knuth's TwoSum for + and -, an FMA residual for
// *, / and sqrt); the bound then steps one ulp outward when the error points
// that way.  The results are the correctly rounded-down / rounded-up values
// (to one ulp, always on the safe side) and are bit-reproducible wherever + -
// * / sqrt and fma are IEEE operations, which makes the CPU restatement in the
// oracle comparable bit for bit.
#pragma once

#include "optics.hpp"

namespace palz {
namespace iv {

constexpr float kInf = __builtin_huge_valf();
constexpr float kTiny = 0x1p-100f;  // below this an FMA residual may underflow

PALZ_HD float next_up(float x) {
    if (!(x == x) || x == kInf) return x;
    if (x == 0.0f) return 0x1p-126f;  // FLT_MIN: safe with or without denormals
    const int32_t b = __builtin_bit_cast(int32_t, x);
    return __builtin_bit_cast(float, x > 0.0f ? b + 1 : b - 1);
}
PALZ_HD float next_down(float x) { return -next_up(-x); }

// exact error of a + b: (a + b) - fl(a + b)
PALZ_HD float two_sum_err(float a, float b, float s) {
    const float bp = s - a;
    const float ap = s - bp;
    return (a - ap) + (b - bp);
}

PALZ_HD float add_rd(float a, float b) {
    const float s = a + b;
    if (s == kInf) return a == kInf || b == kInf ? kInf : 3.40282347e+38f;
    if (!(s - s == 0.0f)) return s;
    return two_sum_err(a, b, s) < 0.0f ? next_down(s) : s;
}
PALZ_HD float add_ru(float a, float b) {
    const float s = a + b;
    if (s == -kInf) return a == -kInf || b == -kInf ? -kInf : -3.40282347e+38f;
    if (!(s - s == 0.0f)) return s;
    return two_sum_err(a, b, s) > 0.0f ? next_up(s) : s;
}
PALZ_HD float mul_rd(float a, float b) {
    if (a == 0.0f || b == 0.0f) return 0.0f;
    const float p = a * b;
    if (!(p - p == 0.0f)) return p == kInf ? 3.40282347e+38f : p;
    if (fabsf(p) < kTiny) return p - kTiny;
    return fmaf(a, b, -p) < 0.0f ? next_down(p) : p;
}
PALZ_HD float mul_ru(float a, float b) {
    if (a == 0.0f || b == 0.0f) return 0.0f;
    const float p = a * b;
    if (!(p - p == 0.0f)) return p == -kInf ? -3.40282347e+38f : p;
    if (fabsf(p) < kTiny) return p + kTiny;
    return fmaf(a, b, -p) > 0.0f ? next_up(p) : p;
}
// b != 0
PALZ_HD float div_rd(float a, float b) {
    if (a == 0.0f) return 0.0f;
    const float q = a / b;
    if (!(q - q == 0.0f)) return q == kInf ? 3.40282347e+38f : q;
    if (fabsf(q) < kTiny) return q - kTiny;
    const float r = fmaf(-q, b, a);  // a - q b, exact
    return (b > 0.0f ? r < 0.0f : r > 0.0f) ? next_down(q) : q;
}
PALZ_HD float div_ru(float a, float b) {
    if (a == 0.0f) return 0.0f;
    const float q = a / b;
    if (!(q - q == 0.0f)) return q == -kInf ? -3.40282347e+38f : q;
    if (fabsf(q) < kTiny) return q + kTiny;
    const float r = fmaf(-q, b, a);
    return (b > 0.0f ? r > 0.0f : r < 0.0f) ? next_up(q) : q;
}
// a >= 0
PALZ_HD float sqrt_rd(float a) {
    if (a == 0.0f || a == kInf) return a;
    const float s = sqrtf(a);
    if (s < kTiny) return 0.0f;
    return fmaf(-s, s, a) < 0.0f ? next_down(s) : s;
}
PALZ_HD float sqrt_ru(float a) {
    if (a == 0.0f || a == kInf) return a;
    const float s = sqrtf(a);
    if (s < kTiny) return kTiny;
    return fmaf(-s, s, a) > 0.0f ? next_up(s) : s;
}

// ---------------------------------------------------------------------------
// intervals [l, u]
struct I {
    float l, u;
};
struct I3 {
    I x, y, z;
};

PALZ_HD I pt(float v) { return {v, v}; }
PALZ_HD I operator-(I a) { return {-a.u, -a.l}; }
PALZ_HD I operator+(I a, I b) { return {add_rd(a.l, b.l), add_ru(a.u, b.u)}; }
PALZ_HD I operator-(I a, I b) { return {add_rd(a.l, -b.u), add_ru(a.u, -b.l)}; }
PALZ_HD I operator*(I a, I b) {
    const float l = fminf(fminf(mul_rd(a.l, b.l), mul_rd(a.l, b.u)), fminf(mul_rd(a.u, b.l), mul_rd(a.u, b.u)));
    const float u = fmaxf(fmaxf(mul_ru(a.l, b.l), mul_ru(a.l, b.u)), fmaxf(mul_ru(a.u, b.l), mul_ru(a.u, b.u)));
    return {l, u};
}
// a divisor that contains zero gives the whole line (interval.h:74-80 has no such case)
PALZ_HD I operator/(I a, I b) {
    if (b.l <= 0.0f && 0.0f <= b.u) return {-kInf, kInf};
    const float l = fminf(fminf(div_rd(a.l, b.l), div_rd(a.l, b.u)), fminf(div_rd(a.u, b.l), div_rd(a.u, b.u)));
    const float u = fmaxf(fmaxf(div_ru(a.l, b.l), div_ru(a.l, b.u)), fmaxf(div_ru(a.u, b.l), div_ru(a.u, b.u)));
    return {l, u};
}
// interval.h:97-113
PALZ_HD I square(I x) {
    const float hi = fmaxf(mul_ru(x.l, x.l), mul_ru(x.u, x.u));
    if (x.l <= 0.0f && 0.0f <= x.u) return {0.0f, hi};
    return {fminf(mul_rd(x.l, x.l), mul_rd(x.u, x.u)), hi};
}
// of the non-negative part (interval.h:90-96 takes sqrtf of the raw bounds)
PALZ_HD I sqrt(I x) { return {sqrt_rd(fmaxf(x.l, 0.0f)), sqrt_ru(fmaxf(x.u, 0.0f))}; }

PALZ_HD I3 box(v3 lo, v3 hi) { return {{lo.x, hi.x}, {lo.y, hi.y}, {lo.z, hi.z}}; }
PALZ_HD I3 pt3(v3 p) { return {pt(p.x), pt(p.y), pt(p.z)}; }
PALZ_HD I3 operator-(I3 a) { return {-a.x, -a.y, -a.z}; }
PALZ_HD I3 operator+(I3 a, I3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
PALZ_HD I3 operator-(I3 a, I3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
PALZ_HD I3 operator*(I3 a, I s) { return {a.x * s, a.y * s, a.z * s}; }
PALZ_HD I dot(I3 a, I3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PALZ_HD I length_squared(I3 v) { return square(v.x) + square(v.y) + square(v.z); }
PALZ_HD I3 cross(I3 a, I3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
PALZ_HD float mag(I a) { return fmaxf(fabsf(a.l), fabsf(a.u)); }
PALZ_HD float mag3(I3 a) { return fmaxf(fmaxf(mag(a.x), mag(a.y)), mag(a.z)); }
// every component meets [-tau, tau] (NaN bounds count as meeting it)
PALZ_HD bool near_zero(I3 v, float tau) {
    return !(v.x.l > tau || v.x.u < -tau || v.y.l > tau || v.y.u < -tau || v.z.l > tau || v.z.u < -tau);
}

// ---------------------------------------------------------------------------
// The Newton residual of one chain vertex (pk.h:505-548) can reach the
// solver's tolerance somewhere in the boxes, for one orientation of (wi, wo).
// kResidualTol: cost < 1e-10 bounds each residual component by 1e-5; the FP32
// evaluation of the residual at a point differs from its exact value by a few
// ulps of the magnitudes involved (kEvalTol of the product of the magnitudes).
constexpr float kResidualTol = 1.0e-5f;
constexpr float kEvalTol = 1.0e-5f;

PALZ_HD bool vertex_admissible_oriented(I3 wi, I3 wo, I3 n, bool refract, I eta) {
    const I non = length_squared(n);
    const I win = dot(wi, n);
    I3 wopt;
    if (!refract) {
        // reflection(wi, n) = n * (2 dot(wi, n) / dot(n, n)) - wi (saka.h:192-195)
        wopt = n * ((win * pt(2.0f)) / non) - wi;
    } else {
        // refraction_norm_free(wi, n, eta) (saka.h:197-208); where k < 0 it is
        // the zero vector and the residual vanishes, so such a box is kept
        const I wiwi = length_squared(wi);
        const I k = non * wiwi * (square(eta) - pt(1.0f)) + square(win);
        if (!(k.l >= 0.0f)) return true;
        wopt = -wi * non + n * (win - sqrt(k));
    }
    const I3 c = cross(wo, wopt);
    const float tau = kResidualTol + kEvalTol * fmaxf(1.0f, mag3(wo) * mag3(wopt));
    return near_zero(c, tau);
}

// solveConstraints swaps wi and wo where dot(wi, n) < 0 (pk.h:517-520): both
// orientations the boxes allow are tried.
PALZ_HD bool vertex_admissible(I3 wi, I3 wo, I3 n, bool refract, I eta) {
    const I s = dot(wi, n);
    if (!(s.u < 0.0f) && vertex_admissible_oriented(wi, wo, n, refract, eta)) return true;
    if (!(s.l >= 0.0f) && vertex_admissible_oriented(wo, wi, n, refract, eta)) return true;
    return false;
}

// Image of the part of the parameter rectangle [u0,u1] x [v0,v1] inside the
// triangle u, v >= 0, u + v <= 1 under the affine map p0 + (p1 - p0) u +
// (p2 - p0) v (a vertex or an unnormalised shading normal): the hull of the
// clipped polygon's corners, each evaluated in intervals.  The bounds are
// multiples of 1/8, so 1 - u and 1 - v are exact.
PALZ_HD I3 affine_box(v3 p0, v3 p1, v3 p2, float u0, float u1, float v0, float v1) {
    const I3 a = pt3(p0), e0 = pt3(p1) - a, e1 = pt3(p2) - a;
    const float cu[8] = {u0, u1, u0, u1, 1.0f - v0, 1.0f - v1, u0, u1};
    const float cv[8] = {v0, v0, v1, v1, v0, v1, 1.0f - u0, 1.0f - u1};
    I3 r = {{kInf, -kInf}, {kInf, -kInf}, {kInf, -kInf}};
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        // corners inside the triangle; edge crossings of u + v = 1 on the rectangle
        const bool on_rect = u0 <= cu[c] && cu[c] <= u1 && v0 <= cv[c] && cv[c] <= v1;
        if (!(on_rect && cu[c] + cv[c] <= 1.0f)) continue;
        const I3 q = a + e0 * pt(cu[c]) + e1 * pt(cv[c]);
        r.x = {fminf(r.x.l, q.x.l), fmaxf(r.x.u, q.x.u)};
        r.y = {fminf(r.y.l, q.y.l), fmaxf(r.y.u, q.y.u)};
        r.z = {fminf(r.z.l, q.z.l), fmaxf(r.z.u, q.z.u)};
    }
    return r;
}

constexpr int kMaxSplit = 8;     // parameter rectangles per triangle axis
constexpr int kSplitBudget = 64;  // rectangles per chain

// Per-tuple admissibility (pk.h:2101-2207 at leaf level): vertex boxes over
// the triangles, normal boxes over their shading normals, the end points as
// points, eta spanning the wavelengths solved.  A triangle much larger than
// the others of the chain is split into S x S parameter rectangles (S a power
// of two, chosen from the box extents within a budget); the chain is kept if
// any combination of rectangles passes every vertex.  False only if no point
// of the triangles can bring every vertex's residual within the solver's
// tolerance.
// The split of each chain triangle's parameter square (a power of two per
// axis, at most kMaxSplit, within kSplitBudget rectangles in all).
template <int K>
PALZ_HD void chain_splits(const ChainTri* tri, int* split) {
    float ext[K], ext_min = kInf;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const v3 lo = cmin(cmin(tri[k].v0, tri[k].v1), tri[k].v2), hi = cmax(cmax(tri[k].v0, tri[k].v1), tri[k].v2);
        ext[k] = fmaxf(fmaxf(hi.x - lo.x, hi.y - lo.y), hi.z - lo.z);
        ext_min = fminf(ext_min, ext[k]);
    }
    int combos = 1;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        int sk = 1;
        while (sk < kMaxSplit && ext[k] > 2.0f * sk * ext_min) sk *= 2;
        split[k] = sk;
        combos *= sk * sk;
    }
    while (combos > kSplitBudget) {  // halve the largest split until within budget
        int kk = 0;
#pragma unroll
        for (int k = 1; k < K; ++k)
            if (split[k] > split[kk]) kk = k;
        combos /= 4;
        split[kk] /= 2;
    }
}

template <int K>
PALZ_HD bool chain_admissible(v3 p_beg, v3 p_end, const ChainTri* tri, uint32_t event, I eta) {
    int split[K];
    chain_splits<K>(tri, split);
    int combos = 1;
#pragma unroll
    for (int k = 0; k < K; ++k) combos *= split[k] * split[k];
    for (int c = 0; c < combos; ++c) {
        I3 vert[K + 2];
        I3 nrm[K];
        vert[0] = pt3(p_beg);
        vert[K + 1] = pt3(p_end);
        int rest = c;
        bool inside = true;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int sk = split[k];
            const int iu = rest % sk, iv = (rest / sk) % sk;
            rest /= sk * sk;
            const float inv = 1.0f / (float)sk;  // a power of two: exact
            const float u0 = iu * inv, u1 = (iu + 1) * inv, v0 = iv * inv, v1 = (iv + 1) * inv;
            if (u0 + v0 > 1.0f) inside = false;  // rectangle beyond the triangle's u + v <= 1
            vert[k + 1] = affine_box(tri[k].v0, tri[k].v1, tri[k].v2, u0, u1, v0, v1);
            nrm[k] = affine_box(tri[k].s0, tri[k].s1, tri[k].s2, u0, u1, v0, v1);
        }
        if (!inside) continue;
        bool ok = true;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const I3 wi = vert[k] - vert[k + 1];
            const I3 wo = vert[k + 2] - vert[k + 1];
            if (!vertex_admissible(wi, wo, nrm[k], ev_is_T(event, k), eta)) {
                ok = false;
                break;
            }
        }
        if (ok) return true;
    }
    return false;
}

}  // namespace iv
}  // namespace palz
