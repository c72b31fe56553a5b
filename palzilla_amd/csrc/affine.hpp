// Affine arithmetic with rigorous rounding, and the per-tuple admissibility
// test of path cuts in affine form.
//
// The interval test (interval.hpp) encloses each chain vertex and normal in a
// box on its own, so wi = x[k-1] - x[k] and wo = x[k+1] - x[k] lose the fact
// that they share x[k], and a product of two quantities that move together is
// bounded as if they moved apart.  Here every quantity is an affine form
//     x = c + sum_i a_i e_i + r [-1, 1],   e_i in [-1, 1],
// over a fixed set of noise symbols: the two barycentric parameters of each
// chain triangle (over the rectangle being tested) and the refractive index
// over the solved wavelengths.  Vertices and interpolated shading normals are
// affine in the parameters, so the chain's residual keeps its correlations up
// to the nonlinear steps, whose remainders go into r.
//
// Rounding: every coefficient is computed round-to-nearest and its exact
// error recovered by an error-free transformation (TwoSum for sums, an FMA
// residual for products); those errors are added into r with upward
// rounding, so each form encloses the exact real-number result of the same
// operations on its inputs.  Nonlinear steps (square root, reciprocal) use a
// linear approximation over the argument's range with rigorous (directed)
// bounds on the remainder.
#pragma once

#include "interval.hpp"

namespace palz {
namespace aa {

using iv::add_rd;
using iv::add_ru;
using iv::div_rd;
using iv::div_ru;
using iv::mul_rd;
using iv::mul_ru;
using iv::sqrt_rd;
using iv::sqrt_ru;

constexpr int kSym = 7;  // u, v of up to three chain triangles, and eta

struct A {
    float c;
    float a[kSym];
    float r;  // >= 0
};
struct A3 {
    A x, y, z;
};

PALZ_HD A pt(float v) {
    A z;
    z.c = v;
#pragma unroll
    for (int i = 0; i < kSym; ++i) z.a[i] = 0.0f;
    z.r = 0.0f;
    return z;
}

// |exact - rounded| of a + b, a * b
PALZ_HD float err_add(float a, float b, float s) { return fabsf(iv::two_sum_err(a, b, s)); }
PALZ_HD float err_mul(float a, float b, float p) {
    if (!(p - p == 0.0f)) return iv::kInf;
    if (fabsf(p) < iv::kTiny) return iv::kTiny;  // the residual may underflow there
    return fabsf(fmaf(a, b, -p));
}

// total radius of the noise part: sum |a_i| + r, rounded up
PALZ_HD float rad(const A& x) {
    float s = x.r;
#pragma unroll
    for (int i = 0; i < kSym; ++i) s = add_ru(s, fabsf(x.a[i]));
    return s;
}
PALZ_HD iv::I range(const A& x) {
    const float rr = rad(x);
    return {add_rd(x.c, -rr), add_ru(x.c, rr)};
}

PALZ_HD A operator+(const A& x, const A& y) {
    A z;
    float e = add_ru(x.r, y.r);
    z.c = x.c + y.c;
    e = add_ru(e, err_add(x.c, y.c, z.c));
#pragma unroll
    for (int i = 0; i < kSym; ++i) {
        z.a[i] = x.a[i] + y.a[i];
        e = add_ru(e, err_add(x.a[i], y.a[i], z.a[i]));
    }
    z.r = e;
    return z;
}
PALZ_HD A operator-(const A& x) {
    A z;
    z.c = -x.c;
#pragma unroll
    for (int i = 0; i < kSym; ++i) z.a[i] = -x.a[i];
    z.r = x.r;
    return z;
}
PALZ_HD A operator-(const A& x, const A& y) { return x + (-y); }

PALZ_HD A operator*(const A& x, const A& y) {
    A z;
    z.c = x.c * y.c;
    float e = err_mul(x.c, y.c, z.c);
#pragma unroll
    for (int i = 0; i < kSym; ++i) {
        const float p1 = x.c * y.a[i], p2 = y.c * x.a[i];
        z.a[i] = p1 + p2;
        e = add_ru(e, add_ru(add_ru(err_mul(x.c, y.a[i], p1), err_mul(y.c, x.a[i], p2)), err_add(p1, p2, z.a[i])));
    }
    // |x.c| y.r + |y.c| x.r + rad(x) rad(y): the cross terms with r and the
    // product of the two noise parts
    e = add_ru(e, mul_ru(fabsf(x.c), y.r));
    e = add_ru(e, mul_ru(fabsf(y.c), x.r));
    e = add_ru(e, mul_ru(rad(x), rad(y)));
    z.r = e;
    return z;
}

// x * x: the square of the noise part lies in [0, rad^2], so its midpoint
// moves into the centre and half its width into r
PALZ_HD A sqr(const A& x) {
    A z;
    const float rr = rad(x);
    const float q = mul_ru(rr, rr);
    const float cc = x.c * x.c;
    const float h = q * 0.5f;  // exact (a power-of-two scale), rounding of q already up
    z.c = cc + h;
    float e = add_ru(err_mul(x.c, x.c, cc), err_add(cc, h, z.c));
    e = add_ru(e, h);
#pragma unroll
    for (int i = 0; i < kSym; ++i) {
        const float p = x.c * x.a[i];
        z.a[i] = p + p;  // exact doubling
        e = add_ru(e, 2.0f * err_mul(x.c, x.a[i], p));
    }
    // 2 |x.c| r: the cross term of the centre with r
    e = add_ru(e, mul_ru(2.0f * fabsf(x.c), x.r));
    z.r = e;
    return z;
}

// alpha x + zeta + delta [-1, 1] with alpha, zeta chosen and delta >= 0
PALZ_HD A lin(const A& x, float alpha, float zeta_lo, float zeta_hi) {
    A z;
    // the intercept interval [zeta_lo, zeta_hi]: its midpoint as the centre
    // shift, half its width into r
    const float zm = 0.5f * zeta_lo + 0.5f * zeta_hi;
    const float zr = pmax(add_ru(zeta_hi, -zm), add_ru(zm, -zeta_lo));
    const float p = alpha * x.c;
    z.c = p + zm;
    float e = add_ru(add_ru(err_mul(alpha, x.c, p), err_add(p, zm, z.c)), zr);
#pragma unroll
    for (int i = 0; i < kSym; ++i) {
        z.a[i] = alpha * x.a[i];
        e = add_ru(e, err_mul(alpha, x.a[i], z.a[i]));
    }
    e = add_ru(e, mul_ru(fabsf(alpha), x.r));
    z.r = e;
    return z;
}

// sqrt over the range [lo, hi], lo >= 0.  For any slope alpha > 0, g(t) =
// sqrt(t) - alpha t is concave: its minimum over [lo, hi] is at an end, its
// maximum at t* = 1 / (4 alpha^2) (value 1 / (4 alpha)) or at the nearer end.
PALZ_HD A sqrt_af(const A& x) {
    const iv::I rg = range(x);
    const float lo = pmax(rg.l, 0.0f), hi = pmax(rg.u, 0.0f);
    if (hi == 0.0f) return pt(0.0f);
    const float slo_d = sqrt_rd(lo), shi_d = sqrt_rd(hi), slo_u = sqrt_ru(lo), shi_u = sqrt_ru(hi);
    if (!(hi > lo)) {  // a point: [sqrt_rd, sqrt_ru]
        return lin(x, 0.0f, slo_d, shi_u);
    }
    // the secant slope, rounded to nearest: any positive value is rigorous
    float alpha = (sqrtf(hi) - sqrtf(lo)) / (hi - lo);
    if (!(alpha > 0.0f) || !(alpha < iv::kInf)) alpha = 0.5f / sqrtf(hi);
    const float g_lo = add_rd(slo_d, -mul_ru(alpha, lo));
    const float g_hi = add_rd(shi_d, -mul_ru(alpha, hi));
    const float m = pmin(g_lo, g_hi);
    // t* = 1 / (4 alpha^2): the maximum 1 / (4 alpha) if t* lies in [lo, hi]
    const float t_lo = div_rd(1.0f, mul_ru(4.0f, mul_ru(alpha, alpha)));
    const float t_hi = div_ru(1.0f, mul_rd(4.0f, mul_rd(alpha, alpha)));
    float M;
    if (t_hi >= lo && t_lo <= hi) {
        M = div_ru(1.0f, mul_rd(4.0f, alpha));
    } else {
        M = pmax(add_ru(slo_u, -mul_rd(alpha, lo)), add_ru(shi_u, -mul_rd(alpha, hi)));
    }
    return lin(x, alpha, m, M);
}

// 1 / x over [lo, hi] with lo > 0: for a slope alpha < 0, g(t) = 1/t - alpha t
// is convex: its maximum is at an end, its minimum at t* = 1/sqrt(-alpha)
// (value 2 sqrt(-alpha)) or at the nearer end.  ok = false when the range
// reaches zero.
PALZ_HD A recip_af(const A& x, bool* ok) {
    const iv::I rg = range(x);
    if (!(rg.l > 0.0f)) {
        *ok = false;
        return pt(0.0f);
    }
    const float lo = rg.l, hi = rg.u;
    if (!(hi > lo)) return lin(x, 0.0f, div_rd(1.0f, hi), div_ru(1.0f, lo));
    float alpha = -1.0f / (lo * hi);
    if (!(alpha < 0.0f) || !(alpha > -iv::kInf)) alpha = -1.0f / (hi * hi);
    const float M = pmax(add_ru(div_ru(1.0f, lo), mul_ru(-alpha, lo)), add_ru(div_ru(1.0f, hi), mul_ru(-alpha, hi)));
    const float na_lo = -alpha, na_hi = -alpha;
    const float ts_lo = div_rd(1.0f, sqrt_ru(na_hi)), ts_hi = div_ru(1.0f, sqrt_rd(na_lo));
    float m;
    if (ts_hi >= lo && ts_lo <= hi) {
        m = mul_rd(2.0f, sqrt_rd(na_lo));
    } else {
        m = pmin(add_rd(div_rd(1.0f, lo), mul_rd(-alpha, lo)), add_rd(div_rd(1.0f, hi), mul_rd(-alpha, hi)));
    }
    return lin(x, alpha, m, M);
}

PALZ_HD A3 operator+(const A3& a, const A3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
PALZ_HD A3 operator-(const A3& a, const A3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
PALZ_HD A3 operator-(const A3& a) { return {-a.x, -a.y, -a.z}; }
PALZ_HD A3 operator*(const A3& a, const A& s) { return {a.x * s, a.y * s, a.z * s}; }
PALZ_HD A dot(const A3& a, const A3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PALZ_HD A length_squared(const A3& a) { return sqr(a.x) + sqr(a.y) + sqr(a.z); }
PALZ_HD A3 cross(const A3& a, const A3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
PALZ_HD A3 pt3(v3 p) { return {pt(p.x), pt(p.y), pt(p.z)}; }
PALZ_HD float mag(const A& a) {
    const iv::I r = range(a);
    return pmax(fabsf(r.l), fabsf(r.u));
}
PALZ_HD float mag3(const A3& a) { return pmax(pmax(mag(a.x), mag(a.y)), mag(a.z)); }
PALZ_HD bool near_zero(const A3& v, float tau) {
    const iv::I x = range(v.x), y = range(v.y), z = range(v.z);
    return !(x.l > tau || x.u < -tau || y.l > tau || y.u < -tau || z.l > tau || z.u < -tau);
}

// p0 + (p1 - p0) u + (p2 - p0) v for u = um + ur e_su, v = vm + vr e_sv over a
// parameter rectangle (um, ur, vm, vr multiples of 1/16: exact); the edge
// differences are rounded, so their errors times |u| <= 1, |v| <= 1 go into r.
PALZ_HD A affine_point1(float p0, float p1, float p2, float um, float ur, float vm, float vr, int su, int sv) {
    const float e0 = p1 - p0, e1 = p2 - p0;
    float e = add_ru(err_add(p1, -p0, e0), err_add(p2, -p0, e1));
    const float t0 = e0 * um, t1 = e1 * vm;
    const float s = t0 + t1;
    A z = pt(0.0f);
    z.c = p0 + s;
    e = add_ru(e, add_ru(add_ru(err_mul(e0, um, t0), err_mul(e1, vm, t1)), add_ru(err_add(t0, t1, s), err_add(p0, s, z.c))));
    z.a[su] = e0 * ur;
    z.a[sv] = e1 * vr;
    e = add_ru(e, add_ru(err_mul(e0, ur, z.a[su]), err_mul(e1, vr, z.a[sv])));
    z.r = e;
    return z;
}
PALZ_HD A3 affine_point(v3 p0, v3 p1, v3 p2, float um, float ur, float vm, float vr, int su, int sv) {
    return {affine_point1(p0.x, p1.x, p2.x, um, ur, vm, vr, su, sv),
            affine_point1(p0.y, p1.y, p2.y, um, ur, vm, vr, su, sv),
            affine_point1(p0.z, p1.z, p2.z, um, ur, vm, vr, su, sv)};
}

// The residual cross(wo, wopt) of one oriented chain vertex (pk.h:505-548) in
// affine form, with the magnitude product |wo| |wopt| the tolerance scales
// with.  Returns false when it cannot be enclosed: total internal reflection
// possible (refraction), or |n|^2 reaching zero (reflection).
PALZ_HD bool residual_af(const A3& wi, const A3& wo, const A3& n, bool refract, const A& eta, A3* c, float* scale) {
    const A non = length_squared(n);
    const A win = dot(wi, n);
    A3 wopt;
    if (!refract) {
        bool ok = true;
        const A inv = recip_af(non, &ok);
        if (!ok) return false;
        wopt = n * (win * pt(2.0f) * inv) - wi;
    } else {
        const A wiwi = length_squared(wi);
        const A k = non * wiwi * (sqr(eta) - pt(1.0f)) + sqr(win);
        if (!(range(k).l >= 0.0f)) return false;
        wopt = -(wi * non) + n * (win - sqrt_af(k));
    }
    *c = cross(wo, wopt);
    *scale = mag3(wo) * mag3(wopt);
    return true;
}

// The residual of one oriented chain vertex can reach the solver's tolerance
// (the criteria of iv::vertex_admissible_oriented, in affine form).  Where it
// cannot be enclosed the vertex is kept: with total internal reflection the
// solver's residual is zero.  Kept out of line: inlined at each of the 2K
// call sites of every chain length, the straight-line affine code takes the
// device compiler tens of minutes.
static __host__ __device__ __attribute__((noinline)) bool vertex_admissible_oriented(const A3& wi, const A3& wo, const A3& n, bool refract, const A& eta) {
    A3 c;
    float scale;
    if (!residual_af(wi, wo, n, refract, eta, &c, &scale)) return true;
    const float tau = iv::kResidualTol + iv::kEvalTol * pmax(1.0f, scale);
    return near_zero(c, tau);
}
PALZ_HD bool vertex_admissible(const A3& wi, const A3& wo, const A3& n, bool refract, const A& eta) {
    const iv::I s = range(dot(wi, n));
    if (!(s.u < 0.0f) && vertex_admissible_oriented(wi, wo, n, refract, eta)) return true;
    if (!(s.l >= 0.0f) && vertex_admissible_oriented(wo, wi, n, refract, eta)) return true;
    return false;
}

// Per-tuple admissibility in affine form over the same parameter rectangles
// as iv::chain_admissible (a rectangle's part beyond u + v <= 1 is included:
// a superset).  eta spans [eta.l, eta.u] on its own noise symbol.
template <int K>
PALZ_HD bool chain_admissible(v3 p_beg, v3 p_end, const ChainTri* tri, uint32_t event, iv::I eta_iv) {
    int split[K];
    iv::chain_splits<K>(tri, split);
    int combos = 1;
#pragma unroll
    for (int k = 0; k < K; ++k) combos *= split[k] * split[k];
    A eta = pt(0.5f * eta_iv.l + 0.5f * eta_iv.u);
    {
        const float half = pmax(add_ru(eta_iv.u, -eta.c), add_ru(eta.c, -eta_iv.l));
        eta.a[kSym - 1] = half;
    }
    for (int c = 0; c < combos; ++c) {
        A3 vert[K + 2];
        A3 nrm[K];
        vert[0] = pt3(p_beg);
        vert[K + 1] = pt3(p_end);
        int rest = c;
        bool inside = true;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int sk = split[k];
            const int iu = rest % sk, iv_ = (rest / sk) % sk;
            rest /= sk * sk;
            const float inv = 1.0f / (float)sk;
            const float u0 = iu * inv, v0 = iv_ * inv;
            if (u0 + v0 > 1.0f) inside = false;
            const float hr = 0.5f * inv;  // half the rectangle's side
            vert[k + 1] = affine_point(tri[k].v0, tri[k].v1, tri[k].v2, u0 + hr, hr, v0 + hr, hr, 2 * k, 2 * k + 1);
            nrm[k] = affine_point(tri[k].s0, tri[k].s1, tri[k].s2, u0 + hr, hr, v0 + hr, hr, 2 * k, 2 * k + 1);
        }
        if (!inside) continue;
        bool ok = true;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const A3 wi = vert[k] - vert[k + 1];
            const A3 wo = vert[k + 2] - vert[k + 1];
            if (!vertex_admissible(wi, wo, nrm[k], ev_is_T(event, k), eta)) {
                ok = false;
                break;
            }
        }
        if (ok) return true;
    }
    return false;
}

}  // namespace aa
}  // namespace palz
