// Device-side optics of the caustic path: forward-mode duals, Householder QR,
// ray/triangle and BVH traversal, Sobol photons, Fresnel, refraction, the
// specular-chain Newton solver, path throughput and the footprint dA/dw.
//
// The FP32 arithmetic reproduces the reference's operation order literally so
// that, with -ffp-contract=off and correctly rounded / and sqrt, every result
// is bit-identical to the CPU restatement:
//   dual numbers ........ saka.h:11-208
//   QR least squares .... sen.h:347-356, :586-747
//   solver .............. pk.h:466-592
//   throughput .......... pk.h:760-842, occlusion pk.h:88-115
//   dA/dw ............... pk.h:844-930
//   traversal ........... minimum_lbvh.h:82-96, :372-422, :927-988, :1078-1152
//   Sobol ............... sobol.h:17-129
// Host min/max semantics (helper_math.h:134-142) are used everywhere but in
// the slab test's folds (fminf/fmaxf, whose box decisions are the same: see
// slabs; tests/slab_check.cpp checks them on the CPU).
#pragma once

#include "common.hpp"

namespace palz {

// ---------------------------------------------------------------------------
// small float helpers (helper_math host forms)
PALZ_HD float hmin(float a, float b) { return a < b ? a : b; }
PALZ_HD float hmax(float a, float b) { return a > b ? a : b; }
PALZ_HD double hmin(double a, double b) { return a < b ? a : b; }
PALZ_HD double hmax(double a, double b) { return a > b ? a : b; }
PALZ_HD float length(v3 v) { return sqrtf(dot(v, v)); }
PALZ_HD v3 normalize(v3 v) { return v * (1.0f / sqrtf(dot(v, v))); }  // rsqrtf = 1/sqrtf
PALZ_HD v3 vmin(v3 a, v3 b) { return {hmin(a.x, b.x), hmin(a.y, b.y), hmin(a.z, b.z)}; }
PALZ_HD v3 vmax(v3 a, v3 b) { return {hmax(a.x, b.x), hmax(a.y, b.y), hmax(a.z, b.z)}; }

// double 3-vectors for the FP64 refinement
struct d3 {
    double x, y, z;
};
PALZ_HD d3 mkd(v3 a) { return {(double)a.x, (double)a.y, (double)a.z}; }
PALZ_HD d3 operator+(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
PALZ_HD d3 operator-(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }

PALZ_HD d3 operator*(d3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }

template <class F>
struct vec3_of;
template <>
struct vec3_of<float> {
    using type = v3;
};
template <>
struct vec3_of<double> {
    using type = d3;
};
template <class F>
PALZ_HD typename vec3_of<F>::type to_w(v3 a);
template <>
PALZ_HD v3 to_w<float>(v3 a) { return a; }
template <>
PALZ_HD d3 to_w<double>(v3 a) { return mkd(a); }

// ---------------------------------------------------------------------------
// Forward-mode dual with NT tangents.  Each tangent is computed with exactly
// the expression saka uses for its single tangent (details::binary, saka.h:
// 38-45), including the terms whose tangent is a constant zero, so NT == 1
// reproduces one saka pass and NT == 2K reproduces 2K passes at once.
template <int NT, class F>
struct dual {
    F v;
    F g[NT];
};

template <int NT, class F>
PALZ_HD dual<NT, F> dc(F x) {
    dual<NT, F> r;
    r.v = x;
#pragma unroll
    for (int t = 0; t < NT; ++t) r.g[t] = F(0);
    return r;
}
template <int NT, class F>
PALZ_HD dual<NT, F> operator+(const dual<NT, F>& a, const dual<NT, F>& b) {
    dual<NT, F> r;
    r.v = a.v + b.v;
#pragma unroll
    for (int t = 0; t < NT; ++t) r.g[t] = a.g[t] + b.g[t];  // g*1 + g*1
    return r;
}
template <int NT, class F>
PALZ_HD dual<NT, F> operator-(const dual<NT, F>& a) {
    dual<NT, F> r;
    r.v = -a.v;
#pragma unroll
    for (int t = 0; t < NT; ++t) r.g[t] = -a.g[t];  // g * -1
    return r;
}
template <int NT, class F>
PALZ_HD dual<NT, F> operator-(const dual<NT, F>& a, const dual<NT, F>& b) {
    dual<NT, F> r;
    r.v = a.v - b.v;
#pragma unroll
    for (int t = 0; t < NT; ++t) r.g[t] = a.g[t] - b.g[t];  // g*1 + g*(-1)
    return r;
}
template <int NT, class F>
PALZ_HD dual<NT, F> operator*(const dual<NT, F>& a, const dual<NT, F>& b) {
    dual<NT, F> r;
    r.v = a.v * b.v;
#pragma unroll
    for (int t = 0; t < NT; ++t) r.g[t] = a.g[t] * b.v + b.g[t] * a.v;
    return r;
}
template <int NT, class F>
PALZ_HD dual<NT, F> operator/(const dual<NT, F>& a, const dual<NT, F>& b) {
    dual<NT, F> r;
    r.v = a.v / b.v;
    const F dx = F(1) / b.v;
    const F dy = -a.v / (b.v * b.v);
#pragma unroll
    for (int t = 0; t < NT; ++t) r.g[t] = a.g[t] * dx + b.g[t] * dy;
    return r;
}
template <int NT, class F>
PALZ_HD dual<NT, F> dsqrt(const dual<NT, F>& a) {
    dual<NT, F> r;
    r.v = sqrt(a.v);
    const F d = F(0.5) / sqrt(a.v);
#pragma unroll
    for (int t = 0; t < NT; ++t) r.g[t] = a.g[t] * d;
    return r;
}

template <int NT, class F>
struct dual3 {
    dual<NT, F> x, y, z;
};
template <int NT, class F, class V>
PALZ_HD dual3<NT, F> dc3(V p) {
    return {dc<NT, F>((F)p.x), dc<NT, F>((F)p.y), dc<NT, F>((F)p.z)};
}
template <int NT, class F>
PALZ_HD dual3<NT, F> operator+(const dual3<NT, F>& a, const dual3<NT, F>& b) {
    return {a.x + b.x, a.y + b.y, a.z + b.z};
}
template <int NT, class F>
PALZ_HD dual3<NT, F> operator-(const dual3<NT, F>& a, const dual3<NT, F>& b) {
    return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <int NT, class F>
PALZ_HD dual3<NT, F> operator-(const dual3<NT, F>& a) {
    return {-a.x, -a.y, -a.z};
}
template <int NT, class F>
PALZ_HD dual3<NT, F> operator*(const dual3<NT, F>& a, const dual<NT, F>& s) {
    return {a.x * s, a.y * s, a.z * s};
}
template <int NT, class F>
PALZ_HD dual3<NT, F> operator/(const dual3<NT, F>& a, const dual<NT, F>& s) {
    return {a.x / s, a.y / s, a.z / s};
}
template <int NT, class F>
PALZ_HD dual<NT, F> ddot(const dual3<NT, F>& a, const dual3<NT, F>& b) {
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <int NT, class F>
PALZ_HD dual3<NT, F> dcross(const dual3<NT, F>& a, const dual3<NT, F>& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// saka::reflection (saka.h:186-189): n * dot(wi, n) * 2 / dot(n, n) - wi
template <int NT, class F>
PALZ_HD dual3<NT, F> dreflection(const dual3<NT, F>& wi, const dual3<NT, F>& n) {
    return n * ddot(wi, n) * dc<NT, F>(F(2)) / ddot(n, n) - wi;
}
// saka::refraction_norm_free (saka.h:191-202); zero vector on TIR
template <int NT, class F>
PALZ_HD dual3<NT, F> drefraction(const dual3<NT, F>& wi, const dual3<NT, F>& n, F eta) {
    const dual<NT, F> NoN = ddot(n, n);
    const dual<NT, F> WIoN = ddot(wi, n);
    const dual<NT, F> WoW = ddot(wi, wi);
    const dual<NT, F> k = NoN * WoW * dc<NT, F>(eta * eta - F(1)) + WIoN * WIoN;
    if (k.v < F(0)) return {dc<NT, F>(F(0)), dc<NT, F>(F(0)), dc<NT, F>(F(0))};
    return -wi * NoN + n * (WIoN - dsqrt(k));
}

// ---------------------------------------------------------------------------
// Householder QR least squares of a rows x cols system (sen.h:658-747).
// A is column-major: A[col * rows + row].  The reference keeps every
// Householder vector and applies them to b after the factorisation
// (applyQTransposed, sen.h:618-647); here each is applied to b as soon as it
// is formed, so that only one vector is live.  b's updates read only that
// vector and b, and run in the same order, so the arithmetic is the same.
template <int R, int C, class F>
PALZ_HD void solve_qr(F* A, F* b, F* x) {
    constexpr int H = (C < R - 1) ? C : R - 1;
#pragma unroll
    for (int i = 0; i < H; ++i) {
        F v[R];
#pragma unroll
        for (int j = 0; j < R; ++j) v[j] = j < i ? F(0) : -A[i * R + j];
        const F sgn = F(0) < v[i] ? F(1) : F(-1);
        F cd = F(0);
#pragma unroll
        for (int j = 0; j < R; ++j) cd += v[j] * v[j];
        const F x_len = sqrt(cd);
        v[i] += sgn * x_len;
        const F v_len2 = F(2) * x_len * (x_len + fabs(A[i * R + i]));
        A[i * R + i] = sgn * x_len;
#pragma unroll
        for (int r = i + 1; r < R; ++r) A[i * R + r] = F(0);
#pragma unroll
        for (int c = i + 1; c < C; ++c) {
            F vdc = F(0);
#pragma unroll
            for (int r = i; r < R; ++r) vdc += v[r] * A[c * R + r];
            const F k = F(2) * vdc / v_len2;
#pragma unroll
            for (int r = i; r < R; ++r) A[c * R + r] -= k * v[r];
        }
        // applyQTransposed's step i
        F vdc = F(0), vdv = F(0);
#pragma unroll
        for (int r = i; r < R; ++r) {
            vdc += v[r] * b[r];
            vdv += v[r] * v[r];
        }
        const F k = F(2) * vdc / vdv;
#pragma unroll
        for (int r = i; r < R; ++r) b[r] -= k * v[r];
    }
    // back substitution (sen.h:736-744)
#pragma unroll
    for (int r = C - 1; r >= 0; --r) {
        F sum = F(0);
#pragma unroll
        for (int i = r + 1; i < C; ++i) sum += A[i * R + r] * x[i];
        x[r] = (b[r] - sum) / A[r * R + r];
    }
}

// ---------------------------------------------------------------------------
// The 32.32 fixed-point word of a contribution: c * 2^32 truncated.  The
// conversion is defined for 0 <= c * 2^32 < 2^64 only, so a NaN or
// non-positive c gives 0 and c >= 2^32 saturates at the largest float below
// 2^64 (with the radiance clamp checked at the API neither occurs).
PALZ_HD unsigned long long fixed_word(float c) {
    if (!(c > 0.0f)) return 0ull;
    const float s = c < 4294967296.0f ? c * 4294967296.0f : 18446742974197923840.0f;
    return (unsigned long long)s;
}

// One contribution into a scene's irradiance image: float atomics (the
// reference's, main_gpu.cu:473-483) or, when the scene names a fixed-point
// image, a 64-bit integer atomic of its fixed-point word (order-independent).
__device__ __forceinline__ void add_contribution(const palz_scene_desc& sc, size_t idx, float c) {
    if (sc.image_fixed) {
        const unsigned long long w = fixed_word(c);
        if (w) atomicAdd(reinterpret_cast<unsigned long long*>(sc.image_fixed) + idx, w);
    } else {
        atomicAdd(sc.image + idx, c);
    }
}

// EventDescriptor (pk.h:122-162): bits [0,24) events (1 = T), [24,32) size.
PALZ_HD int ev_size(uint32_t e) { return (int)(e >> 24); }
PALZ_HD bool ev_is_T(uint32_t e, int k) { return (e >> k) & 1u; }

// Geometry of one chain vertex: the raw triangle and shading normals, from
// which the reference forms its edges per use (v1 - v0, v2 - v0, ...).
struct ChainTri {
    v3 v0, v1, v2;  // Triangle::vs
    v3 s0, s1, s2;  // TriangleAttrib::shadingNormals
    int material;
};

PALZ_HD ChainTri load_chain_tri(const palz_triangle* tris, const palz_triangle_attrib* attribs,
                                                   int idx) {
    const palz_triangle t = tris[idx];
    const palz_triangle_attrib a = attribs[idx];
    ChainTri c;
    c.v0 = mk(t.vs[0]);
    c.v1 = mk(t.vs[1]);
    c.v2 = mk(t.vs[2]);
    c.s0 = mk(a.shading_normals[0]);
    c.s1 = mk(a.shading_normals[1]);
    c.s2 = mk(a.shading_normals[2]);
    c.material = a.material;
    return c;
}

// One residual evaluation of solveConstraints (pk.h:484-548) for the NT
// parameters [first, first + NT): fills those columns of A, all of b and the
// cost (when first == 0).
template <int K, int NT, class F, class V>
PALZ_HD void chain_residual(const F* params, int first, V p_beg, V p_end, const ChainTri* tri, F eta,
                            uint32_t event, F* A, F* b, F* cost) {
    constexpr int R = 3 * K;
    dual3<NT, F> vert[K + 2];
    dual3<NT, F> nrm[K];
    vert[0] = dc3<NT, F>(p_beg);
    vert[K + 1] = dc3<NT, F>(p_end);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        dual<NT, F> pu = dc<NT, F>(params[2 * k]);
        dual<NT, F> pv = dc<NT, F>(params[2 * k + 1]);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            pu.g[t] = (2 * k + 0 == first + t) ? F(1) : F(0);
            pv.g[t] = (2 * k + 1 == first + t) ? F(1) : F(0);
        }
        // e0 = vs[1] - vs[0], e1 = vs[2] - vs[0] (pk.h:501-502, :506-507) in the
        // working precision: float ops for F = float, exact for F = double
        using W = typename vec3_of<F>::type;
        const W v0 = to_w<F>(tri[k].v0), s0 = to_w<F>(tri[k].s0);
        const W e0 = to_w<F>(tri[k].v1) - v0, e1 = to_w<F>(tri[k].v2) - v0;
        const W ne0 = to_w<F>(tri[k].s1) - s0, ne1 = to_w<F>(tri[k].s2) - s0;
        vert[k + 1] = dc3<NT, F>(v0) + dc3<NT, F>(e0) * pu + dc3<NT, F>(e1) * pv;
        nrm[k] = dc3<NT, F>(s0) + dc3<NT, F>(ne0) * pu + dc3<NT, F>(ne1) * pv;
    }
    F c_acc = F(0);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        dual3<NT, F> wi = vert[k] - vert[k + 1];
        dual3<NT, F> wo = vert[k + 2] - vert[k + 1];
        const dual3<NT, F>& n = nrm[k];
        if (ddot(wi, n).v < F(0)) {
            const dual3<NT, F> tmp = wi;
            wi = wo;
            wo = tmp;
        }
        const dual3<NT, F> wopt = ev_is_T(event, k) ? drefraction(wi, n, eta) : dreflection(wi, n);
        const dual3<NT, F> c = dcross(wo, wopt);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            A[(first + t) * R + 3 * k + 0] = c.x.g[t];
            A[(first + t) * R + 3 * k + 1] = c.y.g[t];
            A[(first + t) * R + 3 * k + 2] = c.z.g[t];
        }
        b[3 * k + 0] = c.x.v;
        b[3 * k + 1] = c.y.v;
        b[3 * k + 2] = c.z.v;
        c_acc += ddot(c, c).v;
    }
    if (first == 0) *cost = c_acc;
}

// solveConstraints<K> (pk.h:466-592).  NT tangents per residual pass: NT == 1
// is the reference's 2K passes, NT == 2K computes the Jacobian in one pass;
// both give bit-identical A, b and cost.  Returns 1 on convergence, 0 on
// rejection or iteration exhaustion; *iters is the iteration index at exit.
template <int K, int NT, class F, class V>
PALZ_HD bool solve_constraints(F* params, V p_beg, V p_end, const ChainTri* tri, F eta, uint32_t event,
                               int max_iter, F tol, int warmup, int* iters) {
    constexpr int P = 2 * K;
    constexpr int R = 3 * K;
    F edge_eps = F(0.25);
    for (int iter = 0; iter < max_iter; ++iter) {
        F A[R * P];
        F b[R];
        F cost = F(0);
#pragma unroll
        for (int first = 0; first < P; first += NT)
            chain_residual<K, NT, F, V>(params, first, p_beg, p_end, tri, eta, event, A, b, &cost);
        if (cost < tol) {
            *iters = iter;
            return true;
        }
        F dp[P];
        solve_qr<R, P, F>(A, b, dp);
#pragma unroll
        for (int i = 0; i < P; ++i) params[i] = params[i] - dp[i];
        if (warmup <= iter) {
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const F u = params[2 * k], v = params[2 * k + 1];
                if (!(-edge_eps < u && -edge_eps < v && u + v < F(1) + edge_eps)) {
                    *iters = iter;
                    return false;
                }
            }
            edge_eps *= edge_eps;
        }
    }
    *iters = max_iter;
    return false;
}

// ---------------------------------------------------------------------------
// Fresnel / refraction / reflection on float3 (pk.h:181-205, :427-450)
PALZ_HD float fresnel_exact_norm_free(v3 wi, v3 n, float eta) {
    const float nn = dot(n, n);
    const float wiwi = dot(wi, wi);
    float C = dot(n, wi);
    if (C < 0.0f) {
        C = -C;
        n = -n;
        eta = 1.0f / eta;
    }
    const float K = nn * wiwi * (eta * eta - 1.0f) + C * C;
    if (K < 0.0f) return 1.0f;
    const float G = sqrtf(K);
    const float GmC = G - C;
    const float GpC = G + C;
    const float a = GmC / GpC;
    const float q = (C * GpC - nn * wiwi) / (C * GmC + nn * wiwi);
    return 0.5f * (a * a) * (1.0f + q * q);
}
PALZ_HD bool refraction_norm_free(v3* wo, v3 wi, v3 n, float eta) {
    const float NoN = dot(n, n);
    float WIoN = dot(wi, n);
    if (WIoN < 0.0f) {
        WIoN = -WIoN;
        n = -n;
        eta = 1.0f / eta;
    }
    const float WIoWI = dot(wi, wi);
    const float k = NoN * WIoWI * (eta * eta - 1.0f) + WIoN * WIoN;
    if (k < 0.0f) return false;
    *wo = -wi * NoN + n * (WIoN - sqrtf(k));
    return true;
}
PALZ_HD v3 reflection(v3 wi, v3 n) { return n * dot(wi, n) * 2.0f / dot(n, n) - wi; }

PALZ_HD float ray_offset_scale(v3 p) {  // rayOffsetScale (pk.h:43-49)
    const v3 a = mk(fabsf(p.x), fabsf(p.y), fabsf(p.z));
    const float m = hmax(hmax(hmax(a.x, a.y), a.z), 1.0f);
    return m * 1.192092896e-07F * 128.0f;
}

// ---------------------------------------------------------------------------
// ray / box / triangle (minimum_lbvh.h:82-96, :372-422, :913-916)
PALZ_HD v3 inv_rd(v3 rd) {
    const v3 q = mk(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
    // clamp(v, -FLT_MAX, FLT_MAX) = fmaxf(-FLT_MAX, fminf(v, FLT_MAX))
    return mk(hmax(-kFltMax, hmin(q.x, kFltMax)), hmax(-kFltMax, hmin(q.y, kFltMax)),
              hmax(-kFltMax, hmin(q.z, kFltMax)));
}
// The slab test (minimum_lbvh.h:82-96).  The reference folds with a<b?a:b;
// here the folds are fminf/fmaxf, one v_min/v_max instruction each instead
// of a compare and a select (the traversals issue VALU instructions at the
// device's ceiling).  The two differ only for NaN operands and in the sign
// of a zero result.  No operand here is NaN: ro, lo and hi are finite and
// ird is clamped to +-FLT_MAX, so every t is finite or an infinity.  A zero's
// sign changes no later comparison: the max/min folds, the scaling by
// 1.00000024f and the caller's a <= b treat -0 and +0 alike.  So every box
// decision, and the order of children by entry distance, is the reference's.
PALZ_HD float fold_min(float a, float b) { return fminf(a, b); }
PALZ_HD float fold_max(float a, float b) { return fmaxf(a, b); }
PALZ_HD void slabs(v3 ro, v3 ird, v3 lo, v3 hi, float known_t, float* rmin, float* rmax) {
    const v3 t0 = (lo - ro) * ird;
    const v3 t1 = (hi - ro) * ird;
    const v3 tmin = {fold_min(t0.x, t1.x), fold_min(t0.y, t1.y), fold_min(t0.z, t1.z)};
    const v3 tmax = {fold_max(t0.x, t1.x), fold_max(t0.y, t1.y), fold_max(t0.z, t1.z)};
    float a = fold_max(fold_max(tmin.x, tmin.y), tmin.z);
    float b = fold_min(fold_min(tmax.x, tmax.y), tmax.z) * 1.00000024f;
    a = fold_max(a, 0.0f);
    b = fold_min(b, known_t);
    *rmin = a;
    *rmax = b;
}
PALZ_HD bool intersect_tri(float* t_out, float* u_out, float* v_out, v3* ng_out, float t_min, float t_max, v3 ro,
                           v3 rd, v3 v0, v3 v1, v3 v2) {
    const v3 e0 = v1 - v0;
    const v3 e1 = v2 - v1;
    const v3 ng = cross(e0, e1);
    const float t = dot(v0 - ro, ng) / dot(ng, rd);
    if (t_min <= t && t <= t_max) {
        const v3 e2 = v0 - v2;
        const float u_vol = dot(cross(e2, v0 + v2 - ro * 2.0f), rd);
        const float v_vol = dot(cross(e0, v1 + v0 - ro * 2.0f), rd);
        const float w_vol = dot(cross(e1, v2 + v1 - ro * 2.0f), rd);
        if (hmin(hmin(u_vol, v_vol), w_vol) < 0.0f && 0.0f < hmax(hmax(u_vol, v_vol), w_vol)) return false;
        const float vol = u_vol + v_vol + w_vol;
        *t_out = t;
        *u_out = u_vol / vol;
        *v_out = v_vol / vol;
        *ng_out = ng;
        return true;
    }
    return false;
}

struct Hit {
    float t;
    float u, v;
    v3 ng;
    uint32_t tri;
};

PALZ_HD palz_triangle ld_tri(const palz_triangle* tris, uint32_t i) { return tris[i]; }

// intersect_escape_link (minimum_lbvh.h:1078-1152) in any-hit mode, used by
// occluded() (pk.h:88-115).  A leaf root (one-triangle scene) is tested
// directly; the reference would index nodes[invalid] there.
__host__ __device__ inline bool any_hit_escape(const palz_internal_node* nodes, const palz_triangle* tris, uint32_t root, v3 ro,
                                      v3 rd, v3 ird, float t_max) {
    float hit_t = t_max;
    if (root & PALZ_NODE_LEAF_BIT) {
        const palz_triangle tr = ld_tri(tris, root & ~PALZ_NODE_LEAF_BIT);
        float t, u, v;
        v3 ng;
        return intersect_tri(&t, &u, &v, &ng, 0.0f, hit_t, ro, rd, mk(tr.vs[0]), mk(tr.vs[1]), mk(tr.vs[2]));
    }
    uint32_t curr = root, prev = PALZ_NODE_INVALID;
    while (curr != PALZ_NODE_INVALID) {
        if (curr & PALZ_NODE_LEAF_BIT) {
            const palz_triangle tr = ld_tri(tris, curr & ~PALZ_NODE_LEAF_BIT);
            float t, u, v;
            v3 ng;
            if (intersect_tri(&t, &u, &v, &ng, 0.0f, hit_t, ro, rd, mk(tr.vs[0]), mk(tr.vs[1]), mk(tr.vs[2])))
                return true;
            if (nodes[prev].children[0] == curr) {
                const uint32_t s = curr;
                curr = prev;
                prev = s;
                continue;
            }
            const uint32_t idx = nodes[prev].context;
            prev = curr;
            curr = idx != 0xFFFFFFFFu ? idx : PALZ_NODE_INVALID;
            continue;
        }
        const palz_internal_node& nd = nodes[curr];
        const bool descent = prev == nd.parent;
        palz_aabb box;
        uint32_t hit_child, miss;
        if (descent) {
            box = nd.aabbs[0];
            hit_child = nd.children[0];
            miss = curr;
        } else {
            box = nd.aabbs[1];
            hit_child = nd.children[1];
            miss = nd.context != 0xFFFFFFFFu ? nd.context : PALZ_NODE_INVALID;
        }
        float a, b;
        slabs(ro, ird, mk(box.lower), mk(box.upper), hit_t, &a, &b);
        prev = curr;
        curr = a <= b ? hit_child : miss;
    }
    return false;
}

// intersect_stackfull (minimum_lbvh.h:927-988): near-first closest hit.  The
// stack lives in LDS as stack[depth * stride + lane] (one column per lane, so
// pushes and pops of a wave hit distinct banks).  Returns false if the stack
// would overflow `depth` entries; the hit is then incomplete.
// The traversal stack of closest_hit: one lane's entries interleaved with its
// block's other lanes (entry i at stack[i * stride]).
struct StridedStack {
    uint32_t* base;
    int stride;
    PALZ_HD void put(int i, uint32_t v) { base[i * stride] = v; }
    PALZ_HD uint32_t get(int i) const { return base[i * stride]; }
};
// A stack whose first kNear entries are a StridedStack (in LDS) and the rest a
// private array: deep entries are rare, so the LDS per lane (and with it the
// block's LDS, which bounds how many waves a CU holds) stays small.
template <int kNear, int kFar>
struct SplitStack {
    StridedStack near;
    uint32_t far[kFar];
    PALZ_HD void put(int i, uint32_t v) {
        if (i < kNear)
            near.put(i, v);
        else
            far[i - kNear] = v;
    }
    PALZ_HD uint32_t get(int i) const { return i < kNear ? near.get(i) : far[i - kNear]; }
};

template <class Stack>
__host__ __device__ inline bool closest_hit_on(Hit* hit, const palz_internal_node* nodes, const palz_triangle* tris,
                                               uint32_t root, v3 ro, v3 rd, v3 ird, Stack& stack, int depth);

__host__ __device__ inline bool closest_hit(Hit* hit, const palz_internal_node* nodes, const palz_triangle* tris,
                                   uint32_t root, v3 ro, v3 rd, v3 ird, uint32_t* stack, int stride,
                                   int depth) {
    StridedStack st{stack, stride};
    return closest_hit_on(hit, nodes, tris, root, ro, rd, ird, st, depth);
}

// intersect_stackfull (minimum_lbvh.h:927-988) on any stack type: near child
// first, the far one pushed; false when `depth` entries would not suffice.
template <class Stack>
__host__ __device__ inline bool closest_hit_on(Hit* hit, const palz_internal_node* nodes, const palz_triangle* tris,
                                               uint32_t root, v3 ro, v3 rd, v3 ird, Stack& stack, int depth) {
    int sp = 0;
    uint32_t node = root;
    while (node != PALZ_NODE_INVALID) {
        if (node & PALZ_NODE_LEAF_BIT) {
            const uint32_t ti = node & ~PALZ_NODE_LEAF_BIT;
            const palz_triangle tr = ld_tri(tris, ti);
            float t, u, v;
            v3 ng;
            if (intersect_tri(&t, &u, &v, &ng, 0.0f, hit->t, ro, rd, mk(tr.vs[0]), mk(tr.vs[1]), mk(tr.vs[2]))) {
                hit->t = t;
                hit->u = u;
                hit->v = v;
                hit->ng = ng;
                hit->tri = ti;
            }
            node = sp == 0 ? PALZ_NODE_INVALID : stack.get(--sp);
            continue;
        }
        const palz_internal_node& nd = nodes[node];
        float l0, l1, r0, r1;
        slabs(ro, ird, mk(nd.aabbs[0].lower), mk(nd.aabbs[0].upper), hit->t, &l0, &l1);
        slabs(ro, ird, mk(nd.aabbs[1].lower), mk(nd.aabbs[1].upper), hit->t, &r0, &r1);
        const bool hit_l = l0 <= l1;
        const bool hit_r = r0 <= r1;
        if (hit_l && hit_r) {
            uint32_t near_node = nd.children[0], far_node = nd.children[1];
            if (r0 < l0) {
                near_node = nd.children[1];
                far_node = nd.children[0];
            }
            if (sp >= depth) return false;
            node = near_node;
            stack.put(sp++, far_node);
        } else if (hit_l || hit_r) {
            node = nd.children[hit_l ? 0 : 1];
        } else {
            node = sp == 0 ? PALZ_NODE_INVALID : stack.get(--sp);
        }
    }
    return true;
}

// intersect_stackfree (minimum_lbvh.h:996-1076): the parent-pointer walk with
// no stack.  At each internal node both child boxes are tested against the
// current hit distance and ordered near-first; the node it came from decides
// where to go next (down to the near child, across to the far one, or up to
// the parent).  `any`: stop at the first triangle hit (RAY_QUERY_ANY).  A
// leaf root (one-triangle scene) is tested directly.
__host__ __device__ inline void closest_hit_stackfree(Hit* hit, const palz_internal_node* nodes,
                                                      const palz_triangle* tris, uint32_t root, v3 ro, v3 rd,
                                                      v3 ird, bool any) {
    uint32_t curr = root, prev = PALZ_NODE_INVALID;
    while (curr != PALZ_NODE_INVALID) {
        if (curr & PALZ_NODE_LEAF_BIT) {
            const uint32_t ti = curr & ~PALZ_NODE_LEAF_BIT;
            const palz_triangle tr = ld_tri(tris, ti);
            float t, u, v;
            v3 ng;
            if (intersect_tri(&t, &u, &v, &ng, 0.0f, hit->t, ro, rd, mk(tr.vs[0]), mk(tr.vs[1]), mk(tr.vs[2]))) {
                hit->t = t;
                hit->u = u;
                hit->v = v;
                hit->ng = ng;
                hit->tri = ti;
                if (any) break;
            }
            if (prev == PALZ_NODE_INVALID) break;  // the root was a leaf
            const uint32_t x = curr;
            curr = prev;
            prev = x;
            continue;
        }
        const palz_internal_node& nd = nodes[curr];
        float l0, l1, r0, r1;
        slabs(ro, ird, mk(nd.aabbs[0].lower), mk(nd.aabbs[0].upper), hit->t, &l0, &l1);
        slabs(ro, ird, mk(nd.aabbs[1].lower), mk(nd.aabbs[1].upper), hit->t, &r0, &r1);
        const bool hit_l = l0 <= l1, hit_r = r0 <= r1;
        uint32_t near_node = nd.children[0], far_node = nd.children[1];
        int n_hits = 0;
        if (hit_l && hit_r) {
            if (r0 < l0) {
                near_node = nd.children[1];
                far_node = nd.children[0];
            }
            n_hits = 2;
        } else if (hit_l || hit_r) {
            n_hits = 1;
            near_node = hit_r ? far_node : near_node;
        }
        uint32_t next;
        if (prev == nd.parent)
            next = 0 < n_hits ? near_node : nd.parent;
        else if (prev == near_node)
            next = n_hits == 2 ? far_node : nd.parent;
        else
            next = nd.parent;
        prev = curr;
        curr = next;
    }
}

// ---------------------------------------------------------------------------
// Sobol photons (sobol.h:17-129)
PALZ_HD uint32_t reverse_bits(uint32_t v) { return __builtin_bitreverse32(v); }
PALZ_HD uint32_t laine_karras(uint32_t x, uint32_t seed) {
    x += seed;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return x;
}
PALZ_HD uint32_t nested_uniform_scramble(uint32_t x, uint32_t seed) {
    return reverse_bits(laine_karras(reverse_bits(x), seed));
}
PALZ_HD uint32_t pascal_p(uint32_t v) {
    v ^= v << 16;
    v ^= (v & 0x00FF00FFu) << 8;
    v ^= (v & 0x0F0F0F0Fu) << 4;
    v ^= (v & 0x33333333u) << 2;
    v ^= (v & 0x55555555u) << 1;
    return v;
}
PALZ_HD float random_float(uint32_t u) { return (float)(u >> 8) / 16777216.0f; }
PALZ_HD void pcg3d(uint32_t* x, uint32_t* y, uint32_t* z) {
    uint32_t vx = *x * 1664525u + 1013904223u;
    uint32_t vy = *y * 1664525u + 1013904223u;
    uint32_t vz = *z * 1664525u + 1013904223u;
    vx += vy * vz;
    vy += vz * vx;
    vz += vx * vy;
    vx ^= vx >> 16u;
    vy ^= vy >> 16u;
    vz ^= vz >> 16u;
    vx += vy * vz;
    vy += vz * vx;
    vz += vx * vy;
    *x = vx;
    *y = vy;
    *z = vz;
}
PALZ_HD void sobol2d(float* x, float* y, uint32_t index, uint32_t p0, uint32_t p1, uint32_t p2) {
    pcg3d(&p0, &p1, &p2);
    const uint32_t v = laine_karras(reverse_bits(index), p0);
    *x = random_float(nested_uniform_scramble(v, p1));
    *y = random_float(nested_uniform_scramble(pascal_p(v), p2));
}
// square2triangle (pk.h:165-178)
PALZ_HD void square_to_triangle(float* x, float* y) {
    if (*y > *x) {
        *x *= 0.5f;
        *y -= *x;
    } else {
        *y *= 0.5f;
        *x -= *y;
    }
}

// occluded (pk.h:88-115)
__host__ __device__ inline bool occluded(const palz_internal_node* nodes, const palz_triangle* tris, uint32_t root, v3 from,
                                v3 from_n, v3 to, v3 to_n) {
    if (dot(to - from, from_n) < 0.0f) from_n = -from_n;
    if (dot(from - to, to_n) < 0.0f) to_n = -to_n;
    const v3 from_safe = from + from_n * ray_offset_scale(from);
    const v3 to_safe = to + to_n * ray_offset_scale(to);
    const v3 rd = to_safe - from_safe;
    return any_hit_escape(nodes, tris, root, from_safe, rd, inv_rd(rd), 1.0f);
}

// getVertex (pk.h:752-758)
PALZ_HD v3 chain_vertex(const ChainTri& c, float u, float v) { return c.v0 + (c.v1 - c.v0) * u + (c.v2 - c.v0) * v; }

// The occlusion ray of occluded() (pk.h:88-115): origin and direction of the
// offset segment, tested over t in [0, 1].
struct Segment {
    v3 ro, rd;
};
PALZ_HD Segment occlusion_segment(v3 from, v3 from_n, v3 to, v3 to_n) {
    if (dot(to - from, from_n) < 0.0f) from_n = -from_n;
    if (dot(from - to, to_n) < 0.0f) to_n = -to_n;
    const v3 from_safe = from + from_n * ray_offset_scale(from);
    const v3 to_safe = to + to_n * ray_offset_scale(to);
    return {from_safe, to_safe - from_safe};
}

// Loads of records at a wave-uniform address through the constant address
// space, so that they become scalar loads into SGPRs (one load per wave
// instead of one per lane).  The tree and the triangles are read-only while a
// kernel runs.
template <class T>
__device__ __forceinline__ T load_uniform(const T* p) {
    static_assert(sizeof(T) % 4 == 0, "dword records");
    typedef const __attribute__((address_space(4))) uint32_t cword;
    cword* q = (cword*)(p);
    T r;
    uint32_t* d = reinterpret_cast<uint32_t*>(&r);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) d[i] = q[i];
    return r;
}
// load_uniform of the record `byte_offset` bytes past `base` (a wave-uniform
// offset, added in the constant address space so that it can ride in the
// scalar load's offset register).
template <class T>
__device__ __forceinline__ T load_uniform_at(const void* base, uint32_t byte_offset) {
    static_assert(sizeof(T) % 4 == 0, "dword records");
    typedef const __attribute__((address_space(4))) char cbyte;
    typedef const __attribute__((address_space(4))) uint32_t cword;
    cword* q = (cword*)((cbyte*)base + byte_offset);
    T r;
    uint32_t* d = reinterpret_cast<uint32_t*>(&r);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) d[i] = q[i];
    return r;
}

// Any-hit traversal of one occlusion ray per lane, run by the whole wave as a
// packet: the wave walks the tree once, loading each node with scalar loads,
// and every lane tests its own ray against the node's child boxes; a subtree
// is entered with the mask of the lanes whose ray hit its box (and every box
// above it), so each lane tests exactly the leaves the escape-link traversal
// (any_hit_escape) would test for it, in another order.  Any-hit is
// independent of the order, so the result equals any_hit_escape's.  `stack`
// is this wave's LDS stack; if it would overflow, the lanes still pending fall
// back to any_hit_escape.  Called by every lane of the wave that reaches it
// (lanes with query == false only take part in the ballots).
constexpr int kPacketDepth = 64;
struct PacketStack {
    uint64_t mask[kPacketDepth];
    uint32_t node[kPacketDepth];
};

// A lane's membership in a wave-uniform lane mask as a branch condition: the
// mask becomes the exec mask, with no vector instruction.
__device__ __forceinline__ bool lane_in(uint64_t lanes) { return __builtin_amdgcn_inverse_ballot_w64(lanes); }
// The first active lane's bit (the lane readfirstlane reads).
__device__ __forceinline__ uint64_t first_lane() {
    const uint64_t active = __ballot(1);
    return active & (~active + 1ull);
}
// A value every lane holds alike, read back as wave-uniform (an LDS load's
// result is otherwise taken as per-lane, and so is all that is computed from
// it).
__device__ __forceinline__ uint32_t uniform32(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
    return (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32)) << 32 |
           __builtin_amdgcn_readfirstlane((uint32_t)v);
}

// The box test of a packet walk: every lane runs it (lanes outside the walk's
// mask hold rays that are never read back), and the caller keeps the ballot's
// bits of the lanes it asks about.
__device__ __forceinline__ bool box_test(v3 ro, v3 ird, const palz_aabb& b, float t_max) {
    float a, c;
    slabs(ro, ird, mk(b.lower), mk(b.upper), t_max, &a, &c);
    return a <= c;
}

__device__ inline bool any_hit_packet(const palz_internal_node* __restrict__ nodes,
                                      const palz_triangle* __restrict__ tris, uint32_t root, bool query, v3 ro,
                                      v3 rd, v3 ird, float t_max, PacketStack* stack) {
    const uint64_t me = 1ull << (threadIdx.x & 63);
    bool hit = false;
    uint64_t pending = __ballot(query);
    if (pending == 0ull) return false;
    auto test_leaf = [&](uint32_t leaf, uint64_t lanes) {
        const palz_triangle tr = load_uniform(tris + (__builtin_amdgcn_readfirstlane(leaf) & ~PALZ_NODE_LEAF_BIT));
        if (lane_in(lanes)) {
            float t, u, v;
            v3 ng;
            if (intersect_tri(&t, &u, &v, &ng, 0.0f, t_max, ro, rd, mk(tr.vs[0]), mk(tr.vs[1]), mk(tr.vs[2])))
                hit = true;
        }
        pending &= ~__ballot(hit);
    };
    if (root & PALZ_NODE_LEAF_BIT) {
        test_leaf(root, pending);
        return hit;
    }
    int sp = 0;
    bool overflow = false;
    uint32_t node = root;
    uint64_t mask = pending;
    for (;;) {
        const palz_internal_node nd = load_uniform(nodes + __builtin_amdgcn_readfirstlane(node));
        const uint32_t c0 = nd.children[0], c1 = nd.children[1];
        // every lane tests the boxes; the ballots keep the lanes of `mask`
        uint64_t m0 = __ballot(box_test(ro, ird, nd.aabbs[0], t_max)) & mask;
        uint64_t m1 = __ballot(box_test(ro, ird, nd.aabbs[1], t_max)) & mask;
        if (m0 != 0ull && (c0 & PALZ_NODE_LEAF_BIT)) {
            test_leaf(c0, m0);
            m0 = 0ull;
        }
        if (m1 != 0ull && (c1 & PALZ_NODE_LEAF_BIT)) {
            test_leaf(c1, m1);
            m1 = 0ull;
        }
        m0 &= pending;
        m1 &= pending;
        if (m0 != 0ull && m1 != 0ull) {
            if (sp == kPacketDepth) {
                overflow = true;
                break;
            }
            if (lane_in(first_lane())) {
                stack->mask[sp] = m1;
                stack->node[sp] = c1;
            }
            ++sp;
            node = c0;
            mask = m0;
        } else if (m0 != 0ull) {
            node = c0;
            mask = m0;
        } else if (m1 != 0ull) {
            node = c1;
            mask = m1;
        } else {
            mask = 0ull;
            while (sp > 0 && mask == 0ull) {
                --sp;
                __builtin_amdgcn_wave_barrier();
                node = uniform32(stack->node[sp]);
                mask = uniform64(stack->mask[sp]) & pending;
            }
            if (mask == 0ull) break;
        }
        if (pending == 0ull) break;
    }
    if (overflow && (pending & me)) hit = any_hit_escape(nodes, tris, root, ro, rd, ird, t_max);
    return hit;
}

// The entry and exit planes, on one axis, of a wide node's six boxes (slots
// as in palz_wide_node): plane[a][0..1] of the node for rays whose direction
// component a is positive, plane[a][1..2] for negative ones.
struct WidePlanes {
    float entry[6], exit[6];
};
// Per wave and ray: the byte offset of each axis' WidePlanes in a wide node,
// and whether every lane that queries has one direction sign per axis
// (`ordered`: then the planes at those offsets are each lane's entry and exit
// planes).
struct WaveSlabs {
    bool ordered;
    uint32_t offset[3];
};
__device__ inline WaveSlabs wave_slabs(v3 ird, uint64_t lanes) {
    const uint64_t neg[3] = {__ballot(ird.x < 0.0f) & lanes, __ballot(ird.y < 0.0f) & lanes,
                             __ballot(ird.z < 0.0f) & lanes};
    WaveSlabs w;
    w.ordered = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        w.ordered = w.ordered && (neg[a] == 0ull || neg[a] == lanes);
        w.offset[a] = (uint32_t)offsetof(palz_wide_node, plane) + (uint32_t)(a * 3 * 6 * sizeof(float)) +
                      (neg[a] != 0ull ? (uint32_t)(6 * sizeof(float)) : 0u);
    }
    return w;
}
// slabs() for a ray whose direction signs are known: `nr` holds each axis'
// entry plane (lo.x where ird.x > 0, hi.x where ird.x < 0) and `fr` its exit
// plane.  On an axis with ird.x > 0, lo.x <= hi.x gives lo.x - ro.x <=
// hi.x - ro.x, and (lo.x - ro.x) * ird.x <= (hi.x - ro.x) * ird.x since
// rounding is monotone; so slabs()'s fold_min of the two is the entry plane's
// t and its fold_max the exit plane's, up to the sign of a zero (inv_rd never
// returns 0, and no t is NaN).  With ird.x < 0 the roles swap.  For boxes
// with lower <= upper (every box the tree holds) the results compare as
// slabs()'s do.
PALZ_HD void slabs_ordered(v3 ro, v3 ird, v3 nr, v3 fr, float known_t, float* rmin, float* rmax) {
    const v3 tn = (nr - ro) * ird;
    const v3 tf = (fr - ro) * ird;
    *rmin = fold_max(fold_max(fold_max(tn.x, tn.y), tn.z), 0.0f);
    *rmax = fold_min(fold_min(fold_min(tf.x, tf.y), tf.z) * 1.00000024f, known_t);
}

// any_hit_packet over the BVH4 view (palz_wide_node): one scalar load brings
// a node's two child boxes and its internal children's child boxes, so the
// wave's walk takes one dependent load per two levels.  A lane still tests a
// grandchild's box only if it hit the child's box, so each lane tests exactly
// the leaves any_hit_escape would, and the result is the same.  The boxes are
// read by plane at the offsets of `ws`; kOrdered (ws.ordered) tests them with
// slabs_ordered, else with slabs(), whose result does not depend on which of
// the two planes of an axis comes first.
template <bool kOrdered>
__device__ inline bool any_hit_wide(const palz_internal_node* __restrict__ nodes,
                                    const palz_wide_node* __restrict__ wide, const palz_triangle* __restrict__ tris,
                                    uint32_t root, bool query, v3 ro, v3 rd, v3 ird, const WaveSlabs& ws,
                                    float t_max, PacketStack* stack) {
    const uint64_t me = 1ull << (threadIdx.x & 63);
    bool hit = false;
    uint64_t pending = __ballot(query);
    if (pending == 0ull) return false;
    auto test_leaf = [&](uint32_t leaf, uint64_t lanes) {
        const palz_triangle tr = load_uniform(tris + (__builtin_amdgcn_readfirstlane(leaf) & ~PALZ_NODE_LEAF_BIT));
        if (lane_in(lanes)) {
            float t, u, v;
            v3 ng;
            if (intersect_tri(&t, &u, &v, &ng, 0.0f, t_max, ro, rd, mk(tr.vs[0]), mk(tr.vs[1]), mk(tr.vs[2])))
                hit = true;
        }
        pending &= ~__ballot(hit);
    };
    if (root & PALZ_NODE_LEAF_BIT) {
        test_leaf(root, pending);
        return hit;
    }
    struct Head {
        palz_node_index child[2];
        palz_node_index grand[4];
        uint32_t pad_[2];
    };
    int sp = 0;
    bool overflow = false;
    uint32_t node = root;
    uint64_t mask = pending;
    for (;;) {
        const palz_wide_node* at = wide + __builtin_amdgcn_readfirstlane(node);
        const Head w = load_uniform(reinterpret_cast<const Head*>(at));
        const WidePlanes px = load_uniform_at<WidePlanes>(at, ws.offset[0]);
        const WidePlanes py = load_uniform_at<WidePlanes>(at, ws.offset[1]);
        const WidePlanes pz = load_uniform_at<WidePlanes>(at, ws.offset[2]);
        // every lane tests the box; the ballot keeps the lanes of `lanes`
        auto box_hit = [&](int s, uint64_t lanes) -> uint64_t {
            const v3 nr = mk(px.entry[s], py.entry[s], pz.entry[s]);
            const v3 fr = mk(px.exit[s], py.exit[s], pz.exit[s]);
            float a, c;
            if (kOrdered)
                slabs_ordered(ro, ird, nr, fr, t_max, &a, &c);
            else
                slabs(ro, ird, nr, fr, t_max, &a, &c);
            return __ballot(a <= c) & lanes;
        };
        // the first internal grandchild hit is walked next, the others pushed
        uint32_t first = PALZ_NODE_INVALID;
        uint64_t first_mask = 0ull;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const uint32_t ch = w.child[c];
            const uint64_t mc = box_hit(c, mask) & pending;
            if (mc == 0ull) continue;
            if (ch & PALZ_NODE_LEAF_BIT) {
                test_leaf(ch, mc);
                continue;
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const uint32_t g = w.grand[2 * c + j];
                const uint64_t mg = box_hit(2 + 2 * c + j, mc & pending) & pending;
                if (mg == 0ull) continue;
                if (g & PALZ_NODE_LEAF_BIT) {
                    test_leaf(g, mg);
                } else if (first_mask == 0ull) {
                    first = g;
                    first_mask = mg;
                } else if (sp == kPacketDepth) {
                    overflow = true;
                } else {
                    if (lane_in(first_lane())) {
                        stack->mask[sp] = mg;
                        stack->node[sp] = g;
                    }
                    ++sp;
                }
            }
        }
        if (overflow || pending == 0ull) break;
        if (first_mask != 0ull) {
            node = first;
            mask = first_mask;
        } else {
            mask = 0ull;
            while (sp > 0 && mask == 0ull) {
                --sp;
                __builtin_amdgcn_wave_barrier();
                node = uniform32(stack->node[sp]);
                mask = uniform64(stack->mask[sp]) & pending;
            }
            if (mask == 0ull) break;
        }
    }
    if (overflow && (pending & me)) hit = any_hit_escape(nodes, tris, root, ro, rd, ird, t_max);
    return hit;
}

// contributableThroughput<K> (pk.h:760-842) without its occlusion rays: the
// throughput of a chain that passes every other test (0 otherwise) and, when
// it is positive, the K + 1 occlusion segments to test.
template <int K>
__host__ __device__ inline float throughput_unoccluded(const float* params, v3 p_beg, v3 p_end, const ChainTri* tri,
                                                      uint32_t event, float eta, Segment* seg) {
    v3 vert[K + 2];
    v3 sn[K + 2];
    vert[0] = p_beg;
    vert[K + 1] = p_end;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const float u = params[2 * k], v = params[2 * k + 1];
        if (u < 0.0f || v < 0.0f || 1.0f < u + v) return 0.0f;
        vert[k + 1] = chain_vertex(tri[k], u, v);
        sn[k + 1] = normalize(tri[k].s0 + (tri[k].s1 - tri[k].s0) * u + (tri[k].s2 - tri[k].s0) * v);
    }
    sn[0] = normalize(vert[1] - vert[0]);
    sn[K + 1] = normalize(vert[K] - vert[K + 1]);
    float throughput = 1.0f;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const v3 wi = vert[k] - vert[k + 1];
        const v3 wo = vert[k + 2] - vert[k + 1];
        const v3 n = sn[k + 1];
        const bool e_T = !(0.0f < dot(wi, n) * dot(wo, n));
        if (ev_is_T(event, k) != e_T) return 0.0f;
        const v3 ng = cross(tri[k].v1 - tri[k].v0, tri[k].v2 - tri[k].v0);  // unnormalizedNormalOf
        const bool eg_T = !(0.0f < dot(wi, ng) * dot(wo, ng));
        if (e_T != eg_T) return 0.0f;
        if (tri[k].material == PALZ_MATERIAL_DIELECTRIC) {
            const float r = fresnel_exact_norm_free(wi, n, eta);
            throughput *= e_T ? 1.0f - r : r;
        }
    }
#pragma unroll
    for (int i = 0; i < K + 1; ++i) {
        const v3 d = vert[i] - vert[i + 1];
        if (dot(d, d) < 1.0e-5f) return 0.0f;
    }
#pragma unroll
    for (int i = 0; i < K + 1; ++i) seg[i] = occlusion_segment(vert[i], sn[i], vert[i + 1], sn[i + 1]);
    return throughput;
}

// contributableThroughput<K> (pk.h:760-842)
template <int K>
__host__ __device__ inline float contributable_throughput(const float* params, v3 p_beg, v3 p_end, const ChainTri* tri,
                                                 uint32_t event, const palz_internal_node* nodes,
                                                 const palz_triangle* tris, uint32_t root, float eta) {
    v3 vert[K + 2];
    v3 sn[K + 2];
    vert[0] = p_beg;
    vert[K + 1] = p_end;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const float u = params[2 * k], v = params[2 * k + 1];
        if (u < 0.0f || v < 0.0f || 1.0f < u + v) return 0.0f;
        vert[k + 1] = chain_vertex(tri[k], u, v);
        sn[k + 1] = normalize(tri[k].s0 + (tri[k].s1 - tri[k].s0) * u + (tri[k].s2 - tri[k].s0) * v);
    }
    sn[0] = normalize(vert[1] - vert[0]);
    sn[K + 1] = normalize(vert[K] - vert[K + 1]);
    float throughput = 1.0f;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const v3 wi = vert[k] - vert[k + 1];
        const v3 wo = vert[k + 2] - vert[k + 1];
        const v3 n = sn[k + 1];
        const bool e_T = !(0.0f < dot(wi, n) * dot(wo, n));
        if (ev_is_T(event, k) != e_T) return 0.0f;
        const v3 ng = cross(tri[k].v1 - tri[k].v0, tri[k].v2 - tri[k].v0);  // unnormalizedNormalOf
        const bool eg_T = !(0.0f < dot(wi, ng) * dot(wo, ng));
        if (e_T != eg_T) return 0.0f;
        if (tri[k].material == PALZ_MATERIAL_DIELECTRIC) {
            const float r = fresnel_exact_norm_free(wi, n, eta);
            throughput *= e_T ? 1.0f - r : r;
        }
    }
#pragma unroll
    for (int i = 0; i < K + 1; ++i) {
        const v3 d = vert[i] - vert[i + 1];
        if (dot(d, d) < 1.0e-5f) return 0.0f;
    }
    for (int i = 0; i < K + 1; ++i)
        if (occluded(nodes, tris, root, vert[i], sn[i], vert[i + 1], sn[i + 1])) return 0.0f;
    return throughput;
}

// GetOrthonormalBasis (pk.h:453-459)
PALZ_HD void orthonormal_basis(v3 z, v3* x, v3* y) {
    const float sign = copysignf(1.0f, z.z);
    const float a = -1.0f / (sign + z.z);
    const float b = z.x * z.y * a;
    *x = mk(1.0f + sign * z.x * z.x * a, sign * b, -sign * z.x);
    *y = mk(b, sign + z.y * z.y * a, -z.y);
}

// intersect_p_ray_plane (pk.h:594-598)
template <int NT, class F>
PALZ_HD dual3<NT, F> ray_plane(const dual3<NT, F>& ro, const dual3<NT, F>& rd, const dual3<NT, F>& ng,
                               const dual3<NT, F>& v0) {
    const dual<NT, F> t = ddot(v0 - ro, ng) / ddot(ng, rd);
    return ro + rd * t;
}

// dAdw (pk.h:844-930).  Both light-direction tangents run in one pass (NT=2),
// bit-identical to the reference's two passes.  `nev` = number of chain
// vertices (eDescriptor.size()).
template <int K>
__host__ __device__ inline float dadw(v3 ro, v3 rd, v3 p_end, const ChainTri* tri, uint32_t event, float eta) {
    using D = dual<2, float>;
    using D3 = dual3<2, float>;
    rd = normalize(rd);
    D d0 = dc<2, float>(0.0f), d1 = dc<2, float>(0.0f);
    d0.g[0] = 1.0f;
    d1.g[1] = 1.0f;
    v3 T0, T1;
    orthonormal_basis(rd, &T0, &T1);
    D3 ro_j = dc3<2, float>(ro);
    D3 rd_j = dc3<2, float>(rd) + dc3<2, float>(T0) * d0 + dc3<2, float>(T1) * d1;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const ChainTri& c = tri[j];
        const v3 ng = cross(c.v1 - c.v0, c.v2 - c.v0);
        const D3 p = ray_plane(ro_j, rd_j, dc3<2, float>(ng), dc3<2, float>(c.v0));
        D v = ddot(p - dc3<2, float>(c.v1), dc3<2, float>(cross(ng, c.v1 - c.v0)));
        D w = ddot(p - dc3<2, float>(c.v2), dc3<2, float>(cross(ng, c.v2 - c.v1)));
        D u = ddot(p - dc3<2, float>(c.v0), dc3<2, float>(cross(ng, c.v0 - c.v2)));
        const D area = u + v + w;
        u = u / area;
        v = v / area;
        w = w / area;
        const D3 n = dc3<2, float>(c.s0) * w + dc3<2, float>(c.s1) * u + dc3<2, float>(c.s2) * v;
        const D3 wi = -rd_j;
        D3 wo;
        if (ev_is_T(event, j)) {
            if (ddot(wi, n).v < 0.0f) {
                wo = drefraction(wi, -n, 1.0f / eta);
                if (wo.x.v == 0.0f && wo.y.v == 0.0f && wo.z.v == 0.0f) return kFltMax;
            } else {
                wo = drefraction(wi, n, eta);
            }
        } else {
            wo = dreflection(wi, n);
        }
        ro_j = p;
        rd_j = wo;
    }
    const v3 p_last = mk(ro_j.x.v, ro_j.y.v, ro_j.z.v);
    const D3 pf = ray_plane(ro_j, rd_j, dc3<2, float>(p_last - p_end), dc3<2, float>(p_end));
    const v3 a0 = mk(pf.x.g[0], pf.y.g[0], pf.z.g[0]);
    const v3 a1 = mk(pf.x.g[1], pf.y.g[1], pf.z.g[1]);
    const float len0 = length(a0);
    const float len1 = length(a1);
    const float sin_theta = length(cross(a0 / len0, a1 / len1));
    return len0 * len1 * hmax(sin_theta, 0.01f);
}

}  // namespace palz
