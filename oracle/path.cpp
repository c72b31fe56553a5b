// Oracle: Sobol photons, photon tracing, the tuple set, the specular-chain
// Newton solver (FP32 and FP64), path throughput, occlusion and dA/dw.
// TEST INFRASTRUCTURE — a scalar restatement written from the reference text;
// functions cite the reference lines they follow.
#include <algorithm>
#include <cstring>
#include <map>
#include <thread>
#include <tuple>
#include <vector>

#include "oracle.hpp"

using namespace orc;

namespace {

// ---------------------------------------------------------------------------
// helper_math host semantics (helper_math.h:134-142)
template <class T>
T hfmin(T a, T b) { return a < b ? a : b; }
template <class T>
T hfmax(T a, T b) { return a > b ? a : b; }
template <class T>
vec3<T> normalize(vec3<T> v) { return v * (T(1) / std::sqrt(dot(v, v))); }  // rsqrtf = 1/sqrtf
template <class T>
T length(vec3<T> v) { return std::sqrt(dot(v, v)); }
template <class T>
vec3<T> cvt(const palz_float3& p) { return {(T)p.x, (T)p.y, (T)p.z}; }
template <class T>
vec3<T> cvt(float3 p) { return {(T)p.x, (T)p.y, (T)p.z}; }

// ---------------------------------------------------------------------------
// sobol.h:17-129
uint32_t reverseBits(uint32_t v) {
    v = ((v >> 1) & 0x55555555) | ((v & 0x55555555) << 1);
    v = ((v >> 2) & 0x33333333) | ((v & 0x33333333) << 2);
    v = ((v >> 4) & 0x0F0F0F0F) | ((v & 0x0F0F0F0F) << 4);
    v = ((v >> 8) & 0x00FF00FF) | ((v & 0x00FF00FF) << 8);
    v = (v >> 16) | (v << 16);
    return v;
}
uint32_t laine_karras_permutation(uint32_t x, uint32_t seed) {
    x += seed;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return x;
}
uint32_t nested_uniform_scramble(uint32_t x, uint32_t seed) {
    x = reverseBits(x);
    x = laine_karras_permutation(x, seed);
    x = reverseBits(x);
    return x;
}
uint32_t P(uint32_t v) {
    v ^= v << 16;
    v ^= (v & 0x00FF00FF) << 8;
    v ^= (v & 0x0F0F0F0F) << 4;
    v ^= (v & 0x33333333) << 2;
    v ^= (v & 0x55555555) << 1;
    return v;
}
float random_float(uint32_t u) { return (float)(u >> 8u) / ((0xFFFFFFFFu >> 8u) + 1u); }
void pcg3d(uint32_t* x, uint32_t* y, uint32_t* z) {
    uint32_t vx = *x, vy = *y, vz = *z;
    vx = vx * 1664525u + 1013904223u;
    vy = vy * 1664525u + 1013904223u;
    vz = vz * 1664525u + 1013904223u;
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
void shuffled_scrambled_sobol_2d(float* x, float* y, uint32_t index, uint32_t p0, uint32_t p1, uint32_t p2) {
    pcg3d(&p0, &p1, &p2);
    uint32_t v = laine_karras_permutation(reverseBits(index), p0);
    *x = random_float(nested_uniform_scramble(v, p1));
    *y = random_float(nested_uniform_scramble(P(v), p2));
}
void square2triangle(float* x, float* y) {  // pk.h:165-178
    if (*y > *x) {
        *x *= 0.5f;
        *y -= *x;
    } else {
        *y *= 0.5f;
        *x -= *y;
    }
}

// ---------------------------------------------------------------------------
// FLOP-counting scalar: every + - * / and sqrt on it counts one operation
// (sign flips and comparisons count none).  Instantiating the solver with it
// counts the reference iteration's arithmetic exactly.
struct Flop {
    double v;
    static thread_local uint64_t n;
    Flop() : v(0) {}
    Flop(double x) : v(x) {}
};
thread_local uint64_t Flop::n = 0;
inline Flop operator+(Flop a, Flop b) { ++Flop::n; return a.v + b.v; }
inline Flop operator-(Flop a, Flop b) { ++Flop::n; return a.v - b.v; }
inline Flop operator*(Flop a, Flop b) { ++Flop::n; return a.v * b.v; }
inline Flop operator/(Flop a, Flop b) { ++Flop::n; return a.v / b.v; }
inline Flop operator-(Flop a) { return -a.v; }
inline Flop& operator+=(Flop& a, Flop b) { return a = a + b; }
inline Flop& operator-=(Flop& a, Flop b) { return a = a - b; }
inline Flop& operator*=(Flop& a, Flop b) { return a = a * b; }
inline bool operator<(Flop a, Flop b) { return a.v < b.v; }
inline bool operator>(Flop a, Flop b) { return a.v > b.v; }
template <class T>
T tsqrt(T x) { return std::sqrt(x); }
inline Flop tsqrt(Flop x) { ++Flop::n; return std::sqrt(x.v); }
template <class T>
T tabs(T x) { return std::fabs(x); }
inline Flop tabs(Flop x) { return std::fabs(x.v); }

// ---------------------------------------------------------------------------
// saka::dval (saka.h:11-208): one tangent, generic scalar
template <class T>
struct dval {
    T v, g;
    dval() : v(0), g(0) {}
    dval(T x) : v(x), g(0) {}
    dval(T x, bool rg) : v(x), g(rg ? T(1) : T(0)) {}
};
template <class T>
dval<T> binary(dval<T> x, dval<T> y, T f, T dfdx, T dfdy) {
    dval<T> u;
    u.v = f;
    u.g = x.g * dfdx + y.g * dfdy;
    return u;
}
template <class T>
dval<T> operator+(dval<T> x, dval<T> y) { return binary(x, y, x.v + y.v, T(1), T(1)); }
template <class T>
dval<T> operator-(dval<T> x) {
    dval<T> u;
    u.v = -x.v;
    u.g = x.g * T(-1);
    return u;
}
template <class T>
dval<T> operator-(dval<T> x, dval<T> y) { return binary(x, y, x.v - y.v, T(1), T(-1)); }
template <class T>
dval<T> operator*(dval<T> x, dval<T> y) { return binary(x, y, x.v * y.v, y.v, x.v); }
template <class T>
dval<T> operator/(dval<T> x, dval<T> y) { return binary(x, y, x.v / y.v, T(1) / y.v, -x.v / (y.v * y.v)); }
template <class T>
dval<T> sqrt(dval<T> x) {
    dval<T> u;
    u.v = tsqrt(x.v);
    u.g = x.g * (T(0.5) / tsqrt(x.v));
    return u;
}
template <class T>
struct dval3 {
    dval<T> x, y, z;
};
template <class T>
dval3<T> mkd(vec3<T> v) { return {dval<T>(v.x), dval<T>(v.y), dval<T>(v.z)}; }
template <class T>
dval3<T> operator+(dval3<T> a, dval3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T>
dval3<T> operator-(dval3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T>
dval3<T> operator-(dval3<T> a, dval3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T>
dval3<T> operator*(dval3<T> a, dval<T> s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T>
dval3<T> operator/(dval3<T> a, dval<T> s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T>
dval<T> dot(dval3<T> a, dval3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T>
dval3<T> cross(dval3<T> a, dval3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T>
dval3<T> reflection(dval3<T> wi, dval3<T> n) { return n * dot(wi, n) * dval<T>(T(2)) / dot(n, n) - wi; }
template <class T>
dval3<T> refraction_norm_free(dval3<T> wi, dval3<T> n, T eta) {
    dval<T> NoN = dot(n, n);
    dval<T> WIoN = dot(wi, n);
    dval<T> WoW = dot(wi, wi);
    dval<T> k = NoN * WoW * dval<T>(eta * eta - T(1)) + WIoN * WIoN;
    if (k.v < T(0)) return {dval<T>(T(0)), dval<T>(T(0)), dval<T>(T(0))};
    return -wi * NoN + n * (WIoN - sqrt(k));
}

// ---------------------------------------------------------------------------
// sen.h: column-major dense matrices and Householder least squares
template <class T, int R, int C>
struct Mat {
    T m[R * C];
    T& operator()(int r, int c) { return m[c * R + r]; }
    const T& operator()(int r, int c) const { return m[c * R + r]; }
};
template <class T, int R, int C>
T column_dot(const Mat<T, R, C>& a, int c0, int c1) {  // sen.h:347-356
    T r = T(0);
    for (int i = 0; i < R; ++i) r += a.m[c0 * R + i] * a.m[c1 * R + i];
    return r;
}
template <class T, int R, int C>
void solve_qr_overdetermined(Mat<T, R, C> A, Mat<T, R, 1> b, T* x) {  // sen.h:658-747
    constexpr int H = C < R - 1 ? C : R - 1;
    Mat<T, R, H> vs;
    for (int i = 0; i < H; i++) {
        for (int j = 0; j < R; j++) vs(j, i) = j < i ? T(0) : -A(j, i);
        T sgn = T(0) < vs(i, i) ? T(1) : T(-1);
        T x_len = tsqrt(column_dot(vs, i, i));
        vs(i, i) += sgn * x_len;
        T v_len2 = T(2) * x_len * (x_len + tabs(A(i, i)));
        A(i, i) = sgn * x_len;
        for (int r = i + 1; r < R; r++) A(r, i) = T(0);
        for (int c = i + 1; c < C; c++) {
            T VdotCol = T(0);
            for (int r = i; r < R; r++) VdotCol += vs(r, i) * A(r, c);
            T k = T(2) * VdotCol / v_len2;
            for (int r = i; r < R; r++) A(r, c) -= k * vs(r, i);
        }
    }
    for (int i = 0; i < H; i++) {  // applyQTransposed
        T VdotCol = T(0), VdotV = T(0);
        for (int r = i; r < R; r++) {
            T v = vs(r, i);
            VdotCol += v * b(r, 0);
            VdotV += v * v;
        }
        T k = T(2) * VdotCol / VdotV;
        for (int r = i; r < R; r++) b(r, 0) -= k * vs(r, i);
    }
    for (int r = C - 1; 0 <= r; r--) {
        T sum = T(0);
        for (int i = r + 1; i < C; i++) sum += A(r, i) * x[i];
        x[r] = (b(r, 0) - sum) / A(r, r);
    }
}

// ---------------------------------------------------------------------------
struct Scene {
    const palz_triangle* tris;
    const palz_triangle_attrib* attribs;
    const palz_internal_node* nodes;
    uint32_t root;
};

bool is_T(uint32_t event, int k) { return (event >> k) & 1u; }

// The residual b, its Jacobian A and the cost of one solveConstraints
// iteration (pk.h:479-548): 2K single-tangent passes.
template <int K, class T>
void evalConstraints(const T* parameters, vec3<T> p_beg, vec3<T> p_end, const palz_triangle* tris,
                     const palz_triangle_attrib* attribs, T eta, uint32_t event, Mat<T, K * 3, K * 2>& A,
                     Mat<T, K * 3, 1>& b, T* cost_out) {
    const int nParameters = K * 2;
    T cost = T(0);
    for (int i = 0; i < nParameters; i++) {
        dval3<T> vertices[K + 2];
        vertices[0] = mkd(p_beg);
        vertices[K + 1] = mkd(p_end);
        dval3<T> shadingNormals[K];
        for (int k = 0; k < K; k++) {
            dval<T> param_u(parameters[k * 2 + 0], k * 2 + 0 == i);
            dval<T> param_v(parameters[k * 2 + 1], k * 2 + 1 == i);
            // edges in the working precision (for T = float: the reference's float3 ops)
            vec3<T> v0 = cvt<T>(tris[k].vs[0]), v1 = cvt<T>(tris[k].vs[1]), v2 = cvt<T>(tris[k].vs[2]);
            vec3<T> e0 = v1 - v0, e1 = v2 - v0;
            vertices[k + 1] = mkd(v0) + mkd(e0) * param_u + mkd(e1) * param_v;
            vec3<T> n0 = cvt<T>(attribs[k].shading_normals[0]);
            vec3<T> ne0 = cvt<T>(attribs[k].shading_normals[1]) - n0;
            vec3<T> ne1 = cvt<T>(attribs[k].shading_normals[2]) - n0;
            shadingNormals[k] = mkd(n0) + mkd(ne0) * param_u + mkd(ne1) * param_v;
        }
        for (int k = 0; k < K; k++) {
            dval3<T> wi = vertices[k] - vertices[k + 1];
            dval3<T> wo = vertices[k + 2] - vertices[k + 1];
            dval3<T> n = shadingNormals[k];
            if (dot(wi, n).v < T(0)) std::swap(wi, wo);
            dval3<T> wo_optimize = is_T(event, k) ? refraction_norm_free(wi, n, eta) : reflection(wi, n);
            dval3<T> c = cross(wo, wo_optimize);
            A(k * 3 + 0, i) = c.x.g;
            A(k * 3 + 1, i) = c.y.g;
            A(k * 3 + 2, i) = c.z.g;
            b(k * 3 + 0, 0) = c.x.v;
            b(k * 3 + 1, 0) = c.y.v;
            b(k * 3 + 2, 0) = c.z.v;
            if (i == 0) cost += dot(c, c).v;
        }
    }
    *cost_out = cost;
}

// solveConstraints<K> (pk.h:466-592)
template <int K, class T>
bool solveConstraints(T* parameters, vec3<T> p_beg, vec3<T> p_end, const palz_triangle* tris,
                      const palz_triangle_attrib* attribs, T eta, uint32_t event, int maxIterations, T costTolerance,
                      int earlyTermWarmUp, int* iters) {
    const int nParameters = K * 2;
    T edge_eps = T(0.25);
    for (int iter = 0; iter < maxIterations; iter++) {
        Mat<T, K * 3, K * 2> A;
        Mat<T, K * 3, 1> b;
        T cost = T(0);
        evalConstraints<K, T>(parameters, p_beg, p_end, tris, attribs, eta, event, A, b, &cost);
        if (cost < costTolerance) {
            *iters = iter;
            return true;
        }
        T dparams[K * 2];
        solve_qr_overdetermined(A, b, dparams);
        for (int i = 0; i < nParameters; i++) parameters[i] = parameters[i] - dparams[i];
        if (earlyTermWarmUp <= iter) {
            for (int k = 0; k < K; k++) {
                T param_u = parameters[k * 2 + 0];
                T param_v = parameters[k * 2 + 1];
                if (-edge_eps < param_u && -edge_eps < param_v && param_u + param_v < T(1) + edge_eps) continue;
                *iters = iter;
                return false;
            }
            edge_eps *= edge_eps;
        }
    }
    *iters = maxIterations;
    return false;
}

// The refinement step of the batched solver: one Newton step from the FP32
// solution with the residual evaluated in FP64 at the FP32 parameters, the
// Jacobian of the FP32 iteration at the same point (float dual passes,
// entries widened to double) and the least-squares correction in FP64 (mixed
// precision iterative refinement: the residual carries the precision).
template <int K>
void refineStep(const float* params32, float3 p_beg, float3 p_end, const palz_triangle* tris,
                const palz_triangle_attrib* attribs, float eta, uint32_t event, double* out) {
    Mat<float, K * 3, K * 2> A32;
    Mat<float, K * 3, 1> b32;
    float cost32;
    evalConstraints<K, float>(params32, p_beg, p_end, tris, attribs, eta, event, A32, b32, &cost32);
    double p64[K * 2];
    for (int i = 0; i < 2 * K; i++) p64[i] = (double)params32[i];
    Mat<double, K * 3, K * 2> A;
    Mat<double, K * 3, 1> b;
    double cost;
    evalConstraints<K, double>(p64, cvt<double>(p_beg), cvt<double>(p_end), tris, attribs, (double)eta, event, A,
                               b, &cost);
    for (int i = 0; i < K * 3 * K * 2; i++) A.m[i] = (double)A32.m[i];
    double dp[K * 2];
    solve_qr_overdetermined(A, b, dp);
    for (int i = 0; i < 2 * K; i++) out[i] = p64[i] - dp[i];
}

// ---------------------------------------------------------------------------
// ray casting (minimum_lbvh.h:82-96, :372-422, :913-988, :1078-1152)
float3 invRd(float3 rd) {
    float3 q = {1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z};
    const float M = 3.402823466e+38F;
    return {hfmax(-M, hfmin(q.x, M)), hfmax(-M, hfmin(q.y, M)), hfmax(-M, hfmin(q.z, M))};
}
void slabs(float3 ro, float3 ird, float3 lo, float3 hi, float knownHitT, float* a, float* b) {
    float3 t0 = {(lo.x - ro.x) * ird.x, (lo.y - ro.y) * ird.y, (lo.z - ro.z) * ird.z};
    float3 t1 = {(hi.x - ro.x) * ird.x, (hi.y - ro.y) * ird.y, (hi.z - ro.z) * ird.z};
    float3 tmin = {hfmin(t0.x, t1.x), hfmin(t0.y, t1.y), hfmin(t0.z, t1.z)};
    float3 tmax = {hfmax(t0.x, t1.x), hfmax(t0.y, t1.y), hfmax(t0.z, t1.z)};
    float region_min = hfmax(hfmax(tmin.x, tmin.y), tmin.z);
    float region_max = hfmin(hfmin(tmax.x, tmax.y), tmax.z) * 1.00000024f;
    *a = hfmax(region_min, 0.0f);
    *b = hfmin(region_max, knownHitT);
}
bool intersectRayTriangle(float* tOut, float* uOut, float* vOut, float3* ngOut, float t_min, float t_max, float3 ro,
                          float3 rd, float3 v0, float3 v1, float3 v2) {
    float3 e0 = v1 - v0;
    float3 e1 = v2 - v1;
    float3 ng = cross(e0, e1);
    float t = dot(v0 - ro, ng) / dot(ng, rd);
    if (t_min <= t && t <= t_max) {
        float3 e2 = v0 - v2;
        float u_vol = dot(cross(e2, v0 + v2 - ro * 2.0f), rd);
        float v_vol = dot(cross(e0, v1 + v0 - ro * 2.0f), rd);
        float w_vol = dot(cross(e1, v2 + v1 - ro * 2.0f), rd);
        if (hfmin(hfmin(u_vol, v_vol), w_vol) < 0.0f && 0.0f < hfmax(hfmax(u_vol, v_vol), w_vol)) return false;
        float vol = u_vol + v_vol + w_vol;
        *tOut = t;
        *uOut = u_vol / vol;
        *vOut = v_vol / vol;
        *ngOut = ng;
        return true;
    }
    return false;
}
struct Hit {
    float t = 3.402823466e+38F;
    float u = 0, v = 0;
    float3 ng = {0, 0, 0};
    uint32_t tri = 0xFFFFFFFFu;
};
constexpr uint32_t kLeaf = PALZ_NODE_LEAF_BIT;
constexpr uint32_t kInvalid = PALZ_NODE_INVALID;

// intersect_stackfull; *max_sp receives the deepest stack use
// Algorithmic work of photonTrace, counted when set (orc_photon_trace_work):
// [0] internal nodes visited, [1] triangles tested, [2] attribute records
// read, [3] tuples stored, [4] photons traced.
uint64_t* g_work = nullptr;
uint32_t g_tri_stride = 1;

void intersect_stackfull(Hit* hit, const Scene& s, float3 ro, float3 rd, float3 ird, int* max_sp) {
    std::vector<uint32_t> stack;
    uint32_t node = s.root;
    while (node != kInvalid) {
        if (g_work) ++g_work[(node & kLeaf) ? 1 : 0];
        if (node & kLeaf) {
            const palz_triangle& tri = s.tris[node & ~kLeaf];
            float t, u, v;
            float3 ng;
            if (intersectRayTriangle(&t, &u, &v, &ng, 0.0f, hit->t, ro, rd, load3(tri.vs[0]), load3(tri.vs[1]),
                                     load3(tri.vs[2]))) {
                hit->t = t;
                hit->u = u;
                hit->v = v;
                hit->ng = ng;
                hit->tri = node & ~kLeaf;
            }
            if (stack.empty()) {
                node = kInvalid;
            } else {
                node = stack.back();
                stack.pop_back();
            }
            continue;
        }
        const palz_internal_node& nd = s.nodes[node];
        float l0, l1, r0, r1;
        slabs(ro, ird, load3(nd.aabbs[0].lower), load3(nd.aabbs[0].upper), hit->t, &l0, &l1);
        slabs(ro, ird, load3(nd.aabbs[1].lower), load3(nd.aabbs[1].upper), hit->t, &r0, &r1);
        bool hitL = l0 <= l1, hitR = r0 <= r1;
        if (hitL && hitR) {
            uint32_t near_node = nd.children[0], far_node = nd.children[1];
            if (r0 < l0) std::swap(near_node, far_node);
            node = near_node;
            stack.push_back(far_node);
            if (max_sp && (int)stack.size() > *max_sp) *max_sp = (int)stack.size();
        } else if (hitL || hitR) {
            node = nd.children[hitL ? 0 : 1];
        } else if (stack.empty()) {
            node = kInvalid;
        } else {
            node = stack.back();
            stack.pop_back();
        }
    }
}

// intersect_stackfree (minimum_lbvh.h:996-1076); `any` is RAY_QUERY_ANY.  A
// leaf root is tested directly.
void intersect_stackfree(Hit* hit, const Scene& s, float3 ro, float3 rd, float3 ird, bool any) {
    uint32_t curr = s.root, prev = kInvalid;
    while (curr != kInvalid) {
        if (curr & kLeaf) {
            const palz_triangle& tri = s.tris[curr & ~kLeaf];
            float t, u, v;
            float3 ng;
            if (intersectRayTriangle(&t, &u, &v, &ng, 0.0f, hit->t, ro, rd, load3(tri.vs[0]), load3(tri.vs[1]),
                                     load3(tri.vs[2]))) {
                hit->t = t;
                hit->u = u;
                hit->v = v;
                hit->ng = ng;
                hit->tri = curr & ~kLeaf;
                if (any) break;
            }
            std::swap(curr, prev);
            continue;
        }
        const palz_internal_node& nd = s.nodes[curr];
        float l0, l1, r0, r1;
        slabs(ro, ird, load3(nd.aabbs[0].lower), load3(nd.aabbs[0].upper), hit->t, &l0, &l1);
        slabs(ro, ird, load3(nd.aabbs[1].lower), load3(nd.aabbs[1].upper), hit->t, &r0, &r1);
        const bool hitL = l0 <= l1, hitR = r0 <= r1;
        uint32_t near_node = nd.children[0], far_node = nd.children[1];
        int nHits = 0;
        if (hitL && hitR) {
            if (r0 < l0) std::swap(near_node, far_node);
            nHits = 2;
        } else if (hitL || hitR) {
            nHits = 1;
            near_node = hitR ? far_node : near_node;
        }
        uint32_t next;
        if (prev == nd.parent)
            next = 0 < nHits ? near_node : nd.parent;
        else if (prev == near_node)
            next = nHits == 2 ? far_node : nd.parent;
        else
            next = nd.parent;
        prev = curr;
        curr = next;
    }
}

// Work of the occlusion rays, counted when set (orc_occlusion_work): [0] rays,
// [1] child boxes tested, [2] triangles tested, then the same three per
// segment of the chain (light side first), 4 segments at most.
thread_local uint64_t* g_occl = nullptr;
thread_local int g_occl_segment = 0;
inline void count_occl(int what) {
    ++g_occl[what];
    if (g_occl_segment < 4) ++g_occl[4 + 3 * g_occl_segment + what];
}

// intersect_escape_link, any-hit (a leaf root is tested directly)
bool any_hit(const Scene& s, float3 ro, float3 rd, float3 ird, float t_max) {
    Hit hit;
    hit.t = t_max;
    float t, u, v;
    float3 ng;
    if (g_occl) count_occl(0);
    if (s.root & kLeaf) {
        if (g_occl) count_occl(2);
        const palz_triangle& tri = s.tris[s.root & ~kLeaf];
        return intersectRayTriangle(&t, &u, &v, &ng, 0.0f, hit.t, ro, rd, load3(tri.vs[0]), load3(tri.vs[1]),
                                    load3(tri.vs[2]));
    }
    uint32_t curr = s.root, prev = kInvalid;
    while (curr != kInvalid) {
        if (g_occl) count_occl((curr & kLeaf) ? 2 : 1);
        if (curr & kLeaf) {
            const palz_triangle& tri = s.tris[curr & ~kLeaf];
            if (intersectRayTriangle(&t, &u, &v, &ng, 0.0f, hit.t, ro, rd, load3(tri.vs[0]), load3(tri.vs[1]),
                                     load3(tri.vs[2])))
                return true;
            if (s.nodes[prev].children[0] == curr) {
                std::swap(curr, prev);
                continue;
            }
            uint32_t idx = s.nodes[prev].context;
            prev = curr;
            curr = idx != 0xFFFFFFFFu ? idx : kInvalid;
            continue;
        }
        const palz_internal_node& nd = s.nodes[curr];
        palz_aabb box;
        uint32_t hitChild, miss;
        if (prev == nd.parent) {
            box = nd.aabbs[0];
            hitChild = nd.children[0];
            miss = curr;
        } else {
            box = nd.aabbs[1];
            hitChild = nd.children[1];
            miss = nd.context != 0xFFFFFFFFu ? nd.context : kInvalid;
        }
        float a, b;
        slabs(ro, ird, load3(box.lower), load3(box.upper), hit.t, &a, &b);
        prev = curr;
        curr = a <= b ? hitChild : miss;
    }
    return false;
}

float rayOffsetScale(float3 p) {  // pk.h:43-49
    p = {std::fabs(p.x), std::fabs(p.y), std::fabs(p.z)};
    float maxElem = hfmax(hfmax(hfmax(p.x, p.y), p.z), 1.0f);
    return maxElem * 1.192092896e-07F * 128.0f;
}
bool occluded(const Scene& s, float3 from, float3 from_n, float3 to, float3 to_n) {  // pk.h:88-115
    if (dot(to - from, from_n) < 0.0f) from_n = -from_n;
    if (dot(from - to, to_n) < 0.0f) to_n = -to_n;
    float3 from_safe = from + from_n * rayOffsetScale(from);
    float3 to_safe = to + to_n * rayOffsetScale(to);
    float3 rd = to_safe - from_safe;
    return any_hit(s, from_safe, rd, invRd(rd), 1.0f);
}

float fresnel_exact_norm_free(float3 wi, float3 n, float eta) {  // pk.h:427-450
    float nn = dot(n, n);
    float wiwi = dot(wi, wi);
    float C = dot(n, wi);
    if (C < 0.0f) {
        C = -C;
        n = -n;
        eta = 1.0f / eta;
    }
    float K = nn * wiwi * (eta * eta - 1.0f) + C * C;
    if (K < 0.0f) return 1.0f;
    float G = std::sqrt(K);
    auto sqr = [](float x) { return x * x; };
    float GmC = G - C;
    float GpC = G + C;
    return 0.5f * sqr(GmC / GpC) * (1.0f + sqr((C * GpC - nn * wiwi) / (C * GmC + nn * wiwi)));
}
bool refraction_norm_free(float3* wo, float3 wi, float3 n, float eta) {  // pk.h:181-200
    float NoN = dot(n, n);
    float WIoN = dot(wi, n);
    if (WIoN < 0.0f) {
        WIoN = -WIoN;
        n = -n;
        eta = 1.0f / eta;
    }
    float WIoWI = dot(wi, wi);
    float k = NoN * WIoWI * (eta * eta - 1.0f) + WIoN * WIoN;
    if (k < 0.0f) return false;
    *wo = -wi * NoN + n * (WIoN - std::sqrt(k));
    return true;
}
float3 reflection(float3 wi, float3 n) { return n * dot(wi, n) * 2.0f / dot(n, n) - wi; }  // pk.h:202-205

float3 getVertex(const palz_triangle& tri, float u, float v) {  // pk.h:752-758
    float3 v0 = load3(tri.vs[0]);
    float3 e0 = load3(tri.vs[1]) - v0;
    float3 e1 = load3(tri.vs[2]) - v0;
    return v0 + e0 * u + e1 * v;
}

// contributableThroughput<K> (pk.h:760-842)
template <int K>
float contributableThroughput(const float* parameters, float3 p_beg, float3 p_end, const palz_triangle* tris,
                              const palz_triangle_attrib* attribs, uint32_t event, const Scene& s, float eta) {
    float3 vertices[K + 2];
    vertices[0] = p_beg;
    vertices[K + 1] = p_end;
    float3 shadingNormals[K + 2];
    for (int k = 0; k < K; k++) {
        float param_u = parameters[k * 2 + 0];
        float param_v = parameters[k * 2 + 1];
        if (param_u < 0.0f || param_v < 0.0f || 1.0f < param_u + param_v) return 0.0f;
        vertices[k + 1] = getVertex(tris[k], param_u, param_v);
        float3 n0 = load3(attribs[k].shading_normals[0]);
        float3 ne0 = load3(attribs[k].shading_normals[1]) - n0;
        float3 ne1 = load3(attribs[k].shading_normals[2]) - n0;
        shadingNormals[k + 1] = normalize(n0 + ne0 * param_u + ne1 * param_v);
    }
    shadingNormals[0] = normalize(vertices[1] - vertices[0]);
    shadingNormals[K + 1] = normalize(vertices[K] - vertices[K + 1]);
    float throughput = 1.0f;
    for (int k = 0; k < K; k++) {
        float3 wi = vertices[k] - vertices[k + 1];
        float3 wo = vertices[k + 2] - vertices[k + 1];
        float3 n = shadingNormals[k + 1];
        bool eT = !(0.0f < dot(wi, n) * dot(wo, n));
        if (is_T(event, k) != eT) return 0.0f;
        float3 v0 = load3(tris[k].vs[0]);
        float3 ng = cross(load3(tris[k].vs[1]) - v0, load3(tris[k].vs[2]) - v0);
        bool egT = !(0.0f < dot(wi, ng) * dot(wo, ng));
        if (eT != egT) return 0.0f;
        if (attribs[k].material == PALZ_MATERIAL_DIELECTRIC) {
            float reflectance = fresnel_exact_norm_free(wi, n, eta);
            throughput *= !eT ? reflectance : 1.0f - reflectance;
        }
    }
    for (int i = 0; i < K + 1; i++) {
        float3 d = vertices[i] - vertices[i + 1];
        if (dot(d, d) < 1.0e-5f) return 0.0f;
    }
    for (int i = 0; i < K + 1; i++) {
        g_occl_segment = i;
        if (occluded(s, vertices[i], shadingNormals[i], vertices[i + 1], shadingNormals[i + 1])) return 0.0f;
    }
    return throughput;
}

void GetOrthonormalBasis(float3 zaxis, float3* xaxis, float3* yaxis) {  // pk.h:453-459
    const float sign = std::copysign(1.0f, zaxis.z);
    const float a = -1.0f / (sign + zaxis.z);
    const float b = zaxis.x * zaxis.y * a;
    *xaxis = {1.0f + sign * zaxis.x * zaxis.x * a, sign * b, -sign * zaxis.x};
    *yaxis = {b, sign + zaxis.y * zaxis.y * a, -zaxis.y};
}
dval3<float> intersect_p_ray_plane(dval3<float> ro, dval3<float> rd, dval3<float> ng, dval3<float> v0) {
    auto t = dot(v0 - ro, ng) / dot(ng, rd);
    return ro + rd * t;
}

// dAdw (pk.h:844-930)
float dAdw(float3 ro, float3 rd, float3 p_end, const palz_triangle* tris, const palz_triangle_attrib* attribs,
           uint32_t event, int size, float eta) {
    rd = normalize(rd);
    float3 dAxis[2];
    for (int i = 0; i < 2; i++) {
        dval<float> differentials[2] = {dval<float>(0.0f), dval<float>(0.0f)};
        differentials[i].g = 1.0f;
        dval3<float> ro_j = mkd(ro);
        float3 T0, T1;
        GetOrthonormalBasis(rd, &T0, &T1);
        dval3<float> rd_j = mkd(rd) + mkd(T0) * differentials[0] + mkd(T1) * differentials[1];
        for (int j = 0; j < size; j++) {
            float3 v0 = load3(tris[j].vs[0]), v1 = load3(tris[j].vs[1]), v2 = load3(tris[j].vs[2]);
            float3 ng = cross(v1 - v0, v2 - v0);
            dval3<float> p = intersect_p_ray_plane(ro_j, rd_j, mkd(ng), mkd(v0));
            dval<float> v = dot(p - mkd(v1), mkd(cross(ng, v1 - v0)));
            dval<float> w = dot(p - mkd(v2), mkd(cross(ng, v2 - v1)));
            dval<float> u = dot(p - mkd(v0), mkd(cross(ng, v0 - v2)));
            dval<float> area = u + v + w;
            u = u / area;
            v = v / area;
            w = w / area;
            dval3<float> n = mkd(load3(attribs[j].shading_normals[0])) * w +
                             mkd(load3(attribs[j].shading_normals[1])) * u +
                             mkd(load3(attribs[j].shading_normals[2])) * v;
            dval3<float> wi = -rd_j;
            dval3<float> wo;
            if (is_T(event, j)) {
                if (dot(wi, n).v < 0.0f) {
                    wo = refraction_norm_free(wi, -n, 1.0f / eta);
                    if (wo.x.v == 0.0f && wo.y.v == 0.0f && wo.z.v == 0.0f) return 3.402823466e+38f;
                } else {
                    wo = refraction_norm_free(wi, n, eta);
                }
            } else {
                wo = reflection(wi, n);
            }
            ro_j = p;
            rd_j = wo;
        }
        float3 p_last = {ro_j.x.v, ro_j.y.v, ro_j.z.v};
        dval3<float> p_final = intersect_p_ray_plane(ro_j, rd_j, mkd(p_last - p_end), mkd(p_end));
        dAxis[i] = {p_final.x.g, p_final.y.g, p_final.z.g};
    }
    float len0 = length(dAxis[0]);
    float len1 = length(dAxis[1]);
    float sinTheta = length(cross(dAxis[0] / len0, dAxis[1] / len1));
    return len0 * len1 * hfmax(sinTheta, 0.01f);
}

struct SolveCtx {
    Scene s;
    const palz_first_diffuse* points;
    const float* lambdas;
    uint32_t n_lambdas;
    float3 p_light;
    float light_intensity, radiance_clamp, cauchy_a, cauchy_b;
};

// solveSpecularPath<K> (main_gpu.cu:418-486) for one (path, wavelength)
template <int K>
void solve_item(const SolveCtx& c, const palz_path& path, uint32_t l, palz_path_result* res) {
    palz_triangle tris[K];
    palz_triangle_attrib attribs[K];
    for (int k = 0; k < K; k++) {
        tris[k] = c.s.tris[path.tris[k]];
        attribs[k] = c.s.attribs[path.tris[k]];
    }
    float parameters[K * 2];
    for (int i = 0; i < K * 2; i++) parameters[i] = path.params[i];
    const palz_first_diffuse& q = c.points[path.pixel];
    float3 p = load3(q.p), n = load3(q.ng);
    float lambda = c.lambdas[l];
    float eta = c.cauchy_a + c.cauchy_b / (lambda * lambda);
    int iters = 0;
    bool converged = solveConstraints<K, float>(parameters, c.p_light, p, tris, attribs, eta, path.event, 32, 1.0e-10f,
                                                1, &iters);
    std::memset(res, 0, sizeof(*res));
    res->iterations = iters;
    if (converged) {
        res->status = PALZ_PATH_CONVERGED;
        float throughput =
            contributableThroughput<K>(parameters, c.p_light, p, tris, attribs, path.event, c.s, eta);
        if (0.0f < throughput) {
            float dAdwValue = dAdw(c.p_light, getVertex(tris[0], parameters[0], parameters[1]) - c.p_light, p, tris,
                                   attribs, path.event, K, eta);
            float3 last = getVertex(tris[K - 1], parameters[2 * (K - 1)], parameters[2 * (K - 1) + 1]);
            float contrib = throughput * c.light_intensity / dAdwValue * hfmax(dot(normalize(last - p), n), 0.0f);
            contrib = hfmin(contrib, c.radiance_clamp);
            res->status = PALZ_PATH_CONTRIBUTES;
            res->contribution = contrib;
            res->footprint = dAdwValue;
        }
    }
    for (int i = 0; i < 8; i++) res->params[i] = i < 2 * K ? parameters[i] : 0.0f;
}

void solve_any(const SolveCtx& c, const palz_path& path, uint32_t l, palz_path_result* res) {
    switch (path.event >> 24) {
        case 1: solve_item<1>(c, path, l, res); break;
        case 2: solve_item<2>(c, path, l, res); break;
        case 3: solve_item<3>(c, path, l, res); break;
        case 4: solve_item<4>(c, path, l, res); break;
        default: std::memset(res, 0, sizeof(*res));
    }
}

SolveCtx make_ctx(const palz_scene_desc* sd) {
    SolveCtx c;
    c.s = {sd->tris, sd->attribs, sd->nodes, *sd->root};
    c.points = sd->points;
    c.lambdas = sd->lambdas;
    c.n_lambdas = sd->n_lambdas;
    c.p_light = load3(sd->p_light);
    c.light_intensity = sd->light_intensity;
    c.radiance_clamp = sd->radiance_clamp;
    c.cauchy_a = sd->cauchy_a;
    c.cauchy_b = sd->cauchy_b;
    return c;
}

}  // namespace

extern "C" {

void orc_sobol2d(uint32_t index, uint32_t p0, uint32_t p1, uint32_t p2, float* out) {
    shuffled_scrambled_sobol_2d(out, out + 1, index, p0, p1, p2);
}

// photonTrace (main_gpu.cu:56-211), all photon slots serially.  Host pointers;
// the scene's root is a host value.  Returns the deepest traversal stack use.
int orc_photon_trace(const palz_triangle* tris, const palz_triangle_attrib* attribs, uint32_t n_tris,
                     const palz_internal_node* nodes, uint32_t root, const palz_photon_params* pp,
                     palz_photon_record* out) {
    Scene s{tris, attribs, nodes, root};
    int max_sp = 0;
    const uint32_t Pn = pp->photons_per_triangle;
    const float3 p_light = load3(pp->p_light);
    for (uint32_t iTri = 0; iTri < n_tris; iTri++) {
        for (uint32_t j = 0; j < Pn; j++) {
            palz_photon_record& rec = out[(size_t)iTri * Pn + j];
            std::memset(&rec, 0, sizeof(rec));
            rec.photon = iTri * Pn + j;
            for (int k = 0; k < 4; k++) rec.tris[k] = -1;
            if (iTri % g_tri_stride) continue;
            if (attribs[iTri].material == PALZ_MATERIAL_DIFFUSE) continue;
            const palz_triangle& tri = tris[iTri];
            float3 v0 = load3(tri.vs[0]);
            float3 ng = cross(load3(tri.vs[1]) - v0, load3(tri.vs[2]) - v0);
            float3 ns = load3(attribs[iTri].shading_normals[0]);
            ng = ng * dot(ng, ns);
            if (dot(ng, p_light - v0) < 0.0f) continue;
            float px, py;
            shuffled_scrambled_sobol_2d(&px, &py, j, (uint32_t)pp->iteration, iTri, 789);
            square2triangle(&px, &py);
            float ex, ey;
            shuffled_scrambled_sobol_2d(&ex, &ey, j, (uint32_t)pp->iteration, iTri, 178);
            float eta = pp->eta_min + ex * (pp->eta_max - pp->eta_min);
            float3 e0 = load3(tri.vs[1]) - v0, e1 = load3(tri.vs[2]) - v0;
            float3 p = v0 + e0 * px + e1 * py;
            float3 ro = p_light, rd = p - p_light;
            uint32_t events = 0;
            int size = 0;
            bool admissiblePath = false;
            int triIndices[4];
            float parameters[8];
            float3 p_final = {0, 0, 0};
            float throughtput = 1.0f;
            if (g_work) ++g_work[4];
            for (int d = 0; d < 4 + 1; d++) {
                Hit hit;
                intersect_stackfull(&hit, s, ro, rd, invRd(rd), &max_sp);
                if (hit.t == 3.402823466e+38F) break;
                if (g_work) ++g_work[2];
                const palz_triangle_attrib& attrib = attribs[hit.tri];
                float3 p_hit = ro + rd * hit.t;
                float3 a0 = load3(attrib.shading_normals[0]);
                float3 nsh = a0 + (load3(attrib.shading_normals[1]) - a0) * hit.u +
                             (load3(attrib.shading_normals[2]) - a0) * hit.v;
                if (attrib.material == PALZ_MATERIAL_DIFFUSE) {
                    if (d == 0) break;
                    admissiblePath = true;
                    p_final = p_hit;
                    break;
                }
                if (d == 4) break;
                float3 wi = -rd;
                triIndices[d] = (int)hit.tri;
                parameters[d * 2] = hit.u;
                parameters[d * 2 + 1] = hit.v;
                float rx, ry;
                shuffled_scrambled_sobol_2d(&rx, &ry, j, (uint32_t)d, iTri, (uint32_t)pp->iteration);
                float reflectance = fresnel_exact_norm_free(wi, nsh, eta);
                bool eT = false;
                if (attrib.material == PALZ_MATERIAL_DIELECTRIC) eT = !(rx < reflectance);
                if (eT) events |= 1u << size;
                size++;
                float sign = 1.0f;
                float3 wo;
                if (!eT) {
                    wo = reflection(wi, nsh);
                    if (attrib.material == PALZ_MATERIAL_DIELECTRIC) throughtput *= reflectance;
                } else {
                    if (!refraction_norm_free(&wo, wi, nsh, eta)) break;
                    sign = -1.0f;
                    throughtput *= 1.0f - reflectance;
                }
                if (0.0f < dot(hit.ng, wi) * dot(hit.ng, wo) * sign) {
                    float3 offset_dir = normalize(hit.ng * dot(hit.ng, wo));
                    ro = p_hit + offset_dir * rayOffsetScale(p_hit);
                    rd = wo;
                    continue;
                }
                break;
            }
            if (admissiblePath && pp->min_throughput < throughtput) {
                if (g_work) ++g_work[3];
                rec.event = ((uint32_t)size << 24) | events;
                for (int k = 0; k < size; k++) {
                    rec.tris[k] = triIndices[k];
                    rec.params[2 * k] = parameters[2 * k];
                    rec.params[2 * k + 1] = parameters[2 * k + 1];
                }
                rec.p_final = {p_final.x, p_final.y, p_final.z};
                rec.throughput = throughtput;
            }
        }
    }
    return max_sp;
}

// photonTrace's algorithmic work over the photons of every `stride`-th
// triangle (the reference's stackfull closest-hit traversal, node by node):
// counts = {internal nodes visited, triangles tested, attribute records read,
// tuples stored, photons traced}.
void orc_photon_trace_work(const palz_triangle* tris, const palz_triangle_attrib* attribs, uint32_t n_tris,
                           const palz_internal_node* nodes, uint32_t root, const palz_photon_params* pp,
                           uint32_t stride, uint64_t* counts) {
    std::vector<palz_photon_record> out((size_t)n_tris * pp->photons_per_triangle);
    for (int k = 0; k < 5; ++k) counts[k] = 0;
    g_work = counts;
    g_tri_stride = stride ? stride : 1;
    orc_photon_trace(tris, attribs, n_tris, nodes, root, pp, out.data());
    g_work = nullptr;
    g_tri_stride = 1;
}

// The tuple set and the path list, by direct set construction (std::map):
// for each valid photon in id order, insert (cell, event, tris) for its 2x2x2
// nearest cells (pk.h:302-313) if absent; each receiver point then lists the
// entries of its own cell (pk.h:352-373) whose event matches.  Output paths
// are in (point, event, tris) order; returns the number written (<= cap),
// *total receives the full count.
uint32_t orc_lookup_paths(const palz_photon_record* rec, uint32_t n_rec, float step, const palz_first_diffuse* pts,
                          uint32_t n_pts, uint32_t event, uint32_t scene, palz_path* out, uint32_t cap,
                          uint32_t* total) {
    using Key = std::tuple<int, int, int, uint32_t, int, int, int, int>;
    std::map<Key, uint32_t> table;  // -> photon slot of the first (smallest id) photon
    for (uint32_t i = 0; i < n_rec; i++) {
        const palz_photon_record& r = rec[i];
        if (r.event == 0) continue;
        float fx = r.p_final.x / step, fy = r.p_final.y / step, fz = r.p_final.z / step;
        int x = (int)std::floor(fx), y = (int)std::floor(fy), z = (int)std::floor(fz);
        int dx = fx - x < 0.5f ? -1 : 1, dy = fy - y < 0.5f ? -1 : 1, dz = fz - z < 0.5f ? -1 : 1;
        int size = (int)(r.event >> 24);
        int t[4] = {-1, -1, -1, -1};
        for (int k = 0; k < size; k++) t[k] = r.tris[k];
        for (int iz = 0; iz < 2; iz++)
            for (int iy = 0; iy < 2; iy++)
                for (int ix = 0; ix < 2; ix++) {
                    Key key{x + ix * dx, y + iy * dy, z + iz * dz, r.event, t[0], t[1], t[2], t[3]};
                    table.emplace(key, i);
                }
    }
    uint32_t n = 0;
    for (uint32_t pi = 0; pi < n_pts; pi++) {
        const palz_first_diffuse& q = pts[pi];
        if (q.R.x == 0.0f && q.R.y == 0.0f && q.R.z == 0.0f) continue;
        int x = (int)std::floor(q.p.x / step), y = (int)std::floor(q.p.y / step), z = (int)std::floor(q.p.z / step);
        auto it = table.lower_bound(Key{x, y, z, 0u, INT32_MIN, INT32_MIN, INT32_MIN, INT32_MIN});
        for (; it != table.end(); ++it) {
            const Key& k = it->first;
            if (std::get<0>(k) != x || std::get<1>(k) != y || std::get<2>(k) != z) break;
            if (event != 0 && std::get<3>(k) != event) continue;
            if (n < cap) {
                const palz_photon_record& r = rec[it->second];
                palz_path& pa = out[n];
                pa.scene = scene;
                pa.pixel = pi;
                pa.event = r.event;
                pa.photon = r.photon;
                for (int t = 0; t < 4; t++) pa.tris[t] = r.tris[t];
                for (int t = 0; t < 8; t++) pa.params[t] = r.params[t];
            }
            n++;
        }
    }
    *total = n;
    return std::min(n, cap);
}

// Plain FP32 solve of every (path, wavelength) with `threads` host threads
// (ParallelFor, main.cpp:1745); adds contributions into sd->image (host) in
// item order, or as 32.32 fixed point into sd->image_fixed when that is set,
// and writes results if not NULL.
// The 32.32 fixed-point word of one contribution: c * 2^32 truncated, 0 for a
// NaN or non-positive c, saturated at the largest float below 2^64 for
// c >= 2^32 (where the conversion would be undefined).
uint64_t orc_fixed_word(float c) {
    if (!(c > 0.0f)) return 0;
    const float s = c < 4294967296.0f ? c * 4294967296.0f : 18446742974197923840.0f;
    return (uint64_t)s;
}

void orc_solve_paths(const palz_scene_desc* sd, const palz_path* paths, uint32_t n_paths, palz_path_result* results,
                     int threads) {
    SolveCtx c = make_ctx(sd);
    const uint64_t items = (uint64_t)n_paths * c.n_lambdas;
    std::vector<palz_path_result> local(results ? 0 : items);
    palz_path_result* res = results ? results : local.data();
    threads = std::max(1, threads);
    auto work = [&](uint64_t b, uint64_t e) {
        for (uint64_t it = b; it < e; it++) solve_any(c, paths[it / c.n_lambdas], (uint32_t)(it % c.n_lambdas), &res[it]);
    };
    if (threads == 1) {
        work(0, items);
    } else {
        std::vector<std::thread> pool;
        for (int t = 0; t < threads; t++)
            pool.emplace_back(work, items * t / threads, items * (t + 1) / threads);
        for (auto& th : pool) th.join();
    }
    for (uint64_t it = 0; it < items; it++) {
        if (res[it].status != PALZ_PATH_CONTRIBUTES) continue;
        const size_t idx = (size_t)(it % c.n_lambdas) * sd->n_points + paths[it / c.n_lambdas].pixel;
        if (sd->image_fixed)  // unsigned 32.32 fixed point, contribution truncated
            sd->image_fixed[idx] += orc_fixed_word(res[it].contribution);
        else
            sd->image[idx] += res[it].contribution;
    }
}

// The occlusion rays' work over every (path, wavelength) item of the paths,
// solved serially: out[0..3] = {rays, child boxes tested, triangles tested,
// items}, out[4 + 3 i .. 6 + 3 i] = {rays, boxes, triangles} of chain segment
// i (i = 0 from the light, 4 segments at most).
void orc_occlusion_work(const palz_scene_desc* sd, const palz_path* paths, uint32_t n_paths, uint64_t* out) {
    SolveCtx c = make_ctx(sd);
    for (int i = 0; i < 16; ++i) out[i] = 0;
    g_occl = out;
    palz_path_result r;
    for (uint64_t it = 0; it < (uint64_t)n_paths * c.n_lambdas; it++)
        solve_any(c, paths[it / c.n_lambdas], (uint32_t)(it % c.n_lambdas), &r);
    g_occl = nullptr;
    out[3] = (uint64_t)n_paths * c.n_lambdas;
}

// FP64 Newton of one chain from `params` (FP32 start), to cost < tol64, with
// the reference's iteration structure in double precision.  Writes the
// solution and its K vertices (x,y,z per vertex).  Returns 1 if converged.
int orc_solve_f64(const palz_triangle* tris_all, const palz_triangle_attrib* attribs_all, const palz_path* path,
                  const float* p_light, const float* p_end, double eta, double tol64, int max_iter, double* params_out,
                  double* vertices_out) {
    const int K = (int)(path->event >> 24);
    palz_triangle tris[4];
    palz_triangle_attrib attribs[4];
    for (int k = 0; k < K; k++) {
        tris[k] = tris_all[path->tris[k]];
        attribs[k] = attribs_all[path->tris[k]];
    }
    double prm[8];
    for (int i = 0; i < 2 * K; i++) prm[i] = params_out[i];
    double3 pb = {p_light[0], p_light[1], p_light[2]}, pe = {p_end[0], p_end[1], p_end[2]};
    int iters = 0;
    bool ok = false;
    switch (K) {
        case 1: ok = solveConstraints<1, double>(prm, pb, pe, tris, attribs, eta, path->event, max_iter, tol64, max_iter, &iters); break;
        case 2: ok = solveConstraints<2, double>(prm, pb, pe, tris, attribs, eta, path->event, max_iter, tol64, max_iter, &iters); break;
        case 3: ok = solveConstraints<3, double>(prm, pb, pe, tris, attribs, eta, path->event, max_iter, tol64, max_iter, &iters); break;
        case 4: ok = solveConstraints<4, double>(prm, pb, pe, tris, attribs, eta, path->event, max_iter, tol64, max_iter, &iters); break;
        default: return 0;
    }
    for (int i = 0; i < 2 * K; i++) params_out[i] = prm[i];
    for (int k = 0; k < K; k++) {
        double3 v0 = {tris[k].vs[0].x, tris[k].vs[0].y, tris[k].vs[0].z};
        double3 v1 = {tris[k].vs[1].x, tris[k].vs[1].y, tris[k].vs[1].z};
        double3 v2 = {tris[k].vs[2].x, tris[k].vs[2].y, tris[k].vs[2].z};
        double3 x = v0 + (v1 - v0) * prm[2 * k] + (v2 - v0) * prm[2 * k + 1];
        vertices_out[3 * k + 0] = x.x;
        vertices_out[3 * k + 1] = x.y;
        vertices_out[3 * k + 2] = x.z;
    }
    return ok ? 1 : 0;
}

// The batched solver's refinement step (refineStep above) of one chain from
// its FP32 solution params32; writes 2K doubles.  Returns 0 for K outside 1..3.
int orc_refine_step(const palz_triangle* tris_all, const palz_triangle_attrib* attribs_all, const palz_path* path,
                    const float* p_light, const float* p_end, float eta, const float* params32, double* out) {
    const int K = (int)(path->event >> 24);
    palz_triangle tris[3];
    palz_triangle_attrib attribs[3];
    if (K < 1 || K > 3) return 0;
    for (int k = 0; k < K; k++) {
        tris[k] = tris_all[path->tris[k]];
        attribs[k] = attribs_all[path->tris[k]];
    }
    const float3 pb = {p_light[0], p_light[1], p_light[2]}, pe = {p_end[0], p_end[1], p_end[2]};
    switch (K) {
        case 1: refineStep<1>(params32, pb, pe, tris, attribs, eta, path->event, out); break;
        case 2: refineStep<2>(params32, pb, pe, tris, attribs, eta, path->event, out); break;
        default: refineStep<3>(params32, pb, pe, tris, attribs, eta, path->event, out); break;
    }
    return 1;
}

// Arithmetic of one solveConstraints iteration (pk.h:477-590) on this chain,
// in the reference's form (2K single-tangent passes): out[0] = the residual
// and Jacobian evaluation with the cost, out[1] = the QR step, the update and
// the triangle test that follow it.
int orc_count_flops(const palz_triangle* tris_all, const palz_triangle_attrib* attribs_all, const palz_path* path,
                    const float* p_light, const float* p_end, float eta, uint64_t* out) {
    const int K = (int)(path->event >> 24);
    palz_triangle tris[4];
    palz_triangle_attrib attribs[4];
    for (int k = 0; k < K; k++) {
        tris[k] = tris_all[path->tris[k]];
        attribs[k] = attribs_all[path->tris[k]];
    }
    vec3<Flop> pb = {p_light[0], p_light[1], p_light[2]}, pe = {p_end[0], p_end[1], p_end[2]};
    uint64_t counts[2];
    for (int pass = 0; pass < 2; ++pass) {
        Flop prm[8];
        for (int i = 0; i < 2 * K; i++) prm[i] = path->params[i];
        // pass 0: the cost test succeeds (evaluation only); pass 1: it fails
        const Flop tol = pass == 0 ? Flop(1e300) : Flop(-1.0);
        int iters = 0;
        Flop::n = 0;
        switch (K) {
            case 1: solveConstraints<1, Flop>(prm, pb, pe, tris, attribs, Flop(eta), path->event, 1, tol, 0, &iters); break;
            case 2: solveConstraints<2, Flop>(prm, pb, pe, tris, attribs, Flop(eta), path->event, 1, tol, 0, &iters); break;
            case 3: solveConstraints<3, Flop>(prm, pb, pe, tris, attribs, Flop(eta), path->event, 1, tol, 0, &iters); break;
            case 4: solveConstraints<4, Flop>(prm, pb, pe, tris, attribs, Flop(eta), path->event, 1, tol, 0, &iters); break;
            default: return 0;
        }
        counts[pass] = Flop::n;
    }
    out[0] = counts[0];
    out[1] = counts[1] - counts[0];
    return 1;
}

int orc_max_stack_depth(const palz_triangle* tris, const palz_internal_node* nodes, uint32_t root, const float* ro,
                        const float* rd) {
    Scene s{tris, nullptr, nodes, root};
    Hit hit;
    int m = 0;
    float3 o = {ro[0], ro[1], ro[2]}, d = {rd[0], rd[1], rd[2]};
    intersect_stackfull(&hit, s, o, d, invRd(d), &m);
    return m;
}

// ---- solvePrimary (main_gpu.cu:213-371) and what it needs ----------------

// This library's FP32 exp and log (the reference's NATIVE_EXPF / NATIVE_LOGF
// are the hardware's fast forms): x = k ln2 + r with ln 2 in two parts, e^r
// by its Taylor polynomial to degree 7, times 2^k in two factors, 0 below -87;
// log m for m in [sqrt(1/2), sqrt(2)) as 2 atanh((m - 1) / (m + 1)) by the odd
// series to s^9.
static float pow2i(int e) {
    uint32_t b = (uint32_t)(e + 127) << 23;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}
float orc_exp_f32(float x) {
    if (!(x == x)) return x;
    if (x > 88.72283f) return INFINITY;
    if (x < -87.0f) return 0.0f;
    const float kf = std::floor(x * 1.44269504f + 0.5f);
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
float orc_log_f32(float x) {
    if (!(x > 0.0f)) return x == 0.0f ? -INFINITY : NAN;
    if (x == INFINITY) return x;
    uint32_t b;
    std::memcpy(&b, &x, 4);
    int e = (int)(b >> 23) - 127;
    if ((b >> 23) == 0u) {
        const float y = x * 8388608.0f;
        std::memcpy(&b, &y, 4);
        e = (int)(b >> 23) - 127 - 23;
    }
    const uint32_t mb = (b & 0x7FFFFFu) | 0x3F800000u;
    float m;
    std::memcpy(&m, &mb, 4);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e += 1;
    }
    const float sv = (m - 1.0f) / (m + 1.0f);
    const float s2 = sv * sv;
    float q = 1.11111111e-1f;
    q = q * s2 + 1.42857143e-1f;
    q = q * s2 + 2.0e-1f;
    q = q * s2 + 3.33333333e-1f;
    q = q * s2 + 1.0f;
    const float ef = (float)e;
    return ef * 0.693145751953125f + (2.0f * sv * q + ef * 1.428606765330187e-06f);
}

namespace {
float asym_gauss(float x, float mean, float sigma, float a) {  // pk.h:962-968
    float denom = sigma + a * (x - mean);
    if (denom < 1.0e-15f) denom = 1.0e-15f;
    if (sigma * 2.0f < denom) denom = sigma * 2.0f;
    float k = (x - mean) / denom;
    return orc_exp_f32(-k * k);
}
float cmfX(float x) {  // pk.h:969-973
    float a = 0.42f * asym_gauss(x, 445.5849609375f, 31.146467208862305f, 0.06435633450746536f);
    float b = 1.16f * asym_gauss(x, 594.5570068359375f, 48.602108001708984f, -0.04772702232003212f);
    return (a + b - a * b * 42.559776306152344f) * 1.0008530432426956f;
}
float cmfY(float x) {  // pk.h:974-976
    return 1.0f * asym_gauss(x, 556.8383178710938f, 66.54190826416016f, -0.026492968201637268f) * 1.004315086937574f;
}
float cmfZ(float x) {  // pk.h:977-979
    return 2.146832f * asym_gauss(x, 445.9251708984375f, 30.91781997680664f, 0.08379141241312027f) *
           0.9975112815948937f;
}
float logisticPdf(float x, float sv) {  // pk.h:981-985 (double denominator, as the 1.0 literals make it)
    float k = orc_exp_f32(-std::fabs(x) / sv);
    return (float)(k / ((1.0 + k) * (1.0 + k) * sv));
}
float logisticCdf(float x, float sv) { return 1.0f / (1.0f + orc_exp_f32(-x / sv)); }  // pk.h:986-989
float inverseLogisticCdf(float u, float sv) {                                          // pk.h:990-995
    if (0.99999994f < u) u = 0.99999994f;
    if (u < 1.175494351e-38f) u = 1.175494351e-38f;
    return -sv * orc_log_f32(1.0f / u - 1.0f);
}
float cmfYPdf(float x) {  // pk.h:996-1002
    float sx = x - 554.270751953125f;
    float sv = 26.879621505737305f, a = -164.270751953125f, b = 275.729248046875f;
    return logisticPdf(sx, sv) / (logisticCdf(b, sv) - logisticCdf(a, sv));
}
float cmfYSample(float u) {  // pk.h:1004-1010
    float sv = 26.879621505737305f, a = -164.270751953125f, b = 275.729248046875f;
    float Pa = logisticCdf(a, sv), Pb = logisticCdf(b, sv);
    return inverseLogisticCdf(Pa + (Pb - Pa) * u, sv) + 554.270751953125f;
}
float3 xyz2srgblinear(float3 c) {  // pk.h:1014-1022
    return {3.2406f * c.x + -1.5372f * c.y + -0.4986f * c.z, -0.9689f * c.x + 1.8758f * c.y + 0.0415f * c.z,
            0.0557f * c.x + -0.204f * c.y + 1.057f * c.z};
}
float sphIntersectNear(float3 ro, float3 rd, float3 ce, float ra) {  // pk.h:52-65 (.x)
    float len2rd = dot(rd, rd);
    float3 oc = ro - ce;
    float b = dot(oc, rd) / len2rd;
    float c = dot(oc, oc) - ra * ra;
    float h = b * b - c / len2rd;
    if (h < 0.0f) return -1.0f;
    h = std::sqrt(h);
    return -b - h;
}
void solveLens(float3* ro, float3* rd, float3 camOrigin, float3 camForward, float distance, float thickness,
               float R, float eta) {  // pk.h:1042-1073
    float3 lensOrigin = camOrigin + camForward * distance;
    float3 center1 = lensOrigin + camForward * (R - thickness * 0.5f);
    float3 center2 = lensOrigin - camForward * (R - thickness * 0.5f);
    float t1 = sphIntersectNear(*ro, *rd, center1, R);
    if (t1 < 0.0f) return;
    float3 hit1 = *ro + *rd * t1;
    float3 wo;
    refraction_norm_free(&wo, -*rd, hit1 - center1, eta);
    *rd = wo;
    *ro = hit1;
    float t2 = sphIntersectNear(*ro, *rd, center2, R);
    float3 hit2 = *ro + *rd * t2;
    if (!refraction_norm_free(&wo, -*rd, hit2 - center2, eta)) return;
    *rd = wo;
    *ro = hit2;
}
float3 samplePoint(const uint8_t* tex, uint32_t w, uint32_t h, float u, float v) {  // pk.h:1081-1094
    u = u - std::floor(u);
    v = v - std::floor(v);
    float fx = u * (float)w, fy = v * (float)h;
    int ix = (int)(fx < (float)w - 1.0f ? fx : (float)w - 1.0f);
    int iy = (int)(fy < (float)h - 1.0f ? fy : (float)h - 1.0f);
    uint32_t idx = (uint32_t)ix + w * (uint32_t)iy;
    float c[3];
    for (int k = 0; k < 3; k++) {
        float x = (float)tex[idx * 4u + k] / 255.0f;
        c[k] = x == 0.0f ? 0.0f : orc_exp_f32(2.2f * orc_log_f32(x));  // powf(x, 2.2)
    }
    return {c[0], c[1], c[2]};
}
}  // namespace

// solvePrimary (main_gpu.cu:213-371) for every pixel, serially; host
// pointers, host root, host texture (or NULL).
void orc_solve_primary(const palz_triangle* tris, const palz_triangle_attrib* attribs, const palz_internal_node* nodes,
                       uint32_t root, const palz_camera* cam, const palz_primary_params* pp, const uint8_t* texture,
                       palz_first_diffuse* out, float* accum) {
    Scene sc{tris, attribs, nodes, root};
    const float3 origin = load3(cam->origin), right = load3(cam->right), up = load3(cam->up);
    const float3 forward = normalize(cross(up, right));
    for (uint32_t yi = 0; yi < pp->height; yi++) {
        for (uint32_t xi = 0; xi < pp->width; xi++) {
            const uint32_t pixel = xi + yi * pp->width;
            const uint32_t it = (uint32_t)pp->iteration;
            uint32_t dimLevel = 0;
            float jx, jy;
            shuffled_scrambled_sobol_2d(&jx, &jy, it, xi, yi, dimLevel++);
            float u = ((float)xi + jx) / (float)pp->width, v = ((float)yi + jy) / (float)pp->height;
            float3 to = origin + forward + (-right + (right - -right) * u) + (up + (-up - up) * v);
            float3 ro = origin, rd = to - origin;
            float lx, ly;
            shuffled_scrambled_sobol_2d(&lx, &ly, it, xi, yi, dimLevel++);
            float lambda = cmfYSample(lx);
            float p_lambda = cmfYPdf(lambda);
            float eta = pp->cauchy_a + pp->cauchy_b / (lambda * lambda);
            if (pp->use_lens) solveLens(&ro, &rd, origin, forward, pp->lens_distance, pp->lens_thickness, pp->lens_R, eta);
            bool hasDiffuseHit = false;
            Hit hit_last;
            int max_sp = 0;
            for (int d = 0; d < 32; d++) {
                Hit hit;
                intersect_stackfull(&hit, sc, ro, rd, invRd(rd), &max_sp);
                if (hit.t == 3.402823466e+38F) {
                    hit_last = hit;
                    break;
                }
                const palz_triangle_attrib& attrib = attribs[hit.tri];
                float rx, ry;
                shuffled_scrambled_sobol_2d(&rx, &ry, it, xi, yi, dimLevel++);
                if (attrib.material == PALZ_MATERIAL_DIFFUSE) {
                    hasDiffuseHit = true;
                    hit_last = hit;
                    break;
                }
                float3 p_hit = ro + rd * hit.t;
                float3 a0 = load3(attrib.shading_normals[0]);
                float3 ns = a0 + (load3(attrib.shading_normals[1]) - a0) * hit.u +
                            (load3(attrib.shading_normals[2]) - a0) * hit.v;
                float3 wi = -rd;
                bool eT = false;
                if (attrib.material == PALZ_MATERIAL_DIELECTRIC) eT = !(rx < fresnel_exact_norm_free(wi, ns, eta));
                float sign = 1.0f;
                float3 wo;
                if (!eT) {
                    wo = reflection(wi, ns);
                } else {
                    if (!refraction_norm_free(&wo, wi, ns, eta)) break;
                    sign = -1.0f;
                }
                if (0.0f < dot(hit.ng, wi) * dot(hit.ng, wo) * sign) {
                    float3 offset_dir = normalize(hit.ng * dot(hit.ng, wo));
                    ro = p_hit + offset_dir * rayOffsetScale(p_hit);
                    rd = wo;
                    continue;
                }
                break;
            }
            float* acc = accum ? accum + 4 * pixel : nullptr;
            if (!hasDiffuseHit) {
                out[pixel].R = {0.0f, 0.0f, 0.0f};
                if (acc) acc[3] += 1.0f;
                continue;
            }
            float3 p = ro + rd * hit_last.t;
            float3 n = normalize(hit_last.ng);
            if (0.0f < dot(n, rd)) n = -n;
            float3 p_light = load3(pp->p_light);
            float3 toLight = p_light - p;
            float d2 = dot(toLight, toLight);
            float3 refl = texture ? samplePoint(texture, pp->texture_width, pp->texture_height, p.x, p.z)
                                  : load3(pp->albedo);
            float3 L = {0.0f, 0.0f, 0.0f};
            if (!occluded(sc, p, n, p_light, {0.0f, 0.0f, 0.0f})) {
                float cosv = dot(normalize(toLight), n);
                float contrib = pp->light_intensity / d2 * (cosv > 0.0f ? cosv : 0.0f);
                contrib = contrib < pp->radiance_clamp ? contrib : pp->radiance_clamp;
                float3 srgb = xyz2srgblinear(
                    {cmfX(lambda) / 118.51810464018001f, cmfY(lambda) / 118.51810464018001f,
                     cmfZ(lambda) / 118.51810464018001f});
                float3 brdf = refl / 3.14159265358979323846f;
                L = float3{brdf.x * srgb.x, brdf.y * srgb.y, brdf.z * srgb.z} * contrib / p_lambda;
            }
            out[pixel].p = {p.x, p.y, p.z};
            out[pixel].ng = {n.x, n.y, n.z};
            out[pixel].R = {refl.x, refl.y, refl.z};
            out[pixel].lambda = lambda;
            if (acc) {
                acc[0] += L.x;
                acc[1] += L.y;
                acc[2] += L.z;
                acc[3] += 1.0f;
            }
        }
    }
}

float orc_cmf_y_sample(float u) { return cmfYSample(u); }
float orc_cmf_y_pdf(float x) { return cmfYPdf(x); }

// ---- single-function entry points for the pins of the optics and traversal --

// fresnel_exact_norm_free (pk.h:427-450)
float orc_fresnel(const float* wi, const float* n, float eta) {
    return fresnel_exact_norm_free({wi[0], wi[1], wi[2]}, {n[0], n[1], n[2]}, eta);
}
// refraction_norm_free (pk.h:181-200); returns 0 on total internal reflection
int orc_refraction(const float* wi, const float* n, float eta, float* wo) {
    float3 w;
    if (!refraction_norm_free(&w, {wi[0], wi[1], wi[2]}, {n[0], n[1], n[2]}, eta)) return 0;
    wo[0] = w.x, wo[1] = w.y, wo[2] = w.z;
    return 1;
}
// saka::reflection (op 0) / saka::refraction_norm_free (op 1) on one-tangent
// duals laid out {v.x, v.y, v.z, g.x, g.y, g.z} (saka.h:192-208)
void orc_dual_optics(int op, const float* wi, const float* n, float eta, float* out) {
    auto ld = [](const float* p) {
        dval3<float> r;
        r.x.v = p[0], r.y.v = p[1], r.z.v = p[2];
        r.x.g = p[3], r.y.g = p[4], r.z.g = p[5];
        return r;
    };
    const dval3<float> r = op == 0 ? reflection(ld(wi), ld(n)) : refraction_norm_free(ld(wi), ld(n), eta);
    out[0] = r.x.v, out[1] = r.y.v, out[2] = r.z.v;
    out[3] = r.x.g, out[4] = r.y.g, out[5] = r.z.g;
}
// intersectRayTriangle (minimum_lbvh.h:372-422) on one triangle over
// [t_min, t_max]; out = {t, u, v, ng.x, ng.y, ng.z} when it returns 1
int orc_ray_triangle(const palz_triangle* tri, const float* ro, const float* rd, float t_min, float t_max,
                     float* out) {
    float t, u, v;
    float3 ng;
    if (!intersectRayTriangle(&t, &u, &v, &ng, t_min, t_max, {ro[0], ro[1], ro[2]}, {rd[0], rd[1], rd[2]},
                              load3(tri->vs[0]), load3(tri->vs[1]), load3(tri->vs[2])))
        return 0;
    out[0] = t, out[1] = u, out[2] = v, out[3] = ng.x, out[4] = ng.y, out[5] = ng.z;
    return 1;
}
// intersect_stackfull (minimum_lbvh.h:927-988) for n rays (ro, rd: 3 floats
// each): hit_t (FLT_MAX on a miss) and hit_tri (0xFFFFFFFF on a miss)
void orc_closest_hits(const palz_triangle* tris, const palz_internal_node* nodes, uint32_t root, const float* ro,
                      const float* rd, uint32_t n, float* hit_t, uint32_t* hit_tri) {
    const Scene s{tris, nullptr, nodes, root};
    for (uint32_t i = 0; i < n; ++i) {
        Hit hit;
        const float3 o = {ro[3 * i], ro[3 * i + 1], ro[3 * i + 2]}, d = {rd[3 * i], rd[3 * i + 1], rd[3 * i + 2]};
        intersect_stackfull(&hit, s, o, d, invRd(d), nullptr);
        hit_t[i] = hit.t;
        hit_tri[i] = hit.tri;
    }
}
// The ray queries of palz_intersect_rays: query 0 intersect_stackfull,
// 1 intersect_stackfree (closest), 2 intersect_stackfree (any), over
// [0, t_max]; one palz_hit per ray.
void orc_intersect_rays(const palz_triangle* tris, const palz_internal_node* nodes, uint32_t root, const float* ro,
                        const float* rd, uint32_t n, int query, float t_max, palz_hit* out) {
    const Scene s{tris, nullptr, nodes, root};
    for (uint32_t i = 0; i < n; ++i) {
        const float3 o = {ro[3 * i], ro[3 * i + 1], ro[3 * i + 2]}, d = {rd[3 * i], rd[3 * i + 1], rd[3 * i + 2]};
        Hit hit;
        hit.t = t_max;
        if (query == 0)
            intersect_stackfull(&hit, s, o, d, invRd(d), nullptr);
        else
            intersect_stackfree(&hit, s, o, d, invRd(d), query == 2);
        out[i].t = hit.t;
        out[i].u = hit.u;
        out[i].v = hit.v;
        out[i].ng = {hit.ng.x, hit.ng.y, hit.ng.z};
        out[i].tri = hit.tri;
    }
}
// Brute force over every triangle, no BVH: the nearest hit (hit_t, hit_tri as
// above; on equal t the later triangle wins, as when the stackful walk meets
// it later) and whether any triangle is hit over [0, t_max] (any_out)
void orc_brute_hits(const palz_triangle* tris, uint32_t n_tris, const float* ro, const float* rd, uint32_t n,
                    float t_max, float* hit_t, uint32_t* hit_tri, uint8_t* any_out) {
    for (uint32_t i = 0; i < n; ++i) {
        const float3 o = {ro[3 * i], ro[3 * i + 1], ro[3 * i + 2]}, d = {rd[3 * i], rd[3 * i + 1], rd[3 * i + 2]};
        float best = 3.402823466e+38F;
        uint32_t best_tri = 0xFFFFFFFFu;
        bool any = false;
        for (uint32_t k = 0; k < n_tris; ++k) {
            float t, u, v;
            float3 ng;
            const palz_triangle& tri = tris[k];
            if (intersectRayTriangle(&t, &u, &v, &ng, 0.0f, best, o, d, load3(tri.vs[0]), load3(tri.vs[1]),
                                     load3(tri.vs[2]))) {
                best = t;
                best_tri = k;
                if (t <= t_max) any = true;
            }
        }
        hit_t[i] = best;
        hit_tri[i] = best_tri;
        any_out[i] = any ? 1 : 0;
    }
}
// intersect_escape_link (minimum_lbvh.h:1078-1152) over [0, t_max] for n rays
void orc_any_hits(const palz_triangle* tris, const palz_internal_node* nodes, uint32_t root, const float* ro,
                  const float* rd, uint32_t n, float t_max, uint8_t* out) {
    const Scene s{tris, nullptr, nodes, root};
    for (uint32_t i = 0; i < n; ++i) {
        const float3 o = {ro[3 * i], ro[3 * i + 1], ro[3 * i + 2]}, d = {rd[3 * i], rd[3 * i + 1], rd[3 * i + 2]};
        out[i] = any_hit(s, o, d, invRd(d), t_max) ? 1 : 0;
    }
}
// occluded (pk.h:88-115) for n segments (from, from_n, to, to_n: 3 floats each)
void orc_occluded(const palz_triangle* tris, const palz_internal_node* nodes, uint32_t root, const float* from,
                  const float* from_n, const float* to, const float* to_n, uint32_t n, uint8_t* out) {
    const Scene s{tris, nullptr, nodes, root};
    auto ld = [](const float* p, uint32_t i) { return float3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; };
    for (uint32_t i = 0; i < n; ++i)
        out[i] = occluded(s, ld(from, i), ld(from_n, i), ld(to, i), ld(to_n, i)) ? 1 : 0;
}

}  // extern "C"
