// Adjoint of receiver irradiance with respect to heightfield vertex heights.
//
// Every contributing (path, wavelength) item adds c(x*(q), q) to its receiver
// pixel, where q are the item's local inputs (the heights of its chain
// triangles' vertices and their shading normals) and x* the converged
// barycentrics, which solve the chain constraint F(x, q) = 0 of
// solveConstraints (pk.h:477-590).  By the implicit-function theorem
//     dx*/dq = -J_x^+ J_q,   J_x = dF/dx (3K x 2K),  J_q = dF/dq,
// so  dc/dq = dc/dq|_x + dc/dx dx*/dq.  Per item the kernel refines x* in FP64,
// forms J_x with one multi-tangent pass and its Householder QR, and keeps x*
// and J_x^+ in LDS.  Then every local input has a lane of its own: J_q e by a
// one-tangent forward pass of the residual, dx*/dq e = -J_x^+ J_q e, and the
// total dc/dq e by one forward pass of the contribution carrying both
// (throughput, dA/dw with its own two light-direction tangents nested inside,
// the receiver cosine and the clamp: pk.h:760-930, main_gpu.cu:458-471).  One
// tangent per lane keeps the nested numbers small enough for registers: the
// state of 4 tangents in one lane spilled 1.3-2.2k VGPRs.  The item's weight
// dL/dI times dc/dq is then pushed through the heights' normals (central
// differences, normalised) into the height gradient with FP64 atomics.
// Visibility, validity and the tuple set are those of the forward solve.
#include "optics.hpp"

namespace palz {
namespace ad {

// forward-mode number over a scalar S (double, or another Dn for nesting)
template <class S, int N>
struct Dn {
    S v;
    S d[N];
};

PALZ_HD double val(double x) { return x; }
template <class S, int N>
PALZ_HD double val(const Dn<S, N>& a) { return val(a.v); }

template <class S>
struct Cst {
    PALZ_HD static S make(double c) { return c; }
};
template <class S, int N>
struct Cst<Dn<S, N>> {
    PALZ_HD static Dn<S, N> make(double c) {
        Dn<S, N> r;
        r.v = Cst<S>::make(c);
#pragma unroll
        for (int t = 0; t < N; ++t) r.d[t] = Cst<S>::make(0.0);
        return r;
    }
};
template <class S>
PALZ_HD S cst(double c) { return Cst<S>::make(c); }

template <class S, int N>
PALZ_HD Dn<S, N> operator+(const Dn<S, N>& a, const Dn<S, N>& b) {
    Dn<S, N> r;
    r.v = a.v + b.v;
#pragma unroll
    for (int t = 0; t < N; ++t) r.d[t] = a.d[t] + b.d[t];
    return r;
}
template <class S, int N>
PALZ_HD Dn<S, N> operator-(const Dn<S, N>& a, const Dn<S, N>& b) {
    Dn<S, N> r;
    r.v = a.v - b.v;
#pragma unroll
    for (int t = 0; t < N; ++t) r.d[t] = a.d[t] - b.d[t];
    return r;
}
template <class S, int N>
PALZ_HD Dn<S, N> operator-(const Dn<S, N>& a) {
    Dn<S, N> r;
    r.v = -a.v;
#pragma unroll
    for (int t = 0; t < N; ++t) r.d[t] = -a.d[t];
    return r;
}
template <class S, int N>
PALZ_HD Dn<S, N> operator*(const Dn<S, N>& a, const Dn<S, N>& b) {
    Dn<S, N> r;
    r.v = a.v * b.v;
#pragma unroll
    for (int t = 0; t < N; ++t) r.d[t] = a.d[t] * b.v + a.v * b.d[t];
    return r;
}
template <class S, int N>
PALZ_HD Dn<S, N> operator/(const Dn<S, N>& a, const Dn<S, N>& b) {
    Dn<S, N> r;
    r.v = a.v / b.v;
#pragma unroll
    for (int t = 0; t < N; ++t) r.d[t] = (a.d[t] - r.v * b.d[t]) / b.v;
    return r;
}
PALZ_HD double rt(double x) { return sqrt(x); }
template <class S, int N>
PALZ_HD Dn<S, N> rt(const Dn<S, N>& a) {
    Dn<S, N> r;
    r.v = rt(a.v);
    const S k = cst<S>(0.5) / r.v;
#pragma unroll
    for (int t = 0; t < N; ++t) r.d[t] = a.d[t] * k;
    return r;
}

template <class S>
struct V3 {
    S x, y, z;
};
template <class S>
PALZ_HD V3<S> vc(v3 p) { return {cst<S>(p.x), cst<S>(p.y), cst<S>(p.z)}; }
template <class S>
PALZ_HD V3<S> operator+(const V3<S>& a, const V3<S>& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class S>
PALZ_HD V3<S> operator-(const V3<S>& a, const V3<S>& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class S>
PALZ_HD V3<S> operator-(const V3<S>& a) { return {-a.x, -a.y, -a.z}; }
template <class S>
PALZ_HD V3<S> operator*(const V3<S>& a, const S& s) { return {a.x * s, a.y * s, a.z * s}; }
template <class S>
PALZ_HD V3<S> operator/(const V3<S>& a, const S& s) { return {a.x / s, a.y / s, a.z / s}; }
template <class S>
PALZ_HD S dot(const V3<S>& a, const V3<S>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class S>
PALZ_HD V3<S> cross(const V3<S>& a, const V3<S>& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class S>
PALZ_HD S len(const V3<S>& a) { return rt(dot(a, a)); }
template <class S>
PALZ_HD V3<S> nrmz(const V3<S>& a) { return a * (cst<S>(1.0) / rt(dot(a, a))); }
template <class S>
PALZ_HD bool is_zero(const V3<S>& a) { return val(a.x) == 0.0 && val(a.y) == 0.0 && val(a.z) == 0.0; }

// saka::reflection / refraction_norm_free (saka.h:192-208)
template <class S>
PALZ_HD V3<S> reflect(const V3<S>& wi, const V3<S>& n) { return n * (dot(wi, n) * cst<S>(2.0) / dot(n, n)) - wi; }
template <class S>
PALZ_HD V3<S> refract(const V3<S>& wi, const V3<S>& n, const S& eta) {
    const S non = dot(n, n), win = dot(wi, n), ww = dot(wi, wi);
    const S k = non * ww * (eta * eta - cst<S>(1.0)) + win * win;
    if (val(k) < 0.0) return {cst<S>(0.0), cst<S>(0.0), cst<S>(0.0)};
    return -wi * non + n * (win - rt(k));
}
// fresnel_exact_norm_free (pk.h:427-450)
template <class S>
PALZ_HD S fresnel(const V3<S>& wi, V3<S> n, S eta) {
    const S nn = dot(n, n), ww = dot(wi, wi);
    S c = dot(n, wi);
    if (val(c) < 0.0) {
        c = -c;
        n = -n;
        eta = cst<S>(1.0) / eta;
    }
    const S k = nn * ww * (eta * eta - cst<S>(1.0)) + c * c;
    if (val(k) < 0.0) return cst<S>(1.0);
    const S g = rt(k);
    const S gmc = g - c, gpc = g + c;
    const S a = gmc / gpc;
    const S q = (c * gpc - nn * ww) / (c * gmc + nn * ww);
    return cst<S>(0.5) * (a * a) * (cst<S>(1.0) + q * q);
}

// One chain triangle with its inputs: vertices and shading normals.
template <class S>
struct Tri {
    V3<S> v[3], s[3];
    int material;
};
template <class S>
PALZ_HD V3<S> at(const V3<S>* p, const S& u, const S& w) { return p[0] + (p[1] - p[0]) * u + (p[2] - p[0]) * w; }

// solveConstraints' residual (pk.h:505-548)
template <int K, class S>
PALZ_HD void residual(const S* x, const Tri<S>* tri, const V3<S>& pb, const V3<S>& pe, const S& eta, uint32_t event,
                      S* F) {
    V3<S> vert[K + 2];
    vert[0] = pb;
    vert[K + 1] = pe;
#pragma unroll
    for (int k = 0; k < K; ++k) vert[k + 1] = at(tri[k].v, x[2 * k], x[2 * k + 1]);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        V3<S> wi = vert[k] - vert[k + 1];
        V3<S> wo = vert[k + 2] - vert[k + 1];
        const V3<S> n = at(tri[k].s, x[2 * k], x[2 * k + 1]);
        if (val(dot(wi, n)) < 0.0) {
            const V3<S> t = wi;
            wi = wo;
            wo = t;
        }
        const V3<S> wopt = ev_is_T(event, k) ? refract(wi, n, eta) : reflect(wi, n);
        const V3<S> c = cross(wo, wopt);
        F[3 * k + 0] = c.x;
        F[3 * k + 1] = c.y;
        F[3 * k + 2] = c.z;
    }
}

// dAdw (pk.h:844-930) with the light-direction tangents as an inner Dn<S, 2>
template <int K, class S>
PALZ_HD S footprint(const V3<S>& ro, V3<S> rd, const V3<S>& pe, const Tri<S>* tri, uint32_t event, const S& eta) {
    using D = Dn<S, 2>;
    rd = nrmz(rd);
    // GetOrthonormalBasis (pk.h:453-459)
    const S sign = cst<S>(val(rd.z) < 0.0 ? -1.0 : 1.0);
    const S a = cst<S>(-1.0) / (sign + rd.z);
    const S b = rd.x * rd.y * a;
    const V3<S> T0 = {cst<S>(1.0) + sign * rd.x * rd.x * a, sign * b, -sign * rd.x};
    const V3<S> T1 = {b, sign + rd.y * rd.y * a, -rd.y};
    auto lift = [](const V3<S>& p) -> V3<D> {
        V3<D> r;
        r.x.v = p.x;
        r.y.v = p.y;
        r.z.v = p.z;
#pragma unroll
        for (int t = 0; t < 2; ++t) r.x.d[t] = r.y.d[t] = r.z.d[t] = cst<S>(0.0);
        return r;
    };
    V3<D> ro_j = lift(ro);
    V3<D> rd_j = lift(rd);
    rd_j.x.d[0] = T0.x;
    rd_j.y.d[0] = T0.y;
    rd_j.z.d[0] = T0.z;
    rd_j.x.d[1] = T1.x;
    rd_j.y.d[1] = T1.y;
    rd_j.z.d[1] = T1.z;
    const D eta_d = {eta, {cst<S>(0.0), cst<S>(0.0)}};
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const V3<D> v0 = lift(tri[j].v[0]), v1 = lift(tri[j].v[1]), v2 = lift(tri[j].v[2]);
        const V3<D> ng = cross(v1 - v0, v2 - v0);
        const D t = dot(v0 - ro_j, ng) / dot(ng, rd_j);  // intersect_p_ray_plane (pk.h:594-598)
        const V3<D> p = ro_j + rd_j * t;
        D v = dot(p - v1, cross(ng, v1 - v0));
        D w = dot(p - v2, cross(ng, v2 - v1));
        D u = dot(p - v0, cross(ng, v0 - v2));
        const D area = u + v + w;
        u = u / area;
        v = v / area;
        w = w / area;
        const V3<D> n = lift(tri[j].s[0]) * w + lift(tri[j].s[1]) * u + lift(tri[j].s[2]) * v;
        const V3<D> wi = -rd_j;
        V3<D> wo;
        if (ev_is_T(event, j)) {
            if (val(dot(wi, n)) < 0.0) {
                wo = refract(wi, -n, cst<D>(1.0) / eta_d);
                if (is_zero(wo)) return cst<S>(3.402823466e+38);
            } else {
                wo = refract(wi, n, eta_d);
            }
        } else {
            wo = reflect(wi, n);
        }
        ro_j = p;
        rd_j = wo;
    }
    const V3<D> pl = {{ro_j.x.v, {cst<S>(0.0), cst<S>(0.0)}}, {ro_j.y.v, {cst<S>(0.0), cst<S>(0.0)}},
                      {ro_j.z.v, {cst<S>(0.0), cst<S>(0.0)}}};
    const V3<D> pe_d = lift(pe);
    const V3<D> nrm = pl - pe_d;
    const D t = dot(pe_d - ro_j, nrm) / dot(nrm, rd_j);
    const V3<D> pf = ro_j + rd_j * t;
    const V3<S> a0 = {pf.x.d[0], pf.y.d[0], pf.z.d[0]};
    const V3<S> a1 = {pf.x.d[1], pf.y.d[1], pf.z.d[1]};
    const S l0 = len(a0), l1 = len(a1);
    S s = len(cross(a0 / l0, a1 / l1));
    if (val(s) < 0.01) s = cst<S>(0.01);
    return l0 * l1 * s;
}

// The item's irradiance contribution (main_gpu.cu:458-471) on its forward
// visibility and validity: throughput's Fresnel factors, dA/dw, the receiver
// cosine and the clamp.
template <int K, class S>
PALZ_HD S contribution(const S* x, const Tri<S>* tri, const V3<S>& pb, const V3<S>& pe, const V3<S>& ng_end,
                       const S& eta, uint32_t event, double intensity, double clamp) {
    V3<S> vert[K + 2];
    V3<S> sn[K];
    vert[0] = pb;
    vert[K + 1] = pe;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        vert[k + 1] = at(tri[k].v, x[2 * k], x[2 * k + 1]);
        sn[k] = nrmz(at(tri[k].s, x[2 * k], x[2 * k + 1]));
    }
    S thr = cst<S>(1.0);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (tri[k].material != PALZ_MATERIAL_DIELECTRIC) continue;
        const S r = fresnel(vert[k] - vert[k + 1], sn[k], eta);
        thr = thr * (ev_is_T(event, k) ? cst<S>(1.0) - r : r);
    }
    const S area = footprint<K>(pb, vert[1] - pb, pe, tri, event, eta);
    const S cosr = dot(nrmz(vert[K] - pe), ng_end);
    if (!(val(cosr) > 0.0)) return cst<S>(0.0);
    const S c = thr * cst<S>(intensity) / area * cosr;
    if (!(val(c) < clamp)) return cst<S>(clamp);
    return c;
}

// Householder QR of A (R x C, column-major) kept for several right-hand sides.
template <int R, int C>
struct QR {
    double a[R * C];
    double beta[C];
    PALZ_HD void factor() {
#pragma unroll
        for (int j = 0; j < C; ++j) {
            double norm = 0.0;
            for (int i = j; i < R; ++i) norm += a[j * R + i] * a[j * R + i];
            norm = sqrt(norm);
            const double alpha = a[j * R + j] > 0.0 ? -norm : norm;
            const double v0 = a[j * R + j] - alpha;
            a[j * R + j] = v0;
            double vv = 0.0;
            for (int i = j; i < R; ++i) vv += a[j * R + i] * a[j * R + i];
            beta[j] = vv > 0.0 ? 2.0 / vv : 0.0;
            for (int c = j + 1; c < C; ++c) {
                double s = 0.0;
                for (int i = j; i < R; ++i) s += a[j * R + i] * a[c * R + i];
                s *= beta[j];
                for (int i = j; i < R; ++i) a[c * R + i] -= s * a[j * R + i];
            }
            a[j * R + j] = alpha;  // R's diagonal; the reflector's head is kept apart
            vhead[j] = v0;
        }
    }
    double vhead[C];
    // least-squares solution of A x = b
    PALZ_HD void solve(const double* b_in, double* x) const {
        double b[R];
        for (int i = 0; i < R; ++i) b[i] = b_in[i];
        for (int j = 0; j < C; ++j) {
            double s = vhead[j] * b[j];
            for (int i = j + 1; i < R; ++i) s += a[j * R + i] * b[i];
            s *= beta[j];
            b[j] -= s * vhead[j];
            for (int i = j + 1; i < R; ++i) b[i] -= s * a[j * R + i];
        }
        for (int r = C - 1; r >= 0; --r) {
            double s = b[r];
            for (int c = r + 1; c < C; ++c) s -= a[c * R + r] * x[c];
            x[r] = s / a[r * R + r];
        }
    }
};

}  // namespace ad

namespace {

// Items per workgroup (one wave): their FP64 solutions and pseudo-inverses
// stay in LDS between the two phases of k_adjoint.
constexpr int kAdjItems = 64;
constexpr uint32_t kAdjSpread = 4096;  // waves per window (see k_adjoint)

// An item's FP64 solution x* and the pseudo-inverse J_x^+ (P x R, row-major)
// of its constraint Jacobian there.
template <int K>
struct ItemState {
    double x[2 * K];
    double pinv[2 * K * 3 * K];
};

// The chain's triangles as one-tangent numbers: vertex c of triangle k gets
// the unit tangent in its height (what 0) or in one shading-normal component
// (what 1..3); every other input is constant.  Selected per lane without
// indexing registers by the lane's (k, c).
template <int K, class G>
PALZ_HD void seeded_tris(const ChainTri* ct, int ks, int cs, int what, ad::Tri<G>* tg) {
    using namespace ad;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const v3 v[3] = {ct[k].v0, ct[k].v1, ct[k].v2};
        const v3 n[3] = {ct[k].s0, ct[k].s1, ct[k].s2};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            tg[k].v[c] = vc<G>(v[c]);
            tg[k].s[c] = vc<G>(n[c]);
            const bool me = k == ks && c == cs;
            tg[k].v[c].y.d[0] = (me && what == 0) ? 1.0 : 0.0;
            tg[k].s[c].x.d[0] = (me && what == 1) ? 1.0 : 0.0;
            tg[k].s[c].y.d[0] = (me && what == 2) ? 1.0 : 0.0;
            tg[k].s[c].z.d[0] = (me && what == 3) ? 1.0 : 0.0;
        }
        tg[k].material = ct[k].material;
    }
}

// Phase A of one item: the contributing check (stored results, or the FP32
// forward solve recomputed), three FP64 Newton steps from the FP32 solution,
// and J_x^+ at x* from J_x's Householder QR.  Returns the chain vertices
// whose heights are parameters as a bit mask (bit 3k + c), 0 when the item
// adds nothing.
template <int K>
__device__ uint32_t adjoint_prepare(const palz_scene_desc& sc, const palz_height_param& hp, const palz_path& path,
                                    uint32_t l, uint64_t item, const palz_path_result* __restrict__ results,
                                    ItemState<K>* st) {
    using namespace ad;
    constexpr int P = 2 * K;
    constexpr int R = 3 * K;
    const uint32_t event = path.event;
    ChainTri ct[K];
#pragma unroll
    for (int k = 0; k < K; ++k) ct[k] = load_chain_tri(sc.tris, sc.attribs, path.tris[k]);
    const palz_first_diffuse q = sc.points[path.pixel];
    const float lam = sc.lambdas[l];
    const float eta32 = sc.cauchy_a + sc.cauchy_b / (lam * lam);
    float x32[P];
    if (results) {
        const palz_path_result r = results[item];
        if (r.status != PALZ_PATH_CONTRIBUTES) return 0u;
#pragma unroll
        for (int i = 0; i < P; ++i) x32[i] = r.params[i];
    } else {
        // recompute the forward solve instead of storing it
#pragma unroll
        for (int i = 0; i < P; ++i) x32[i] = path.params[i];
        int iters = 0;
        if (!solve_constraints<K, 1, float, v3>(x32, mk(sc.p_light), mk(q.p), ct, eta32, event, 32, 1.0e-10f, 1,
                                                &iters))
            return 0u;
        if (!(0.0f < contributable_throughput<K>(x32, mk(sc.p_light), mk(q.p), ct, event, sc.nodes, sc.tris,
                                                 *sc.root, eta32)))
            return 0u;
    }
    uint32_t mask = 0u;
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            if (hp.tri_vertex[3 * path.tris[k] + c] >= 0) mask |= 1u << (3 * k + c);
    if (mask == 0u) return 0u;
    // the FP64 solution: three Newton steps from the FP32 one, keeping the
    // factorisation at x*
    const double eta = (double)eta32;
    using J = Dn<double, P>;
    double x[P];
#pragma unroll
    for (int i = 0; i < P; ++i) x[i] = (double)x32[i];
    QR<R, P> qr;
    for (int step = 0; step < 3; ++step) {
        J xj[P];
#pragma unroll
        for (int i = 0; i < P; ++i) {
            xj[i] = cst<J>(x[i]);
            xj[i].d[i] = 1.0;
        }
        Tri<J> tj[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const v3 v[3] = {ct[k].v0, ct[k].v1, ct[k].v2};
            const v3 n[3] = {ct[k].s0, ct[k].s1, ct[k].s2};
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                tj[k].v[c] = vc<J>(v[c]);
                tj[k].s[c] = vc<J>(n[c]);
            }
            tj[k].material = ct[k].material;
        }
        J F[R];
        residual<K, J>(xj, tj, vc<J>(mk(sc.p_light)), vc<J>(mk(q.p)), cst<J>(eta), event, F);
        double b[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            b[r] = F[r].v;
#pragma unroll
            for (int c = 0; c < P; ++c) qr.a[c * R + r] = F[r].d[c];
        }
        qr.factor();
        if (step == 2) break;
        double dx[P];
        qr.solve(b, dx);
#pragma unroll
        for (int i = 0; i < P; ++i) x[i] -= dx[i];
    }
#pragma unroll
    for (int i = 0; i < P; ++i) st->x[i] = x[i];
    // J_x^+ column by column: the least-squares solutions of J_x y = e_r
    for (int r = 0; r < R; ++r) {
        double e[R], y[P];
#pragma unroll
        for (int i = 0; i < R; ++i) e[i] = i == r ? 1.0 : 0.0;
        qr.solve(e, y);
#pragma unroll
        for (int i = 0; i < P; ++i) st->pinv[i * R + r] = y[i];
    }
    return mask;
}

// Phase B of one (item, chain vertex, input) triple: with the input's unit
// tangent, J_q = dF/dq by one forward pass of the residual at x*, dx*/dq =
// -J_x^+ J_q, then the total dc/dq by one forward pass of the contribution
// carrying both tangents.  Returns the weighted derivative.
template <int K>
__device__ double adjoint_input(const palz_scene_desc& sc, const palz_path& path, uint32_t l, float w_img,
                                const ItemState<K>& st, int ks, int cs, int what) {
    using namespace ad;
    constexpr int P = 2 * K;
    constexpr int R = 3 * K;
    using G = Dn<double, 1>;
    const uint32_t event = path.event;
    ChainTri ct[K];
#pragma unroll
    for (int k = 0; k < K; ++k) ct[k] = load_chain_tri(sc.tris, sc.attribs, path.tris[k]);
    const palz_first_diffuse q = sc.points[path.pixel];
    const float lam = sc.lambdas[l];
    const G eta = cst<G>((double)(sc.cauchy_a + sc.cauchy_b / (lam * lam)));
    Tri<G> tg[K];
    seeded_tris<K, G>(ct, ks, cs, what, tg);
    const V3<G> pb = vc<G>(mk(sc.p_light)), pe = vc<G>(mk(q.p)), ng = vc<G>(mk(q.ng));
    G xg[P];
#pragma unroll
    for (int i = 0; i < P; ++i) xg[i] = cst<G>(st.x[i]);
    G F[R];
    residual<K, G>(xg, tg, pb, pe, eta, event, F);
#pragma unroll
    for (int i = 0; i < P; ++i) {
        double s = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) s += st.pinv[i * R + r] * F[r].d[0];
        xg[i].d[0] = -s;
    }
    const G c = contribution<K, G>(xg, tg, pb, pe, ng, eta, event, sc.light_intensity, sc.radiance_clamp);
    return (double)w_img * c.d[0];
}

// d(loss)/d(heights) of the contributing items of chain length K.  One wave
// per kAdjItems items: phase A prepares each item in its own lane
// (x*, J_x^+ into LDS); phase B gives every (item, parameter vertex, input)
// triple a lane of its own, four lanes per vertex: the height and the three
// shading-normal components, each a single-tangent forward pass.  The four
// lanes of a vertex then combine the normal's chain rule and one of them adds
// the vertex's share into the heights with FP64 atomics.
template <int K>
__global__ void __launch_bounds__(kAdjItems)
    k_adjoint(palz_scene_desc sc, palz_height_param hp, const palz_path* __restrict__ paths, uint32_t n_paths,
              const palz_path_result* __restrict__ results, const float* __restrict__ weight,
              double* __restrict__ grad) {
    __shared__ ItemState<K> s_state[kAdjItems];
    __shared__ uint16_t s_unit[kAdjItems * 3 * K];  // lane << 4 | chain vertex (3k + c)
    const uint32_t L = sc.n_lambdas;
    const uint32_t lane = threadIdx.x;
    const uint64_t n_items = (uint64_t)n_paths * L;
    // Windows of kAdjSpread waves: lane l of the window's wave j takes item
    // l * kAdjSpread + j.  Callers pass lists ordered by region (the batched
    // solver's), whose neighbouring items add into the same heights; spread
    // kAdjSpread apart, a wave's FP64 atomics seldom meet on one address,
    // while a window stays a contiguous part of the list (one chain length).
    const uint64_t window = (uint64_t)(blockIdx.x / kAdjSpread) * (kAdjItems * kAdjSpread);
    const uint32_t j = blockIdx.x % kAdjSpread;
    // phase A
    uint32_t mask = 0u;
    {
        const uint64_t item = window + (uint64_t)lane * kAdjSpread + j;
        if (item < n_items) {
            const uint32_t pi = (uint32_t)(item / L), l = (uint32_t)(item % L);
            const palz_path path = paths[pi];
            if (ev_size(path.event) == K && weight[(size_t)l * sc.n_points + path.pixel] != 0.0f)
                mask = adjoint_prepare<K>(sc, hp, path, l, item, results, &s_state[lane]);
        }
    }
    // list the (item, vertex) units in lane order
    const uint32_t cnt = (uint32_t)__popc(mask);
    uint32_t incl = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)incl, o, 64);
        if (lane >= (uint32_t)o) incl += y;
    }
    const uint32_t total = (uint32_t)__shfl((int)incl, 63, 64);
    uint32_t pos = incl - cnt;
    for (uint32_t m = mask; m != 0u; m &= m - 1u) s_unit[pos++] = (uint16_t)(lane << 4 | (uint32_t)__ffs(m) - 1u);
    __syncthreads();
    // phase B: four lanes per unit
    for (uint32_t t0 = 0; t0 < 4u * total; t0 += 64u) {
        const uint32_t t = t0 + lane;
        const bool active = t < 4u * total;
        const uint32_t unit = active ? s_unit[t >> 2] : 0u;
        const uint32_t il = unit >> 4, v = unit & 15u;
        const int ks = (int)(v / 3u), cs = (int)(v % 3u), what = (int)(t & 3u);
        double g = 0.0, dmx = 0.0, dmz = 0.0;
        int h = -1;
        palz_path path;
        if (active) {
            const uint64_t item = window + (uint64_t)il * kAdjSpread + j;
            const uint32_t pi = (uint32_t)(item / L), l = (uint32_t)(item % L);
            path = paths[pi];
            const float w_img = weight[(size_t)l * sc.n_points + path.pixel];
            g = adjoint_input<K>(sc, path, l, w_img, s_state[il], ks, cs, what);
            h = hp.tri_vertex[3 * path.tris[ks] + cs];
            if (what > 0) {
                // the normal n = m / |m|, m = (-dh/dx, 1, -dh/dz): this
                // component's share of dL/dm = (I - n n^T) g / |m|, |m| = 1 / |n_y|
                const palz_triangle_attrib a = sc.attribs[path.tris[ks]];
                const double n3[3] = {a.shading_normals[cs].x, a.shading_normals[cs].y, a.shading_normals[cs].z};
                const int comp = what - 1;
                const double inv_m = fabs(n3[1]);
                dmx = ((comp == 0 ? g : 0.0) - n3[0] * n3[comp] * g) * inv_m;
                dmz = ((comp == 2 ? g : 0.0) - n3[2] * n3[comp] * g) * inv_m;
            }
        }
        // the vertex's four lanes are adjacent: sum the normal's shares
#pragma unroll
        for (int o = 1; o < 4; o <<= 1) {
            dmx += __shfl_xor(dmx, o, 64);
            dmz += __shfl_xor(dmz, o, 64);
        }
        if (active && what == 0) {
            atomicAdd(grad + h, g);
            for (int j = 0; j < 4; ++j) {
                const int nb = hp.normal_nb[4 * h + j];
                if (nb < 0) continue;
                // m_x = -hx, m_z = -hz
                const double d = -dmx * hp.normal_coef[8 * h + 2 * j] - dmz * hp.normal_coef[8 * h + 2 * j + 1];
                atomicAdd(grad + nb, d);
            }
        }
    }
}

}  // namespace
}  // namespace palz

extern "C" {

int palz_adjoint_irradiance(const palz_scene_desc* scene, const palz_height_param* hp, const palz_path* d_paths,
                            uint32_t n_paths, const palz_path_result* d_results, const float* d_weight,
                            double* d_grad, void* stream) {
    using namespace palz;
    if (!scene || !hp || !hp->tri_vertex || !hp->normal_nb || !hp->normal_coef || !d_weight || !d_grad ||
        (n_paths && !d_paths))
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_adjoint_irradiance: NULL argument");
    const uint64_t items = (uint64_t)n_paths * scene->n_lambdas;
    if (items == 0) return PALZ_OK;
    if (items >= (1ull << 31)) return fail(PALZ_ERR_CAPACITY, "palz_adjoint_irradiance: more than 2^31 items");
    const uint64_t per_window = (uint64_t)kAdjItems * kAdjSpread;
    const uint32_t blocks = (uint32_t)((items + per_window - 1) / per_window * kAdjSpread);
    hipStream_t st = (hipStream_t)stream;
    k_adjoint<1><<<blocks, kAdjItems, 0, st>>>(*scene, *hp, d_paths, n_paths, d_results, d_weight, d_grad);
    k_adjoint<2><<<blocks, kAdjItems, 0, st>>>(*scene, *hp, d_paths, n_paths, d_results, d_weight, d_grad);
    k_adjoint<3><<<blocks, kAdjItems, 0, st>>>(*scene, *hp, d_paths, n_paths, d_results, d_weight, d_grad);
    PALZ_LAUNCH_CHECK();
    return PALZ_OK;
}

}  // extern "C"
