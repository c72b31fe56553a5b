// Adjoint of receiver irradiance with respect to heightfield vertex heights.
//
// Every contributing (path, wavelength) item adds c(x*(q), q) to its receiver
// pixel, where q are the item's local inputs (the heights of its chain
// triangles' vertices and their shading normals) and x* the converged
// barycentrics, which solve the chain constraint F(x, q) = 0 of
// solveConstraints (pk.h:477-590).  By the implicit-function theorem
//     dx*/dq = -J_x^+ J_q,   J_x = dF/dx (3K x 2K),  J_q = dF/dq,
// so  dc/dq = dc/dq|_x + dc/dx dx*/dq.  Per item the kernel refines x* in FP64,
// forms J_x with one multi-tangent pass and its Householder QR, then for
// chunks of 4 local inputs: J_q by a forward pass, dx*/dq by the QR, and the
// total dc/dq by one forward pass of the contribution (throughput, dA/dw with
// its own two light-direction tangents nested inside, the receiver cosine and
// the clamp: pk.h:760-930, main_gpu.cu:458-471).  The item's weight dL/dI
// times dc/dq is then pushed through the heights' normals (central
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

constexpr int kChunk = 4;

template <int K>
__global__ void __launch_bounds__(64)
    k_adjoint(palz_scene_desc sc, palz_height_param hp, const palz_path* __restrict__ paths, uint32_t n_paths,
              const palz_path_result* __restrict__ results, const float* __restrict__ weight,
              double* __restrict__ grad) {
    using namespace ad;
    constexpr int P = 2 * K;
    constexpr int R = 3 * K;
    const uint32_t L = sc.n_lambdas;
    const uint64_t item = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= (uint64_t)n_paths * L) return;
    const uint32_t pi = (uint32_t)(item / L), l = (uint32_t)(item % L);
    const palz_path path = paths[pi];
    if (ev_size(path.event) != K) return;
    const float w_img = weight[(size_t)l * sc.n_points + path.pixel];
    if (w_img == 0.0f) return;
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
        if (r.status != PALZ_PATH_CONTRIBUTES) return;
#pragma unroll
        for (int i = 0; i < P; ++i) x32[i] = r.params[i];
    } else {
        // recompute the forward solve instead of storing it
#pragma unroll
        for (int i = 0; i < P; ++i) x32[i] = path.params[i];
        int iters = 0;
        if (!solve_constraints<K, 1, float, v3>(x32, mk(sc.p_light), mk(q.p), ct, eta32, event, 32, 1.0e-10f, 1,
                                                &iters))
            return;
        if (!(0.0f < contributable_throughput<K>(x32, mk(sc.p_light), mk(q.p), ct, event, sc.nodes, sc.tris,
                                                 *sc.root, eta32)))
            return;
    }
    // the FP64 solution: two Newton steps from the FP32 one
    Tri<double> td[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        td[k].v[0] = vc<double>(ct[k].v0);
        td[k].v[1] = vc<double>(ct[k].v1);
        td[k].v[2] = vc<double>(ct[k].v2);
        td[k].s[0] = vc<double>(ct[k].s0);
        td[k].s[1] = vc<double>(ct[k].s1);
        td[k].s[2] = vc<double>(ct[k].s2);
        td[k].material = ct[k].material;
    }
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
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                tj[k].v[c] = {cst<J>(td[k].v[c].x), cst<J>(td[k].v[c].y), cst<J>(td[k].v[c].z)};
                tj[k].s[c] = {cst<J>(td[k].s[c].x), cst<J>(td[k].s[c].y), cst<J>(td[k].s[c].z)};
            }
            tj[k].material = td[k].material;
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
        if (step == 2) break;  // keep the factorisation at x*
        double dx[P];
        qr.solve(b, dx);
#pragma unroll
        for (int i = 0; i < P; ++i) x[i] -= dx[i];
    }
    // local inputs: per chain triangle whose vertices are parameters, the
    // height (vertex y) and the shading normal (x, y, z) of each vertex
    int slot_tri[4 * 3 * K];
    int slot_vtx[4 * 3 * K];
    int slot_what[4 * 3 * K];
    int n_slots = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int t = path.tris[k];
        for (int c = 0; c < 3; ++c) {
            const int h = hp.tri_vertex[3 * t + c];
            if (h < 0) continue;
            for (int what = 0; what < 4; ++what) {
                slot_tri[n_slots] = k;
                slot_vtx[n_slots] = c;
                slot_what[n_slots] = what;
                ++n_slots;
            }
        }
    }
    double gq[4 * 3 * K];
    using G = Dn<double, kChunk>;
    for (int s0 = 0; s0 < n_slots; s0 += kChunk) {
        // inputs with their tangents
        Tri<G> tg[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                tg[k].v[c] = {cst<G>(td[k].v[c].x), cst<G>(td[k].v[c].y), cst<G>(td[k].v[c].z)};
                tg[k].s[c] = {cst<G>(td[k].s[c].x), cst<G>(td[k].s[c].y), cst<G>(td[k].s[c].z)};
            }
            tg[k].material = td[k].material;
        }
        for (int t = 0; t < kChunk && s0 + t < n_slots; ++t) {
            const int k = slot_tri[s0 + t], c = slot_vtx[s0 + t], what = slot_what[s0 + t];
            if (what == 0) tg[k].v[c].y.d[t] = 1.0;
            if (what == 1) tg[k].s[c].x.d[t] = 1.0;
            if (what == 2) tg[k].s[c].y.d[t] = 1.0;
            if (what == 3) tg[k].s[c].z.d[t] = 1.0;
        }
        const V3<G> pb = vc<G>(mk(sc.p_light)), pe = vc<G>(mk(q.p)), ng = vc<G>(mk(q.ng));
        const G eta_g = cst<G>(eta);
        // J_q, then dx*/dq = -J_x^+ J_q
        G xg[P];
#pragma unroll
        for (int i = 0; i < P; ++i) xg[i] = cst<G>(x[i]);
        G F[R];
        residual<K, G>(xg, tg, pb, pe, eta_g, event, F);
#pragma unroll
        for (int t = 0; t < kChunk; ++t) {
            double bq[R], dx[P];
#pragma unroll
            for (int r = 0; r < R; ++r) bq[r] = F[r].d[t];
            qr.solve(bq, dx);
#pragma unroll
            for (int i = 0; i < P; ++i) xg[i].d[t] = -dx[i];
        }
        const G c = contribution<K, G>(xg, tg, pb, pe, ng, eta_g, event, sc.light_intensity, sc.radiance_clamp);
        for (int t = 0; t < kChunk && s0 + t < n_slots; ++t) gq[s0 + t] = (double)w_img * c.d[t];
    }
    // into the heights: y directly; the normal through n = m / |m|,
    // m = (-dh/dx, 1, -dh/dz) from central differences of the heights
    for (int s = 0; s < n_slots; s += 4) {
        const int k = slot_tri[s], c = slot_vtx[s];
        const int h = hp.tri_vertex[3 * path.tris[k] + c];
        atomicAdd(grad + h, gq[s]);
        const v3 n = c == 0 ? ct[k].s0 : (c == 1 ? ct[k].s1 : ct[k].s2);
        const double nx = n.x, ny = n.y, nz = n.z;
        const double gx = gq[s + 1], gy = gq[s + 2], gz = gq[s + 3];
        // dL/dm = (I - n n^T) g / |m|, |m| = 1 / |n_y|
        const double ndg = nx * gx + ny * gy + nz * gz;
        const double inv_m = fabs(ny);
        const double dmx = (gx - nx * ndg) * inv_m;
        const double dmz = (gz - nz * ndg) * inv_m;
        for (int j = 0; j < 4; ++j) {
            const int nb = hp.normal_nb[4 * h + j];
            if (nb < 0) continue;
            // m_x = -hx, m_z = -hz
            const double v = -dmx * hp.normal_coef[8 * h + 2 * j] - dmz * hp.normal_coef[8 * h + 2 * j + 1];
            atomicAdd(grad + nb, v);
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
    const uint32_t blocks = (uint32_t)((items + 63) / 64);
    hipStream_t st = (hipStream_t)stream;
    k_adjoint<1><<<blocks, 64, 0, st>>>(*scene, *hp, d_paths, n_paths, d_results, d_weight, d_grad);
    k_adjoint<2><<<blocks, 64, 0, st>>>(*scene, *hp, d_paths, n_paths, d_results, d_weight, d_grad);
    k_adjoint<3><<<blocks, 64, 0, st>>>(*scene, *hp, d_paths, n_paths, d_results, d_weight, d_grad);
    PALZ_LAUNCH_CHECK();
    return PALZ_OK;
}

}  // extern "C"
