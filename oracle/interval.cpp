// Oracle: interval arithmetic with directed rounding and the per-tuple
// admissibility chain of path cuts (interval.h:34-713, pk.h:2101-2207).
// TEST INFRASTRUCTURE — a scalar restatement; the tests pin it against exact
// rational arithmetic and against the FP32/FP64 solver, not against itself.
//
// Rounding: each operation is evaluated round-to-nearest and its exact error
// recovered with an error-free transformation (TwoSum; an FMA residual for
// products, quotients and square roots); a bound moves one ulp outward when
// the error points outward.  Compiled with -ffp-contract=off.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle.hpp"

using namespace orc;

namespace {

constexpr float kInf = INFINITY;
constexpr float kFltMax = 3.40282347e+38f;
constexpr float kTiny = 0x1p-100f;

float step_up(float x) {
    if (std::isnan(x) || x == kInf) return x;
    if (x == 0.0f) return 0x1p-126f;
    int32_t b;
    std::memcpy(&b, &x, 4);
    b += x > 0.0f ? 1 : -1;
    float r;
    std::memcpy(&r, &b, 4);
    return r;
}
float step_down(float x) { return -step_up(-x); }

enum class Dir { down, up };

float outward(float r, float err, Dir d) {
    if (d == Dir::down) return err < 0.0f ? step_down(r) : r;
    return err > 0.0f ? step_up(r) : r;
}

float r_add(float a, float b, Dir d) {
    const float s = a + b;
    if (std::isinf(s) && !std::isinf(a) && !std::isinf(b)) return d == Dir::down ? (s > 0 ? kFltMax : s) : (s < 0 ? -kFltMax : s);
    if (!std::isfinite(s)) return s;
    const float bv = s - a, av = s - bv;
    return outward(s, (a - av) + (b - bv), d);
}
float r_mul(float a, float b, Dir d) {
    if (a == 0.0f || b == 0.0f) return 0.0f;
    const float p = a * b;
    if (std::isinf(p)) return d == Dir::down ? (p > 0 ? kFltMax : p) : (p < 0 ? -kFltMax : p);
    if (!std::isfinite(p)) return p;
    if (std::fabs(p) < kTiny) return d == Dir::down ? p - kTiny : p + kTiny;
    return outward(p, std::fmaf(a, b, -p), d);
}
float r_div(float a, float b, Dir d) {
    if (a == 0.0f) return 0.0f;
    const float q = a / b;
    if (std::isinf(q)) return d == Dir::down ? (q > 0 ? kFltMax : q) : (q < 0 ? -kFltMax : q);
    if (!std::isfinite(q)) return q;
    if (std::fabs(q) < kTiny) return d == Dir::down ? q - kTiny : q + kTiny;
    const float rem = std::fmaf(-q, b, a);  // a - q b exactly; sign(a/b - q) = sign(rem) sign(b)
    return outward(q, b > 0.0f ? rem : -rem, d);
}
float r_sqrt(float a, Dir d) {
    if (a == 0.0f || a == kInf) return a;
    const float s = std::sqrt(a);
    if (s < kTiny) return d == Dir::down ? 0.0f : kTiny;
    return outward(s, std::fmaf(-s, s, a), d);
}

struct Iv {
    float l, u;
    static Iv point(float v) { return {v, v}; }
    bool has_zero() const { return l <= 0.0f && 0.0f <= u; }
};
Iv operator-(Iv a) { return {-a.u, -a.l}; }
Iv operator+(Iv a, Iv b) { return {r_add(a.l, b.l, Dir::down), r_add(a.u, b.u, Dir::up)}; }
Iv operator-(Iv a, Iv b) { return {r_add(a.l, -b.u, Dir::down), r_add(a.u, -b.l, Dir::up)}; }
Iv operator*(Iv a, Iv b) {
    const float c[4][2] = {{a.l, b.l}, {a.l, b.u}, {a.u, b.l}, {a.u, b.u}};
    Iv r = {kInf, -kInf};
    for (auto& p : c) {
        r.l = std::fmin(r.l, r_mul(p[0], p[1], Dir::down));
        r.u = std::fmax(r.u, r_mul(p[0], p[1], Dir::up));
    }
    return r;
}
Iv operator/(Iv a, Iv b) {
    if (b.has_zero()) return {-kInf, kInf};
    const float c[4][2] = {{a.l, b.l}, {a.l, b.u}, {a.u, b.l}, {a.u, b.u}};
    Iv r = {kInf, -kInf};
    for (auto& p : c) {
        r.l = std::fmin(r.l, r_div(p[0], p[1], Dir::down));
        r.u = std::fmax(r.u, r_div(p[0], p[1], Dir::up));
    }
    return r;
}
Iv sq(Iv x) {
    const float hi = std::fmax(r_mul(x.l, x.l, Dir::up), r_mul(x.u, x.u, Dir::up));
    if (x.has_zero()) return {0.0f, hi};
    return {std::fmin(r_mul(x.l, x.l, Dir::down), r_mul(x.u, x.u, Dir::down)), hi};
}
Iv root(Iv x) { return {r_sqrt(std::fmax(x.l, 0.0f), Dir::down), r_sqrt(std::fmax(x.u, 0.0f), Dir::up)}; }

struct Iv3 {
    Iv c[3];
};
Iv3 operator-(Iv3 a) { return {{-a.c[0], -a.c[1], -a.c[2]}}; }
Iv3 operator-(Iv3 a, Iv3 b) { return {{a.c[0] - b.c[0], a.c[1] - b.c[1], a.c[2] - b.c[2]}}; }
Iv3 operator+(Iv3 a, Iv3 b) { return {{a.c[0] + b.c[0], a.c[1] + b.c[1], a.c[2] + b.c[2]}}; }
Iv3 scale(Iv3 a, Iv s) { return {{a.c[0] * s, a.c[1] * s, a.c[2] * s}}; }
Iv dot3(Iv3 a, Iv3 b) { return a.c[0] * b.c[0] + a.c[1] * b.c[1] + a.c[2] * b.c[2]; }
Iv norm2(Iv3 a) { return sq(a.c[0]) + sq(a.c[1]) + sq(a.c[2]); }
Iv3 cross3(Iv3 a, Iv3 b) {
    return {{a.c[1] * b.c[2] - a.c[2] * b.c[1], a.c[2] * b.c[0] - a.c[0] * b.c[2], a.c[0] * b.c[1] - a.c[1] * b.c[0]}};
}
float magnitude(Iv3 a) {
    float m = 0.0f;
    for (auto& x : a.c) m = std::fmax(m, std::fmax(std::fabs(x.l), std::fabs(x.u)));
    return m;
}
bool meets_zero(Iv3 v, float tau) {
    for (auto& x : v.c)
        if (x.l > tau || x.u < -tau) return false;
    return true;
}

// box of three points with the canonical min/max (-0 below +0)
Iv3 hull(const palz_float3* p) {
    Iv3 r;
    const float* v[3] = {&p[0].x, &p[1].x, &p[2].x};
    for (int a = 0; a < 3; ++a) {
        float lo = v[0][a], hi = v[0][a];
        for (int i = 1; i < 3; ++i) {
            lo = box_min(lo, v[i][a]);
            hi = box_max(hi, v[i][a]);
        }
        r.c[a] = {lo, hi};
    }
    return r;
}

// one orientation of one chain vertex: solveConstraints' residual (pk.h:505-548)
// within the tolerance the solver accepts plus FP32 evaluation slack
bool oriented(Iv3 wi, Iv3 wo, Iv3 n, bool refract, Iv eta) {
    const Iv non = norm2(n);
    const Iv win = dot3(wi, n);
    Iv3 wopt;
    if (refract) {
        const Iv k = non * norm2(wi) * (sq(eta) - Iv::point(1.0f)) + sq(win);
        if (!(k.l >= 0.0f)) return true;  // total internal reflection possible: residual 0
        wopt = scale(-wi, non) + scale(n, win - root(k));
    } else {
        wopt = scale(n, (win * Iv::point(2.0f)) / non) - wi;
    }
    const Iv3 c = cross3(wo, wopt);
    const float tau = 1.0e-5f + 1.0e-5f * std::fmax(1.0f, magnitude(wo) * magnitude(wopt));
    return meets_zero(c, tau);
}

bool vertex_ok(Iv3 wi, Iv3 wo, Iv3 n, bool refract, Iv eta) {
    const Iv s = dot3(wi, n);
    if (!(s.u < 0.0f) && oriented(wi, wo, n, refract, eta)) return true;
    if (!(s.l >= 0.0f) && oriented(wo, wi, n, refract, eta)) return true;
    return false;
}

Iv3 point3(const palz_float3& p) { return {{Iv::point(p.x), Iv::point(p.y), Iv::point(p.z)}}; }

// the part of the parameter rectangle inside the triangle, mapped by
// p[0] + (p[1] - p[0]) u + (p[2] - p[0]) v: hull of the clipped polygon's corners
Iv3 param_box(const palz_float3* p, float u0, float u1, float v0, float v1) {
    const Iv3 base = point3(p[0]), e0 = point3(p[1]) - base, e1 = point3(p[2]) - base;
    const float pts[8][2] = {{u0, v0}, {u1, v0}, {u0, v1}, {u1, v1},
                             {1.0f - v0, v0}, {1.0f - v1, v1}, {u0, 1.0f - u0}, {u1, 1.0f - u1}};
    Iv3 r = {{{kInf, -kInf}, {kInf, -kInf}, {kInf, -kInf}}};
    for (auto& q : pts) {
        const float u = q[0], v = q[1];
        if (u < u0 || u > u1 || v < v0 || v > v1 || u + v > 1.0f) continue;
        const Iv3 x = base + scale(e0, Iv::point(u)) + scale(e1, Iv::point(v));
        for (int a = 0; a < 3; ++a) r.c[a] = {std::fmin(r.c[a].l, x.c[a].l), std::fmax(r.c[a].u, x.c[a].u)};
    }
    return r;
}

bool admissible(const palz_triangle* tris, const palz_triangle_attrib* attribs, const palz_path& path,
                const float* p_light, const float* p_end, Iv eta) {
    const int K = (int)(path.event >> 24);
    if (K < 1 || K > 3) return true;
    // splits: a triangle much larger than the chain's smallest one is cut into
    // S x S parameter rectangles, S a power of two <= 8, at most 64 combinations
    int split[3];
    float ext[3], ext_min = kInf;
    for (int k = 0; k < K; ++k) {
        const Iv3 h = hull(tris[path.tris[k]].vs);
        ext[k] = 0.0f;
        for (auto& x : h.c) ext[k] = std::fmax(ext[k], x.u - x.l);
        ext_min = std::fmin(ext_min, ext[k]);
    }
    int combos = 1;
    for (int k = 0; k < K; ++k) {
        int sk = 1;
        while (sk < 8 && ext[k] > 2.0f * sk * ext_min) sk *= 2;
        split[k] = sk;
        combos *= sk * sk;
    }
    while (combos > 64) {
        int kk = 0;
        for (int k = 1; k < K; ++k)
            if (split[k] > split[kk]) kk = k;
        combos /= 4;
        split[kk] /= 2;
    }
    for (int c = 0; c < combos; ++c) {
        Iv3 v[5], n[3];
        v[0] = {{Iv::point(p_light[0]), Iv::point(p_light[1]), Iv::point(p_light[2])}};
        v[K + 1] = {{Iv::point(p_end[0]), Iv::point(p_end[1]), Iv::point(p_end[2])}};
        int rest = c;
        bool inside = true;
        for (int k = 0; k < K; ++k) {
            const int sk = split[k];
            const int iu = rest % sk, iv = (rest / sk) % sk;
            rest /= sk * sk;
            const float w = 1.0f / (float)sk;
            const float u0 = iu * w, u1 = (iu + 1) * w, v0 = iv * w, v1 = (iv + 1) * w;
            if (u0 + v0 > 1.0f) inside = false;
            v[k + 1] = param_box(tris[path.tris[k]].vs, u0, u1, v0, v1);
            n[k] = param_box(attribs[path.tris[k]].shading_normals, u0, u1, v0, v1);
        }
        if (!inside) continue;
        bool ok = true;
        for (int k = 0; k < K && ok; ++k)
            ok = vertex_ok(v[k] - v[k + 1], v[k + 2] - v[k + 1], n[k], (path.event >> k) & 1u, eta);
        if (ok) return true;
    }
    return false;
}

// ---- node-level path cuts (traverseAdmissibleNodes, pk.h:1760-2212) ----

Iv3 aabb_iv(const palz_aabb& b) { return {{{b.lower.x, b.upper.x}, {b.lower.y, b.upper.y}, {b.lower.z, b.upper.z}}}; }

Iv3 union_iv(const palz_aabb& a, const palz_aabb& b) {
    return {{{box_min(a.lower.x, b.lower.x), box_max(a.upper.x, b.upper.x)},
             {box_min(a.lower.y, b.lower.y), box_max(a.upper.y, b.upper.y)},
             {box_min(a.lower.z, b.lower.z), box_max(a.upper.z, b.upper.z)}}};
}

// AABB::surfaceArea (minimum_lbvh.h:256-260) of the union of a node's children
float node_area(const palz_internal_node& n) {
    const Iv3 u = union_iv(n.aabbs[0], n.aabbs[1]);
    const float sx = u.c[0].u - u.c[0].l, sy = u.c[1].u - u.c[1].l, sz = u.c[2].u - u.c[2].l;
    return (sx * sy + sy * sz + sz * sx) * 2.0f;
}

// pk.h:1850-1875: p not behind the triangle, its normal turned to shading normal 0
bool front_side(const palz_triangle& t, const palz_triangle_attrib& a, const float* p) {
    const float e1[3] = {t.vs[1].x - t.vs[0].x, t.vs[1].y - t.vs[0].y, t.vs[1].z - t.vs[0].z};
    const float e2[3] = {t.vs[2].x - t.vs[0].x, t.vs[2].y - t.vs[0].y, t.vs[2].z - t.vs[0].z};
    float ng[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    const palz_float3& s = a.shading_normals[0];
    const float d = ng[0] * s.x + ng[1] * s.y + ng[2] * s.z;
    for (float& c : ng) c = c * d;
    const float lhs = ng[0] * p[0] + ng[1] * p[1] + ng[2] * p[2];
    const float rhs = ng[0] * t.vs[0].x + ng[1] * t.vs[0].y + ng[2] * t.vs[0].z;
    return !(lhs < rhs);
}

bool boxes_ok(const Iv3* vb, const Iv3* nb, int K, const float* pb, const float* pe, uint32_t event, Iv eta) {
    Iv3 v[5];
    v[0] = {{Iv::point(pb[0]), Iv::point(pb[1]), Iv::point(pb[2])}};
    v[K + 1] = {{Iv::point(pe[0]), Iv::point(pe[1]), Iv::point(pe[2])}};
    for (int k = 0; k < K; ++k) v[k + 1] = vb[k];
    for (int k = 0; k < K; ++k)
        if (!vertex_ok(v[k] - v[k + 1], v[k + 2] - v[k + 1], nb[k], (event >> k) & 1u, eta)) return false;
    return true;
}

struct CutScene {
    const palz_triangle* tris;
    const palz_triangle_attrib* attribs;
    const palz_internal_node* nodes;
    const palz_normal_bound* nb;
};

// One query: the tuples in the order the reference's callback receives them.
template <class Emit>
void cut_walk(const CutScene& sc, uint32_t root, int K, const float* pb, const float* pe, uint32_t event, Iv eta,
              Emit emit) {
    const uint32_t kLeaf = PALZ_NODE_LEAF_BIT;
    struct Entry { uint32_t n[3]; };
    std::vector<Entry> stack;
    Entry cur = {{root, root, root}};
    while (true) {
        int cut = -1;
        float best = -1.0f;
        for (int k = 0; k < K; ++k) {
            if (cur.n[k] & kLeaf) continue;
            const float a = node_area(sc.nodes[cur.n[k]]);
            if (best < a) {
                best = a;
                cut = k;
            }
        }
        if (cut < 0) {
            int idx[3];
            for (int k = 0; k < K; ++k) idx[k] = (int)(cur.n[k] & ~kLeaf);
            bool ok = true;
            for (int k = 0; k + 1 < K; ++k)
                if (idx[k] == idx[k + 1]) ok = false;
            if (ok && !front_side(sc.tris[idx[0]], sc.attribs[idx[0]], pb)) ok = false;
            if (ok && !front_side(sc.tris[idx[K - 1]], sc.attribs[idx[K - 1]], pe)) ok = false;
            if (ok) emit(idx);
        } else {
            Iv3 vb[3], nbx[3];
            for (int k = 0; k < K; ++k) {
                if (k == cut) continue;
                const uint32_t n = cur.n[k];
                if (n & kLeaf) {
                    vb[k] = hull(sc.tris[n & ~kLeaf].vs);
                    nbx[k] = hull(sc.attribs[n & ~kLeaf].shading_normals);
                } else {
                    vb[k] = union_iv(sc.nodes[n].aabbs[0], sc.nodes[n].aabbs[1]);
                    nbx[k] = union_iv(sc.nb[n].normal_bounds[0], sc.nb[n].normal_bounds[1]);
                }
            }
            const uint32_t c = cur.n[cut];
            for (int i = 0; i < 2; ++i) {
                vb[cut] = aabb_iv(sc.nodes[c].aabbs[i]);
                nbx[cut] = aabb_iv(sc.nb[c].normal_bounds[i]);
                if (boxes_ok(vb, nbx, K, pb, pe, event, eta)) {
                    Entry e = cur;
                    e.n[cut] = sc.nodes[c].children[i];
                    stack.push_back(e);
                }
            }
        }
        if (stack.empty()) break;
        cur = stack.back();
        stack.pop_back();
    }
}

}  // namespace

extern "C" {

// op: 0 add, 1 sub, 2 mul, 3 div, 4 sqrt (of a); out = {rounded down, rounded up};
// 5: the interval square of [a, b]
void orc_iv_op(int op, float a, float b, float* out) {
    if (op == 5) {
        const Iv r = sq(Iv{a, b});
        out[0] = r.l;
        out[1] = r.u;
        return;
    }
    switch (op) {
        case 0: out[0] = r_add(a, b, Dir::down); out[1] = r_add(a, b, Dir::up); break;
        case 1: out[0] = r_add(a, -b, Dir::down); out[1] = r_add(a, -b, Dir::up); break;
        case 2: out[0] = r_mul(a, b, Dir::down); out[1] = r_mul(a, b, Dir::up); break;
        case 3: out[0] = r_div(a, b, Dir::down); out[1] = r_div(a, b, Dir::up); break;
        default: out[0] = r_sqrt(a, Dir::down); out[1] = r_sqrt(a, Dir::up); break;
    }
}

// Interval operations on {l, u} pairs and 6-float boxes {x.l, x.u, y.l, y.u,
// z.l, z.u}, in the operand layout of the reference's interval.h: op 0 +,
// 1 -, 2 *, 3 /, 4 sqrt(a), 5 square(a), 6 dot, 7 cross (out[0..5]),
// 8 lengthSquared(a).
void orc_iv_ops(int op, const float* a, const float* b, float* out) {
    auto box = [](const float* p) { return Iv3{{{p[0], p[1]}, {p[2], p[3]}, {p[4], p[5]}}}; };
    Iv r;
    switch (op) {
        case 0: r = Iv{a[0], a[1]} + Iv{b[0], b[1]}; break;
        case 1: r = Iv{a[0], a[1]} - Iv{b[0], b[1]}; break;
        case 2: r = Iv{a[0], a[1]} * Iv{b[0], b[1]}; break;
        case 3: r = Iv{a[0], a[1]} / Iv{b[0], b[1]}; break;
        case 4: r = root(Iv{a[0], a[1]}); break;
        case 5: r = sq(Iv{a[0], a[1]}); break;
        case 6: r = dot3(box(a), box(b)); break;
        case 7: {
            const Iv3 c = cross3(box(a), box(b));
            for (int i = 0; i < 3; ++i) out[2 * i] = c.c[i].l, out[2 * i + 1] = c.c[i].u;
            return;
        }
        default: r = norm2(box(a)); break;
    }
    out[0] = r.l;
    out[1] = r.u;
}

// Enclosure of the residual c_k = cross(wo, wopt) of one oriented vertex over
// boxes (6 floats each: lo xyz, hi xyz): out = 3 x {l, u}; returns 0 when the
// refraction box may hold total internal reflection (no enclosure computed).
int orc_iv_residual(const float* wi, const float* wo, const float* n, int refract, float eta_lo, float eta_hi,
                    float* out) {
    auto mk = [](const float* b) { return Iv3{{{b[0], b[3]}, {b[1], b[4]}, {b[2], b[5]}}}; };
    const Iv3 WI = mk(wi), WO = mk(wo), N = mk(n);
    const Iv eta = {eta_lo, eta_hi};
    const Iv non = norm2(N), win = dot3(WI, N);
    Iv3 wopt;
    if (refract) {
        const Iv k = non * norm2(WI) * (sq(eta) - Iv::point(1.0f)) + sq(win);
        if (!(k.l >= 0.0f)) return 0;
        wopt = scale(-WI, non) + scale(N, win - root(k));
    } else {
        wopt = scale(N, (win * Iv::point(2.0f)) / non) - WI;
    }
    const Iv3 c = cross3(WO, wopt);
    for (int i = 0; i < 3; ++i) {
        out[2 * i] = c.c[i].l;
        out[2 * i + 1] = c.c[i].u;
    }
    return 1;
}

// The per-tuple culling of n paths of one scene: out[p] = 1 if admissible.
void orc_interval_cull(const palz_triangle* tris, const palz_triangle_attrib* attribs, const palz_path* paths,
                       uint32_t n, const float* p_light, const palz_first_diffuse* points, float eta_lo,
                       float eta_hi, uint8_t* out) {
    for (uint32_t p = 0; p < n; ++p) {
        const palz_first_diffuse& q = points[paths[p].pixel];
        const float pe[3] = {q.p.x, q.p.y, q.p.z};
        out[p] = admissible(tris, attribs, paths[p], p_light, pe, {eta_lo, eta_hi}) ? 1 : 0;
    }
}

// Path-cuts test of one tuple of boxes (K x 6 floats each: lo xyz, hi xyz):
// 1 if every vertex of light -> boxes -> p_end passes.
int orc_tuple_boxes_ok(const float* vb, const float* nb, int K, const float* p_beg, const float* p_end,
                       uint32_t event, float eta_lo, float eta_hi) {
    Iv3 v[3], n[3];
    for (int k = 0; k < K; ++k) {
        const float* b = vb + 6 * k;
        const float* m = nb + 6 * k;
        v[k] = {{{b[0], b[3]}, {b[1], b[4]}, {b[2], b[5]}}};
        n[k] = {{{m[0], m[3]}, {m[1], m[4]}, {m[2], m[5]}}};
    }
    return boxes_ok(v, n, K, p_beg, p_end, event, {eta_lo, eta_hi}) ? 1 : 0;
}

// traverseAdmissibleNodes<K> (pk.h:1760-2212) for every receiver point with a
// diffuse hit: counts[q] = its tuple count; the tuples (K ints each, points in
// order, each point's tuples in traversal order) fill `tuples` up to
// `capacity` tuples.  Returns the total.
uint64_t orc_path_cuts(const palz_triangle* tris, const palz_triangle_attrib* attribs, const palz_internal_node* nodes,
                       uint32_t root, const palz_normal_bound* nb, const palz_first_diffuse* points, uint32_t n_points,
                       const float* p_light, uint32_t event, float eta_lo, float eta_hi, uint32_t* counts,
                       int32_t* tuples, uint64_t capacity) {
    const int K = (int)(event >> 24);
    const CutScene sc = {tris, attribs, nodes, nb};
    uint64_t total = 0;
    for (uint32_t q = 0; q < n_points; ++q) {
        const palz_first_diffuse& pt = points[q];
        uint32_t found = 0;
        if (!(pt.R.x == 0.0f && pt.R.y == 0.0f && pt.R.z == 0.0f)) {
            const float pe[3] = {pt.p.x, pt.p.y, pt.p.z};
            cut_walk(sc, root, K, p_light, pe, event, {eta_lo, eta_hi}, [&](const int* idx) {
                if (total < capacity)
                    for (int k = 0; k < K; ++k) tuples[total * K + k] = idx[k];
                ++total;
                ++found;
            });
        }
        counts[q] = found;
    }
    return total;
}

}  // extern "C"
