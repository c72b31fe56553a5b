// Host checker of the affine enclosures (palzilla_amd/csrc/affine.hpp),
// built by tests/test_affine.py with hipcc as host code.
//
// For random chains (1-3 vertices on random triangles, random shading
// normals, a refractive index range) and random parameter rectangles, the
// affine residual of every oriented vertex is computed, and the residual is
// evaluated in float64 at random points of the rectangles and index range
// where the orientation holds and the refraction exists: every sample must
// lie inside the affine ranges.  The same residual in the interval
// arithmetic of interval.hpp gives the width ratio printed at the end.
// Prints "cases samples violations mean_width_ratio".
#include <cmath>
#include <cstdio>
#include <random>

#include "affine.hpp"

using namespace palz;

namespace {

struct D3 {
    double x, y, z;
};
D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
D3 to64(v3 p) { return {p.x, p.y, p.z}; }

// the solver's residual of one oriented vertex in float64; false where the
// refraction does not exist (k < 0)
bool residual64(D3 wi, D3 wo, D3 n, bool refract, double eta, D3* c) {
    D3 wopt;
    const double non = dot(n, n), win = dot(wi, n);
    if (!refract) {
        wopt = n * (win * 2.0 / non) - wi;
    } else {
        const double k = non * dot(wi, wi) * (eta * eta - 1.0) + win * win;
        if (k < 0.0) return false;
        wopt = wi * (-non) + n * (win - std::sqrt(k));
    }
    *c = cross(wo, wopt);
    return true;
}

float width(const iv::I& r) { return r.u - r.l; }

}  // namespace

int main(int argc, char** argv) {
    const int n_cases = argc > 1 ? std::atoi(argv[1]) : 3000;
    std::mt19937_64 rng(2026);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    long samples = 0, violations = 0, cases = 0, compared = 0;
    double ratio_sum = 0.0;
    for (int cs = 0; cs < n_cases; ++cs) {
        const int K = 1 + cs % 3;
        // a stack of triangles between a light above and a receiver point below
        ChainTri tri[3];
        for (int k = 0; k < K; ++k) {
            const double y = 1.0 - 0.3 * k + 0.05 * (U(rng) - 0.5);
            const double size = std::pow(10.0, -2.5 + 2.0 * U(rng));
            const double cx = 0.4 * (U(rng) - 0.5), cz = 0.4 * (U(rng) - 0.5);
            auto vtx = [&]() {
                return mk((float)(cx + size * (U(rng) - 0.5)), (float)(y + 0.2 * size * (U(rng) - 0.5)),
                          (float)(cz + size * (U(rng) - 0.5)));
            };
            auto nrm = [&]() {
                return mk((float)(0.3 * (U(rng) - 0.5)), (float)(0.8 + 0.4 * U(rng)), (float)(0.3 * (U(rng) - 0.5)));
            };
            tri[k].v0 = vtx();
            tri[k].v1 = vtx();
            tri[k].v2 = vtx();
            tri[k].s0 = nrm();
            tri[k].s1 = nrm();
            tri[k].s2 = nrm();
            tri[k].material = PALZ_MATERIAL_DIELECTRIC;
        }
        const v3 pb = mk((float)(0.3 * (U(rng) - 0.5)), 2.0f, (float)(0.3 * (U(rng) - 0.5)));
        const v3 pe = mk((float)(U(rng) - 0.5), 0.0f, (float)(U(rng) - 0.5));
        const uint32_t event = (uint32_t)(rng() & 7u) | ((uint32_t)K << 24);
        const float eta_lo = (float)(1.3 + 0.3 * U(rng));
        const float eta_hi = cs % 4 == 0 ? eta_lo : eta_lo + (float)(0.05 * U(rng));
        // one rectangle per triangle: side 1, 1/2, 1/4 or 1/8
        float um[3], vm[3], hr[3];
        for (int k = 0; k < K; ++k) {
            const int sk = 1 << (rng() % 4);
            const int iu = (int)(rng() % sk), iv_ = (int)(rng() % sk);
            hr[k] = 0.5f / (float)sk;
            um[k] = (float)iu / (float)sk + hr[k];
            vm[k] = (float)iv_ / (float)sk + hr[k];
        }
        aa::A eta = aa::pt(0.5f * eta_lo + 0.5f * eta_hi);
        eta.a[aa::kSym - 1] = pmax(iv::add_ru(eta_hi, -eta.c), iv::add_ru(eta.c, -eta_lo));
        aa::A3 vert[5], nrm[3];
        vert[0] = aa::pt3(pb);
        vert[K + 1] = aa::pt3(pe);
        iv::I3 vbox[5], nbox[3];
        vbox[0] = iv::pt3(pb);
        vbox[K + 1] = iv::pt3(pe);
        for (int k = 0; k < K; ++k) {
            vert[k + 1] = aa::affine_point(tri[k].v0, tri[k].v1, tri[k].v2, um[k], hr[k], vm[k], hr[k], 2 * k, 2 * k + 1);
            nrm[k] = aa::affine_point(tri[k].s0, tri[k].s1, tri[k].s2, um[k], hr[k], vm[k], hr[k], 2 * k, 2 * k + 1);
            const float u0 = um[k] - hr[k], u1 = um[k] + hr[k], v0 = vm[k] - hr[k], v1 = vm[k] + hr[k];
            // the interval boxes of the whole rectangle (its four corners)
            iv::I3 b = {{iv::kInf, -iv::kInf}, {iv::kInf, -iv::kInf}, {iv::kInf, -iv::kInf}};
            iv::I3 nb = b;
            for (int c = 0; c < 4; ++c) {
                const float cu = (c & 1) ? u1 : u0, cv = (c & 2) ? v1 : v0;
                const iv::I3 a = iv::pt3(tri[k].v0), e0 = iv::pt3(tri[k].v1) - a, e1 = iv::pt3(tri[k].v2) - a;
                const iv::I3 q = a + e0 * iv::pt(cu) + e1 * iv::pt(cv);
                const iv::I3 sa = iv::pt3(tri[k].s0), se0 = iv::pt3(tri[k].s1) - sa, se1 = iv::pt3(tri[k].s2) - sa;
                const iv::I3 qn = sa + se0 * iv::pt(cu) + se1 * iv::pt(cv);
                b.x = {fminf(b.x.l, q.x.l), fmaxf(b.x.u, q.x.u)};
                b.y = {fminf(b.y.l, q.y.l), fmaxf(b.y.u, q.y.u)};
                b.z = {fminf(b.z.l, q.z.l), fmaxf(b.z.u, q.z.u)};
                nb.x = {fminf(nb.x.l, qn.x.l), fmaxf(nb.x.u, qn.x.u)};
                nb.y = {fminf(nb.y.l, qn.y.l), fmaxf(nb.y.u, qn.y.u)};
                nb.z = {fminf(nb.z.l, qn.z.l), fmaxf(nb.z.u, qn.z.u)};
            }
            vbox[k + 1] = b;
            nbox[k] = nb;
        }
        for (int k = 0; k < K; ++k) {
            const bool refract = (event >> k) & 1u;
            for (int orient = 0; orient < 2; ++orient) {
                aa::A3 wi = vert[k] - vert[k + 1], wo = vert[k + 2] - vert[k + 1];
                if (orient) {
                    const aa::A3 t = wi;
                    wi = wo;
                    wo = t;
                }
                aa::A3 c;
                float scale;
                if (!aa::residual_af(wi, wo, nrm[k], refract, eta, &c, &scale)) continue;
                ++cases;
                const iv::I rx = aa::range(c.x), ry = aa::range(c.y), rz = aa::range(c.z);
                // the interval residual of the same boxes, for the width ratio
                iv::I3 bwi = vbox[k] - vbox[k + 1], bwo = vbox[k + 2] - vbox[k + 1];
                if (orient) {
                    const iv::I3 t = bwi;
                    bwi = bwo;
                    bwo = t;
                }
                const iv::I non = iv::length_squared(nbox[k]), win = iv::dot(bwi, nbox[k]);
                iv::I3 wopt;
                bool iv_ok = true;
                if (!refract) {
                    wopt = nbox[k] * ((win * iv::pt(2.0f)) / non) - bwi;
                } else {
                    const iv::I kk = non * iv::length_squared(bwi) * (iv::square(iv::I{eta_lo, eta_hi}) - iv::pt(1.0f)) +
                                     iv::square(win);
                    iv_ok = kk.l >= 0.0f;
                    if (iv_ok) wopt = -bwi * non + nbox[k] * (win - iv::sqrt(kk));
                }
                if (iv_ok) {
                    const iv::I3 ci = iv::cross(bwo, wopt);
                    const double wa = (double)width(rx) + width(ry) + width(rz);
                    const double wb = (double)width(ci.x) + width(ci.y) + width(ci.z);
                    if (wb > 0.0 && std::isfinite(wa) && std::isfinite(wb)) {
                        ratio_sum += wa / wb;
                        ++compared;
                    }
                }
                for (int s = 0; s < 64; ++s) {
                    D3 x[5];
                    x[0] = to64(pb);
                    x[K + 1] = to64(pe);
                    double u = 0, v = 0;
                    for (int j = 0; j < K; ++j) {
                        const double uu = um[j] + hr[j] * (2.0 * U(rng) - 1.0);
                        const double vv = vm[j] + hr[j] * (2.0 * U(rng) - 1.0);
                        const D3 a = to64(tri[j].v0);
                        x[j + 1] = a + (to64(tri[j].v1) - a) * uu + (to64(tri[j].v2) - a) * vv;
                        if (j == k) u = uu, v = vv;
                    }
                    const D3 sa = to64(tri[k].s0);
                    const D3 n = sa + (to64(tri[k].s1) - sa) * u + (to64(tri[k].s2) - sa) * v;
                    D3 swi = x[k] - x[k + 1], swo = x[k + 2] - x[k + 1];
                    if (orient) {
                        const D3 t = swi;
                        swi = swo;
                        swo = t;
                    }
                    const double e = eta_lo + (eta_hi - eta_lo) * U(rng);
                    D3 r;
                    if (!residual64(swi, swo, n, refract, e, &r)) continue;
                    ++samples;
                    const double tol = 1e-12;
                    if (!(rx.l - tol <= r.x && r.x <= rx.u + tol && ry.l - tol <= r.y && r.y <= ry.u + tol &&
                          rz.l - tol <= r.z && r.z <= rz.u + tol)) {
                        if (violations < 5)
                            std::printf("violation: case %d vertex %d orient %d residual (%g %g %g) ranges "
                                        "[%g %g] [%g %g] [%g %g]\n",
                                        cs, k, orient, r.x, r.y, r.z, rx.l, rx.u, ry.l, ry.u, rz.l, rz.u);
                        ++violations;
                    }
                }
            }
        }
    }
    std::printf("%ld %ld %ld %.4f\n", cases, samples, violations, compared ? ratio_sum / compared : 0.0);
    return violations != 0;
}
