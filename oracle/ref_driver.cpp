// Driver of the reference's own standalone headers — TEST INFRASTRUCTURE.
//
// This file is ours; the three headers it includes are the reference's,
// compiled unmodified from where they lie (oracle/ref_build.py passes
// -I <reference root>) into oracle/_ref/libref.so:
//   sobol.h:17-129   shuffled_scrambled_sobol_2d (Owen-scrambled 2D Sobol)
//   saka.h:11-208    one-tangent forward-mode duals, reflection and
//                    refraction_norm_free
//   interval.h:34-713 round-to-nearest interval arithmetic
// The library is used only to write the golden vectors under tests/golden/
// and, where it is built, by the CPU tests that compare the oracle's
// restatements with the reference's own code.  Nothing in the product loads it.
#include <cfloat>
#include <cmath>
#include <cstdint>

#include "interval.h"
#include "saka.h"
#include "sobol.h"

namespace {

saka::dval3 load_dual(const float* p) {  // {v.x, v.y, v.z, g.x, g.y, g.z}
    saka::dval3 r;
    r.x.v = p[0], r.y.v = p[1], r.z.v = p[2];
    r.x.g = p[3], r.y.g = p[4], r.z.g = p[5];
    return r;
}
void store_dual(const saka::dval3& d, float* p) {
    p[0] = d.x.v, p[1] = d.y.v, p[2] = d.z.v;
    p[3] = d.x.g, p[4] = d.y.g, p[5] = d.z.g;
}
interval::intr3 load_box(const float* p) {  // {x.l, x.u, y.l, y.u, z.l, z.u}
    return {{p[0], p[1]}, {p[2], p[3]}, {p[4], p[5]}};
}
void store_box(const interval::intr3& b, float* p) {
    p[0] = b.x.l, p[1] = b.x.u, p[2] = b.y.l, p[3] = b.y.u, p[4] = b.z.l, p[5] = b.z.u;
}

}  // namespace

extern "C" {

// sobol.h:118-128
void ref_sobol2d(uint32_t index, uint32_t p0, uint32_t p1, uint32_t p2, float* out) {
    sobol::shuffled_scrambled_sobol_2d(out, out + 1, index, p0, p1, p2);
}

// saka.h:192-208: op 0 = reflection(wi, n), op 1 = refraction_norm_free(wi, n, eta)
void ref_saka_optics(int op, const float* wi, const float* n, float eta, float* out) {
    const saka::dval3 w = load_dual(wi), nn = load_dual(n);
    store_dual(op == 0 ? saka::reflection(w, nn) : saka::refraction_norm_free(w, nn, eta), out);
}

// interval.h: scalar ops on {l, u} pairs (op 0 +, 1 -, 2 *, 3 /, 4 sqrt(a),
// 5 square(a)) and vector ops on 6-float boxes (6 dot, 7 cross,
// 8 lengthSquared(a)); scalar results fill out[0..1], vector results out[0..5]
void ref_interval(int op, const float* a, const float* b, float* out) {
    using namespace interval;
    intr r;
    switch (op) {
        case 0: r = intr(a[0], a[1]) + intr(b[0], b[1]); break;
        case 1: r = intr(a[0], a[1]) - intr(b[0], b[1]); break;
        case 2: r = intr(a[0], a[1]) * intr(b[0], b[1]); break;
        case 3: r = intr(a[0], a[1]) / intr(b[0], b[1]); break;
        case 4: r = interval::sqrt(intr(a[0], a[1])); break;
        case 5: r = square(intr(a[0], a[1])); break;
        case 6: r = dot(load_box(a), load_box(b)); break;
        case 7: store_box(cross(load_box(a), load_box(b)), out); return;
        default: r = lengthSquared(load_box(a)); break;
    }
    out[0] = r.l;
    out[1] = r.u;
}

}  // extern "C"
