// The box test's exactness, checked on the CPU against the host build of
// optics.hpp:
//   1. slabs() folds with fminf/fmaxf; the reference folds with a < b ? a : b
//      (minimum_lbvh.h:82-96, restated below as ref_slabs).  Every box
//      decision (rmin <= rmax) must agree, and rmin, rmax must be equal up to
//      the sign of a zero.
//   2. slabs_ordered() takes each axis' entry and exit planes picked by the
//      sign of the ray's inverse direction (the packet walks' plane layout);
//      its decisions and values must equal slabs()'s the same way.
// Rays and boxes are drawn at random with the edge cases mixed in: zero,
// signed-zero, denormal and huge direction components (inv_rd clamps them),
// degenerate boxes, origins on a box plane, t_max of 1 and of FLT_MAX.
//
//   slab_check N        N trials; prints "trials <N> fold <m1> ordered <m2>"
//   slab_check N swap   the same with entry and exit swapped on the x axis of
//                       slabs_ordered's input (a planted mistake that must show)
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "optics.hpp"

using palz::mk;
using palz::v3;

namespace {

uint64_t g_state = 0x853c49e6748fea9bull;
uint32_t next_u32() {  // xorshift64*
    g_state ^= g_state >> 12;
    g_state ^= g_state << 25;
    g_state ^= g_state >> 27;
    return (uint32_t)((g_state * 0x2545F4914F6CDD1Dull) >> 32);
}
float uniform(float lo, float hi) { return lo + (hi - lo) * (float)(next_u32() >> 8) * (1.0f / 16777216.0f); }

float direction_component() {
    switch (next_u32() % 10u) {
        case 0: return 0.0f;
        case 1: return -0.0f;
        case 2: return (next_u32() & 1u) ? 1e-40f : -1e-40f;  // denormal
        case 3: return (next_u32() & 1u) ? 1e30f : -1e30f;
        case 4: return (next_u32() & 1u) ? 1e-20f : -1e-20f;
        default: return uniform(-2.0f, 2.0f);
    }
}

void ref_slabs(v3 ro, v3 ird, v3 lo, v3 hi, float known_t, float* rmin, float* rmax) {
    auto sel_min = [](float a, float b) { return a < b ? a : b; };
    auto sel_max = [](float a, float b) { return a > b ? a : b; };
    const v3 t0 = (lo - ro) * ird;
    const v3 t1 = (hi - ro) * ird;
    const v3 tmin = {sel_min(t0.x, t1.x), sel_min(t0.y, t1.y), sel_min(t0.z, t1.z)};
    const v3 tmax = {sel_max(t0.x, t1.x), sel_max(t0.y, t1.y), sel_max(t0.z, t1.z)};
    float a = sel_max(sel_max(tmin.x, tmin.y), tmin.z);
    float b = sel_min(sel_min(tmax.x, tmax.y), tmax.z) * 1.00000024f;
    a = sel_max(a, 0.0f);
    b = sel_min(b, known_t);
    *rmin = a;
    *rmax = b;
}

bool same(float a, float b) { return a == b || (std::isnan(a) && std::isnan(b)); }

}  // namespace

int main(int argc, char** argv) {
    const long n = argc > 1 ? std::atol(argv[1]) : 1000000;
    const bool swap = argc > 2 && std::strcmp(argv[2], "swap") == 0;
    long fold_bad = 0, ordered_bad = 0;
    for (long i = 0; i < n; ++i) {
        v3 lo = mk(uniform(-10.f, 10.f), uniform(-10.f, 10.f), uniform(-10.f, 10.f));
        v3 ext = mk(uniform(0.f, 5.f), uniform(0.f, 5.f), uniform(0.f, 5.f));
        const uint32_t shape = next_u32() % 8u;
        if (shape == 0) ext.x = 0.0f;                 // flat box
        if (shape == 1) ext = mk(0.f, 0.f, 0.f);      // a point
        if (shape == 2) ext.y = 1e-6f;
        const v3 hi = lo + ext;
        v3 ro = mk(uniform(-20.f, 20.f), uniform(-20.f, 20.f), uniform(-20.f, 20.f));
        const uint32_t where = next_u32() % 8u;
        if (where == 0) ro.x = lo.x;                  // on a plane
        if (where == 1) ro = lo + ext * 0.5f;          // inside
        if (where == 2) ro.z = hi.z;
        const v3 rd = mk(direction_component(), direction_component(), direction_component());
        const v3 ird = palz::inv_rd(rd);
        const float t_max = (next_u32() & 1u) ? 1.0f : FLT_MAX;

        float ra, rb, fa, fb, oa, ob;
        ref_slabs(ro, ird, lo, hi, t_max, &ra, &rb);
        palz::slabs(ro, ird, lo, hi, t_max, &fa, &fb);
        // entry plane: lower where ird > 0, upper where ird < 0 (inv_rd never gives 0)
        v3 nr = mk(ird.x < 0.0f ? hi.x : lo.x, ird.y < 0.0f ? hi.y : lo.y, ird.z < 0.0f ? hi.z : lo.z);
        v3 fr = mk(ird.x < 0.0f ? lo.x : hi.x, ird.y < 0.0f ? lo.y : hi.y, ird.z < 0.0f ? lo.z : hi.z);
        if (swap) {
            const float t = nr.x;
            nr.x = fr.x;
            fr.x = t;
        }
        palz::slabs_ordered(ro, ird, nr, fr, t_max, &oa, &ob);
        if ((ra <= rb) != (fa <= fb) || !same(ra, fa) || !same(rb, fb)) ++fold_bad;
        if ((fa <= fb) != (oa <= ob) || !same(fa, oa) || !same(fb, ob)) ++ordered_bad;
    }
    std::printf("trials %ld fold %ld ordered %ld\n", n, fold_bad, ordered_bad);
    return 0;
}
