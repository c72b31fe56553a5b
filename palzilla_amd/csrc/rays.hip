// Ray queries against a scene's BVH: the reference's three traversals as
// one entry point (minimum_lbvh.h:905-1076).  One lane per ray.
//   PALZ_QUERY_CLOSEST            intersect_stackfull: near-first closest
//                                 hit with a per-lane stack in LDS (a ray whose
//                                 walk would overflow it finishes with the
//                                 stackless walk, which finds the same nearest
//                                 hit distance)
//   PALZ_QUERY_CLOSEST_STACKFREE  intersect_stackfree (RAY_QUERY_CLOSEST): the
//                                 parent-pointer walk the reference's CPU
//                                 renderer uses
//   PALZ_QUERY_ANY                intersect_stackfree (RAY_QUERY_ANY): the
//                                 first triangle found within t_max
#include "optics.hpp"

namespace palz {

namespace {

constexpr int kRayThreads = 128;
constexpr int kRayStack = 64;     // traversal stack entries per ray
constexpr int kRayStackLds = 16;  // of which in LDS; the rest in private memory

__global__ void __launch_bounds__(kRayThreads)
    k_intersect_rays(palz_scene_view sc, const float* __restrict__ ro3, const float* __restrict__ rd3, uint32_t n,
                     int query, float t_max, palz_hit* __restrict__ out) {
    __shared__ uint32_t s_stack[kRayStackLds * kRayThreads];
    const uint32_t i = blockIdx.x * kRayThreads + threadIdx.x;
    if (i >= n) return;
    const v3 ro = mk(ro3[3 * i], ro3[3 * i + 1], ro3[3 * i + 2]);
    const v3 rd = mk(rd3[3 * i], rd3[3 * i + 1], rd3[3 * i + 2]);
    const v3 ird = inv_rd(rd);
    const uint32_t root = *sc.root;
    Hit hit;
    hit.t = t_max;
    hit.u = hit.v = 0.0f;
    hit.ng = mk(0.0f, 0.0f, 0.0f);
    hit.tri = 0xFFFFFFFFu;
    if (query == PALZ_QUERY_CLOSEST) {
        Hit h = hit;
        SplitStack<kRayStackLds, kRayStack - kRayStackLds> stack;
        stack.near = StridedStack{s_stack + threadIdx.x, kRayThreads};
        if (closest_hit_on(&h, sc.nodes, sc.tris, root, ro, rd, ird, stack, kRayStack))
            hit = h;
        else
            closest_hit_stackfree(&hit, sc.nodes, sc.tris, root, ro, rd, ird, false);
    } else {
        closest_hit_stackfree(&hit, sc.nodes, sc.tris, root, ro, rd, ird, query == PALZ_QUERY_ANY);
    }
    palz_hit r;
    r.t = hit.t;
    r.u = hit.u;
    r.v = hit.v;
    r.ng = {hit.ng.x, hit.ng.y, hit.ng.z};
    r.tri = hit.tri;
    out[i] = r;
}

}  // namespace

}  // namespace palz

extern "C" int palz_intersect_rays(const palz_scene_view* view, const float* d_ro, const float* d_rd, uint32_t n,
                                   int query, float t_max, palz_hit* d_hits, void* stream) {
    using namespace palz;
    if (!view || (n && (!d_ro || !d_rd || !d_hits)))
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_intersect_rays: NULL argument");
    if (query < PALZ_QUERY_CLOSEST || query > PALZ_QUERY_ANY)
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_intersect_rays: unknown query");
    if (!(t_max >= 0.0f)) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_intersect_rays: t_max must be >= 0");
    if (n == 0) return PALZ_OK;
    if (!view->tris || !view->root || view->n_tris == 0 || (view->n_tris > 1 && !view->nodes))
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_intersect_rays: the scene needs triangles and a BVH");
    const uint32_t blocks = (n + kRayThreads - 1) / kRayThreads;
    k_intersect_rays<<<blocks, kRayThreads, 0, (hipStream_t)stream>>>(*view, d_ro, d_rd, n, query, t_max, d_hits);
    PALZ_LAUNCH_CHECK();
    return PALZ_OK;
}
