// palz_render: a C++ host program over libpalz.so's renderer, driving it the
// way the reference's main.cpp drives PKRenderer (load a frame, step each
// frame, resolve, write the image).  The mesh is placed over a diffuse floor
// as glass, lit from above and seen from above and in front.
//
// Usage: palz_render MESH.obj OUT.ppm [WIDTH HEIGHT FRAMES]
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "palz.h"

namespace {

void check(int rc, const char* what) {
    if (rc != PALZ_OK) {
        std::fprintf(stderr, "palz_render: %s failed: %s\n", what, palz_last_error());
        std::exit(1);
    }
}

palz_float3 sub(palz_float3 a, palz_float3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
palz_float3 scale(palz_float3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
palz_float3 cross(palz_float3 a, palz_float3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
palz_float3 normalize(palz_float3 a) { return scale(a, 1.0f / std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z)); }

// RayGenerator::lookat (camera.h:10-22)
palz_camera lookat(palz_float3 eye, palz_float3 center, palz_float3 up, float fovy, int w, int h) {
    const palz_float3 f = normalize(sub(center, eye));
    const palz_float3 s = normalize(cross(f, up));
    const palz_float3 u = cross(s, f);
    const float tan_y = std::tan(fovy * 0.5f);
    const float tan_x = tan_y / (float)h * (float)w;
    return {eye, scale(s, tan_x), scale(u, tan_y)};
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s MESH.obj OUT.ppm [WIDTH HEIGHT FRAMES]\n", argv[0]);
        return 2;
    }
    const int W = argc > 3 ? std::atoi(argv[3]) : 256;
    const int H = argc > 4 ? std::atoi(argv[4]) : 256;
    const int frames = argc > 5 ? std::atoi(argv[5]) : 4;
    if (W < 1 || H < 1 || frames < 1) {
        std::fprintf(stderr, "palz_render: bad size or frame count\n");
        return 2;
    }

    // loadFrame: the mesh as glass, then a diffuse floor below it
    palz_mesh* mesh = nullptr;
    check(palz_mesh_load_obj(argv[1], PALZ_MATERIAL_DIELECTRIC, &mesh), "palz_mesh_load_obj");
    uint32_t n = 0;
    check(palz_mesh_triangle_count(mesh, &n), "palz_mesh_triangle_count");
    std::vector<palz_triangle> tris(n + 2);
    std::vector<palz_triangle_attrib> attribs(n + 2);
    check(palz_mesh_copy(mesh, tris.data(), attribs.data()), "palz_mesh_copy");
    palz_mesh_destroy(mesh);
    palz_float3 lo = {1e30f, 1e30f, 1e30f}, hi = {-1e30f, -1e30f, -1e30f};
    for (uint32_t i = 0; i < n; ++i)
        for (const palz_float3& v : tris[i].vs) {
            lo = {std::min(lo.x, v.x), std::min(lo.y, v.y), std::min(lo.z, v.z)};
            hi = {std::max(hi.x, v.x), std::max(hi.y, v.y), std::max(hi.z, v.z)};
        }
    const float ext = std::max(hi.x - lo.x, hi.z - lo.z);
    const float cx = 0.5f * (lo.x + hi.x), cz = 0.5f * (lo.z + hi.z), fy = lo.y - 0.5f * ext;
    const float r = 1.5f * ext;
    const palz_float3 a = {cx - r, fy, cz - r}, b = {cx + r, fy, cz - r}, c = {cx + r, fy, cz + r},
                      d = {cx - r, fy, cz + r};
    tris[n] = {{a, d, c}};
    tris[n + 1] = {{a, c, b}};
    for (uint32_t i = n; i < n + 2; ++i) {
        attribs[i].material = PALZ_MATERIAL_DIFFUSE;
        for (palz_float3& s : attribs[i].shading_normals) s = {0.0f, 1.0f, 0.0f};
    }

    // 16 wavelengths over 380-720 nm with trapezoid-free uniform weights
    const int L = 16;
    std::vector<float> lambdas(L), weights(L, 340.0f / (L - 1));
    for (int l = 0; l < L; ++l) lambdas[l] = 380.0f + 340.0f * (float)l / (L - 1);
    palz_renderer* rd = nullptr;
    check(palz_renderer_create(W, H, lambdas.data(), weights.data(), L, &rd), "palz_renderer_create");
    palz_renderer_params p;
    check(palz_renderer_get_params(rd, &p), "palz_renderer_get_params");
    p.p_light = {cx + 0.1f * ext, hi.y + 1.5f * ext, cz + 0.05f * ext};
    p.photons_per_triangle = 64;
    p.use_lens = 0;
    p.albedo = {0.8f, 0.8f, 0.8f};
    p.path_capacity = 1u << 22;
    check(palz_renderer_set_params(rd, &p), "palz_renderer_set_params");
    const palz_camera cam = lookat({cx, hi.y + 1.2f * ext, cz + 1.6f * ext}, {cx, fy, cz}, {0.0f, 1.0f, 0.0f}, 0.8f, W, H);
    check(palz_renderer_set_camera(rd, &cam), "palz_renderer_set_camera");
    check(palz_renderer_load_scene(rd, tris.data(), attribs.data(), n + 2, nullptr), "palz_renderer_load_scene");

    const auto t0 = std::chrono::steady_clock::now();
    for (int f = 0; f < frames; ++f) {
        check(palz_renderer_step(rd, nullptr), "palz_renderer_step");
        check(palz_stream_sync(nullptr), "palz_stream_sync");
        std::printf("frame %d done\n", f);
        std::fflush(stdout);
    }
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::vector<uint32_t> rgba((size_t)W * H);
    check(palz_renderer_resolve(rd, rgba.data(), nullptr), "palz_renderer_resolve");
    palz_renderer_destroy(rd);

    FILE* out = std::fopen(argv[2], "wb");
    if (!out) {
        std::fprintf(stderr, "palz_render: cannot write %s\n", argv[2]);
        return 1;
    }
    std::fprintf(out, "P6\n%d %d\n255\n", W, H);
    for (uint32_t px : rgba) {
        const unsigned char rgb[3] = {(unsigned char)(px & 0xFF), (unsigned char)((px >> 8) & 0xFF),
                                      (unsigned char)((px >> 16) & 0xFF)};
        std::fwrite(rgb, 1, 3, out);
    }
    std::fclose(out);
    std::printf("%u triangles, %dx%d, %d frames in %.1f ms -> %s\n", n + 2, W, H, frames, ms, argv[2]);
    return 0;
}
