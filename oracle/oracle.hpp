// CPU oracle of the caustic hot path — TEST INFRASTRUCTURE ONLY.
//
// A plain scalar C++ restatement of the reference algorithms, written from the
// reference source text (each function cites the file:line it follows).  It is
// used by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline
// leg as the checker and the CPU baseline; nothing in the product library
// links, loads or calls it.
//
// Floating-point rules (they make the FP32 functions comparable bit for bit
// with the gfx950 kernels): compiled with -ffp-contract=off and without
// fast-math, so + - * / sqrt are single IEEE operations in program order.
// Bounding boxes use the canonical min/max (-0 below +0), which equals the
// reference's a<b?a:b (helper_math.h:134-142) for every input without mixed
// signed zeros and does not depend on evaluation order.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>

#include "../include/palz.h"

namespace orc {

template <class T>
struct vec3 {
    T x, y, z;
};
template <class T>
inline vec3<T> operator+(vec3<T> a, vec3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T>
inline vec3<T> operator-(vec3<T> a, vec3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T>
inline vec3<T> operator-(vec3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T>
inline vec3<T> operator*(vec3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T>
inline vec3<T> operator/(vec3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T>
inline T dot(vec3<T> a, vec3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T>
inline vec3<T> cross(vec3<T> a, vec3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
using float3 = vec3<float>;
using double3 = vec3<double>;

inline float3 load3(const palz_float3& p) { return {p.x, p.y, p.z}; }

inline float box_min(float a, float b) {
    if (a < b) return a;
    if (b < a) return b;
    return std::signbit(a) ? a : b;
}
inline float box_max(float a, float b) {
    if (a < b) return b;
    if (b < a) return a;
    return std::signbit(a) ? b : a;
}

}  // namespace orc
