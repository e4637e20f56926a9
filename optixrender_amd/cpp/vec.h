// Vector/box types of the host facade.
//
// Built inside the reference's tree (common/gdt on the include path, as its
// CMakeLists.txt:13-24 sets up for every target), the facade uses gdt's own
// vec_t / box_t (gdt/math/AffineSpace.h, which the viewer's GLFWindow.h:7
// includes too), so main.cpp and the viewer see one set of types.  Built on
// its own, it defines the subset the facade and main.cpp use with the same
// component-wise semantics: vec2i, vec2f, vec3f, vec3i, box3f, dot, cross,
// normalize (v * 1/sqrt(dot(v, v)), gdt/math/vec.h:290-315), length, box
// center/span/extend.
#pragma once
#if __has_include("gdt/math/AffineSpace.h")
#include "gdt/math/AffineSpace.h"
#define OXR_FACADE_USES_GDT 1
#else
#include <algorithm>
#include <cmath>
#include <cstdint>

namespace gdt {

struct vec2i {
    int x = 0, y = 0;
    vec2i() = default;
    vec2i(int s) : x(s), y(s) {}
    vec2i(int a, int b) : x(a), y(b) {}
};
struct vec2f {
    float x = 0, y = 0;
    vec2f() = default;
    vec2f(float s) : x(s), y(s) {}
    vec2f(float a, float b) : x(a), y(b) {}
};
struct vec3i {
    int x = 0, y = 0, z = 0;
    vec3i() = default;
    vec3i(int a, int b, int c) : x(a), y(b), z(c) {}
};
struct vec3f {
    float x = 0, y = 0, z = 0;
    vec3f() = default;
    vec3f(float s) : x(s), y(s), z(s) {}
    vec3f(float a, float b, float c) : x(a), y(b), z(c) {}
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline vec3f operator+(vec3f a, vec3f b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline vec3f operator-(vec3f a, vec3f b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline vec3f operator*(vec3f a, vec3f b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline vec3f operator*(vec3f a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline vec3f operator*(float s, vec3f a) { return {s * a.x, s * a.y, s * a.z}; }
inline vec3f operator/(vec3f a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline vec3f operator/(vec3f a, vec3f b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
inline vec3f operator-(vec3f a) { return {-a.x, -a.y, -a.z}; }
inline float dot(vec3f a, vec3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline vec3f cross(vec3f a, vec3f b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
inline float length(vec3f v) { return std::sqrt(dot(v, v)); }
inline vec3f normalize(vec3f v) { return v * 1.f / std::sqrt(dot(v, v)); }

struct box3f {
    vec3f lower{1e30f}, upper{-1e30f};
    void extend(vec3f p) {
        lower = {std::min(lower.x, p.x), std::min(lower.y, p.y), std::min(lower.z, p.z)};
        upper = {std::max(upper.x, p.x), std::max(upper.y, p.y), std::max(upper.z, p.z)};
    }
    vec3f center() const { return (lower + upper) * 0.5f; }
    vec3f span() const { return upper - lower; }
};

}  // namespace gdt
#endif
