// Device math for the gfx950 path tracer: vectors, RNG, sampling helpers,
// pixel filters, software bilinear.  Every function rounds exactly as the
// reference's device code does when compiled without FMA contraction
// (the library is built with -ffp-contract=off): component-wise gdt vector
// semantics (common/gdt/gdt/math/vec/functors.h:143-199), CUDA float min/max
// (fminf/fmaxf), left-to-right RNG draws in argument lists.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define OXR_D __device__ __forceinline__
#define OXR_HD __host__ __device__ __forceinline__

namespace oxr {

constexpr float kEps = 1e-4f;                      // Utils.h:10
constexpr float kPi = 3.14159265358979323846f;     // Utils.h:11
constexpr float kTwoPi = 6.28318530717958647692f;  // Utils.h:14
constexpr float kInvPi = 0.318309886183790671538f; // Utils.h:13
constexpr float kFltMax = 3.402823466e+38f;
constexpr float kFltMin = 1.175494351e-38f;

struct f2 { float x, y; };
struct f3 {
    float x, y, z;
    OXR_HD f3() : x(0.0f), y(0.0f), z(0.0f) {}
    OXR_HD f3(float s) : x(s), y(s), z(s) {}
    OXR_HD f3(float a, float b, float c) : x(a), y(b), z(c) {}
    OXR_HD float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
OXR_HD f3 operator+(f3 a, f3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
OXR_HD f3 operator-(f3 a, f3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
OXR_HD f3 operator*(f3 a, f3 b) { return f3(a.x * b.x, a.y * b.y, a.z * b.z); }
OXR_HD f3 operator/(f3 a, f3 b) { return f3(a.x / b.x, a.y / b.y, a.z / b.z); }
OXR_HD f3 operator+(f3 a, float s) { return f3(a.x + s, a.y + s, a.z + s); }
OXR_HD f3 operator-(f3 a, float s) { return f3(a.x - s, a.y - s, a.z - s); }
OXR_HD f3 operator*(f3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
OXR_HD f3 operator/(f3 a, float s) { return f3(a.x / s, a.y / s, a.z / s); }
OXR_HD f3 operator-(float s, f3 a) { return f3(s - a.x, s - a.y, s - a.z); }
OXR_HD f3 operator*(float s, f3 a) { return f3(s * a.x, s * a.y, s * a.z); }
OXR_HD f3 operator-(f3 a) { return f3(-a.x, -a.y, -a.z); }
OXR_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
OXR_HD f3 cross(f3 a, f3 b) { return f3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y); }
OXR_D f3 normalize(f3 v) { return v * 1.f / sqrtf(dot(v, v)); }        // vec.h:306-309
OXR_D float length(f3 v) { return sqrtf(dot(v, v)); }
OXR_D bool any_nan(f3 v) { return isnan(v.x) || isnan(v.y) || isnan(v.z); }
OXR_HD f3 make3(const float* p) { return f3(p[0], p[1], p[2]); }
OXR_D f3 xyz(float4 v) { return f3(v.x, v.y, v.z); }

OXR_D float sqr(float x) { return x * x; }
OXR_D float mixf(float x, float y, float t) { return x * (1.0f - t) + y * t; }
OXR_D float clampf(float v, float lo, float hi) { return fminf(hi, fmaxf(lo, v)); }

// Utils.h:28-38
OXR_D f3 reflect(f3 v, f3 n) { return normalize(v - 2.0f * dot(v, n) * n); }
OXR_D f3 refract(f3 v, f3 n, float eta) {
    float ct = fminf(dot(-v, n), 1.0f);
    f3 perp = eta * (v + ct * n);
    f3 par = -sqrtf(fabsf(1.0f - dot(perp, perp))) * n;
    return normalize(perp + par);
}
// Utils.h:99-130: frame around `up` (C, B = C x up, up).
OXR_D void frame_of(f3 up, f3& B, f3& C) {
    if (fabsf(up.x) > fabsf(up.y)) {
        float li = 1.0f / sqrtf(up.x * up.x + up.z * up.z);
        C = f3(up.z * li, 0.0f, -up.x * li);
    } else {
        float li = 1.0f / sqrtf(up.y * up.y + up.z * up.z);
        C = f3(0.0f, up.z * li, -up.y * li);
    }
    B = cross(C, up);
}
OXR_D f3 to_local(f3 d, f3 up) { f3 B, C; frame_of(up, B, C); return f3(dot(d, B), dot(d, C), dot(d, up)); }
OXR_D f3 to_world(f3 d, f3 up) { f3 B, C; frame_of(up, B, C); return normalize(d.x * B + d.y * C + d.z * up); }

// gdt::LCG<16> (random.h:55-89): TEA with 16 rounds seeds a 32-bit LCG whose
// low 24 bits give the float.
OXR_HD uint32_t lcg_seed(uint32_t v0, uint32_t v1) {
    uint32_t s0 = 0;
#pragma unroll
    for (int n = 0; n < 16; n++) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}
struct Rng {
    uint32_t s;
    OXR_D float operator()() {
        s = 1664525u * s + 1013904223u;
        return (float)(s & 0x00FFFFFFu) * (1.0f / 16777216.0f);
    }
};

// Sampling.h
OXR_D float power_heuristic(float a, float b) {
    float p1 = powf(a, 2.0f), p2 = powf(b, 2.0f);
    return p1 / (p1 + p2);
}
OXR_D f3 plane_to_sphere(f2 uv) {
    float phi = kTwoPi * (uv.x - 0.5f);
    float th = kPi * (uv.y - 0.5f);
    return normalize(f3(cosf(th) * cosf(phi), sinf(th), cosf(th) * sinf(phi)));
}
OXR_D f2 sphere_to_plane(f3 v) {
    f2 uv{atan2f(v.z, v.x), asinf(v.y)};
    uv.x /= kTwoPi;
    uv.y /= kPi;
    uv.x += 0.5f;
    uv.y += 0.5f;
    return uv;
}
OXR_D f3 sample_cone(f2 s, float cos_angle) {
    float phi = kTwoPi * s.x;
    float ct = mixf(cos_angle, 1.0f, s.y);
    float st = sqrtf(1.0f - ct * ct);
    return normalize(f3(st * cosf(phi), st * sinf(phi), ct));
}
OXR_D float cone_pdf(float c) { return 1.0f / (kTwoPi * (1.0f - c)); }
OXR_D f3 cosine_hemisphere(f2 s) {
    float r = sqrtf(s.x);
    float phi = kTwoPi * s.y;
    float px = r * cosf(phi), py = r * sinf(phi);
    float pz = sqrtf(fmaxf(0.0f, 1.0f - px * px - py * py));
    return normalize(f3(px, py, pz));
}
OXR_D float cosine_pdf(float c) { return c * kInvPi; }
OXR_D f3 uniform_sphere(f2 s) {
    float phi = kTwoPi * s.x;
    float ct = 1.0f - 2.0f * s.y;
    float st = sqrtf(1.0f - ct * ct);
    return normalize(f3(st * cosf(phi), st * sinf(phi), ct));
}

// Filter.h:5-57
OXR_D f2 filter_jitter(f2 s, int type) {
    if (type == 0) return f2{s.x - 0.5f, s.y - 0.5f};
    if (type == 1) {
        float jx = s.x * 2.0f, jy = s.y * 2.0f;
        jx = jx < 1.0f ? sqrtf(jx) - 1.0f : 1.0f - sqrtf(2.0f - jx);
        jy = jy < 1.0f ? sqrtf(jy) - 1.0f : 1.0f - sqrtf(2.0f - jy);
        return f2{0.5f + jx, 0.5f + jy};
    }
    if (type == 2) {
        float u1 = s.x, u2 = s.y;
        if (u2 > u1) { u1 *= 0.5f; u2 -= u1; } else { u2 *= 0.5f; u1 -= u2; }
        return f2{0.5f + u1, 0.5f + u2};
    }
    if (type == 3) {
        float r1 = fmaxf(kFltMin, s.x);
        float r = sqrtf(-2.0f * logf(r1));
        float th = kTwoPi * s.y;
        return f2{0.5f + 0.375f * (r * cosf(th)), 0.5f + 0.375f * (r * sinf(th))};
    }
    return s;
}

// FP32 software bilinear, wrap addressing, normalised coordinates: the
// build's definition of tex2D with cudaFilterModeLinear/cudaAddressModeWrap
// (Scene.cpp:340-351, 380-391).
OXR_D int wrap_i(int i, int n) { int m = i % n; return m < 0 ? m + n : m; }
// texel(i): the float4 value of texel i (row-major).
template <class Texel>
OXR_D float4 bilinear_fetch(int w, int h, float u, float v, Texel texel) {
    // a NaN or infinite coordinate (from a NaN direction or garbage texture
    // coordinates) reads at 0
    if (!(fabsf(u) <= kFltMax)) u = 0.0f;
    if (!(fabsf(v) <= kFltMax)) v = 0.0f;
    u = u - floorf(u);
    v = v - floorf(v);
    float x = u * (float)w - 0.5f, y = v * (float)h - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    float ax = x - fx, ay = y - fy;
    int x0 = wrap_i((int)fx, w), x1 = wrap_i((int)fx + 1, w);
    int y0 = wrap_i((int)fy, h), y1 = wrap_i((int)fy + 1, h);
    float4 a = texel((size_t)y0 * w + x0), b = texel((size_t)y0 * w + x1);
    float4 c = texel((size_t)y1 * w + x0), d = texel((size_t)y1 * w + x1);
    float bx = 1.0f - ax, by = 1.0f - ay;
    float4 r;
    r.x = (a.x * bx + b.x * ax) * by + (c.x * bx + d.x * ax) * ay;
    r.y = (a.y * bx + b.y * ax) * by + (c.y * bx + d.y * ax) * ay;
    r.z = (a.z * bx + b.z * ax) * by + (c.z * bx + d.z * ax) * ay;
    r.w = (a.w * bx + b.w * ax) * by + (c.w * bx + d.w * ax) * ay;
    return r;
}
OXR_D float4 bilinear(const float4* __restrict__ img, int w, int h, float u, float v) {
    return bilinear_fetch(w, h, u, v, [&](size_t i) { return img[i]; });
}
// RGBA8 texel read as cudaReadModeNormalizedFloat does: byte / 255.
OXR_D float4 unorm8x4(uint32_t c) {
    return make_float4((float)(c & 255u) / 255.0f, (float)((c >> 8) & 255u) / 255.0f,
                       (float)((c >> 16) & 255u) / 255.0f, (float)(c >> 24) / 255.0f);
}
// Material texture: tex2D<float4> with linear filtering, wrap addressing and
// normalised-float reads of an RGBA8 image (Renderer.cpp:95-144).
OXR_D float4 bilinear_rgba8(const uint32_t* __restrict__ img, int w, int h, float u, float v) {
    return bilinear_fetch(w, h, u, v, [&](size_t i) { return unorm8x4(img[i]); });
}

// NormalFromTangentToWorld (Utils.h:85-94).
OXR_D f3 normal_from_tangent(f3 n, f3 t) {
    t = normalize(t * 2.0f - 1.0f);
    const f3 up = fabsf(n.z) < 0.9f ? f3(0.0f, 0.0f, 1.0f) : f3(1.0f, 0.0f, 0.0f);
    const f3 tx = normalize(cross(up, n));
    const f3 ty = normalize(cross(n, tx));
    return normalize(tx * t.x + ty * t.y + n * t.z);
}

}  // namespace oxr
