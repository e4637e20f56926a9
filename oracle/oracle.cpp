/*
 * CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h).  Not the product.
 *
 * Compiled with g++ -O2 -ffp-contract=off (oracle/Makefile) so that every
 * float operation rounds exactly as written.  Each function cites the
 * reference file:line it restates.  Conventions fixed by the build where the
 * reference leaves them open (kept identical in the GPU kernels):
 *   - Function arguments that draw from the RNG are drawn left to right
 *     (vec2f(random(), random()) -> x first).  C++ leaves the order open.
 *   - min/max on floats follow CUDA device semantics (fminf/fmaxf: a NaN
 *     operand yields the other operand), as the reference's device code did.
 *   - pow(x, int) is powf(x, (float)n).
 *   - tex2D (CUDA linear filter, wrap, normalised coordinates) is FP32
 *     software bilinear: x = frac(u)*W - 0.5, lerp in x then y.
 *   - optixTrace is an LBVH (Karras 2012) traversed closest-hit with a
 *     watertight ray/triangle test (Woop, Benthin, Wald 2013); ties in t go
 *     to the lower triangle id.  tmin = EPS = 1e-4 (devicePrograms.cu:261).
 *   - one path = (pixel, frame): RNG init(ix + W*iy, frameID) and one sample
 *     per frame, i.e. the reference with numPixelSamples = 1 in progressive
 *     mode (devicePrograms.cu:179-180, Renderer.cpp:642-646).
 *   - Russian roulette runs from bounce rr_start (the reference: 0).
 *   - AOV buffers are running means over frames like the colour buffer (the
 *     reference overwrites them each frame).
 *
 * Parity pins (tests/test_oracle_*.py):
 *   LCG                     known-answer vectors from the reference's own
 *                           random.h compiled in oracle/_ref, plus TEA/LCG
 *                           restated from the published algorithms.
 *   Morton / LBVH           brute-force Karras invariants, exhaustive small n.
 *   traversal               brute-force closest hit over all triangles.
 *   Fresnel, GGX, phase     closed forms (normal incidence, TIR, D and HG
 *                           normalisation, mean cosine), Monte-Carlo pdf
 *                           integrals, sample/eval consistency.
 *   env cache               exhaustive lower_bound check on small maps.
 *   tone map                closed form on a value grid.
 *   textures                texel centres, channel and row order, wrap,
 *                           a float64 bilinear; NormalFromTangentToWorld
 *                           frames and rotation invariants; zero-valued
 *                           textures equal zeroed parameters.
 *   direct lighting         whole renders of a Lambertian floor under one
 *                           sphere or quad light equal the closed-form
 *                           irradiance (NEE + MIS-weighted BSDF hits); two
 *                           lights show the reference's light_pdf quirk.
 *   Kulla-Conty table       the 128-point Hammersley estimate of the textbook
 *                           GGX albedo, converging to an independent float64
 *                           quadrature (test_oracle_pins.py).
 *   AverageFresnel fits     numerical 2 int F(mu) mu dmu of the textbook
 *                           (complex-index) Fresnel equations.
 *   FresnelConductor        its values at normal and grazing incidence as
 *                           written (first term + sin^2, not + cos^2).
 *   SampleGGX               chi-square against D(h) cos(theta_h), both branches.
 *   remaining lobes         sampled weights equal their closed forms (G1(L),
 *                           eta^2 G1(L), G1(V), albedo), Fresnel lobe choice,
 *                           white furnaces; sample/eval consistency.
 *   No compile of the reference's device headers is possible here (they
 *   need the OptiX SDK and CUDA headers).
 */
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr float kEps = 1e-4f;                         // Utils.h:10
constexpr float kPi = 3.14159265358979323846f;        // Utils.h:11
constexpr float kTwoPi = 6.28318530717958647692f;     // Utils.h:14
constexpr float kInvPi = 0.318309886183790671538f;    // Utils.h:13
constexpr int kLut = 64;                              // Microfacet.h:218

// ---------------------------------------------------------------- vectors
// Component-wise semantics of gdt vec_t (vec/functors.h:143-199).
struct V2 { float x, y; };
struct V3 {
    float x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(float s) : x(s), y(s), z(s) {}
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V3 operator/(V3 a, V3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
inline V3 operator+(V3 a, float s) { return {a.x + s, a.y + s, a.z + s}; }
inline V3 operator-(V3 a, float s) { return {a.x - s, a.y - s, a.z - s}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3 operator+(float s, V3 a) { return {s + a.x, s + a.y, s + a.z}; }
inline V3 operator-(float s, V3 a) { return {s - a.x, s - a.y, s - a.z}; }
inline V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator/(float s, V3 a) { return {s / a.x, s / a.y, s / a.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }   // vec.h:298-301
inline V3 cross(V3 a, V3 b) {                                                  // vec.h:289-294
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
inline V3 normalize(V3 v) { return v * 1.f / sqrtf(dot(v, v)); }              // vec.h:306-309
inline float length(V3 v) { return sqrtf(dot(v, v)); }                         // vec.h:313-316
inline V3 load3(const float* p) { return {p[0], p[1], p[2]}; }

// ------------------------------------------------------------- Utils.h
inline float sqr(float x) { return x * x; }                                    // Utils.h:16-18
inline float mixf(float x, float y, float t) { return x * (1.0f - t) + y * t; } // Utils.h:20-22
inline float clampf(float v, float lo, float hi) { return fminf(hi, fmaxf(lo, v)); } // functors.h:52
inline V3 reflect(V3 v, V3 n) { return normalize(v - 2.0f * dot(v, n) * n); }  // Utils.h:28-30
inline V3 refract(V3 v, V3 n, float eta) {                                     // Utils.h:32-38
    float cos_theta = fminf(dot(-v, n), 1.0f);
    V3 perp = eta * (v + cos_theta * n);
    V3 par = -sqrtf(fabsf(1.0f - dot(perp, perp))) * n;
    return normalize(perp + par);
}
inline bool is_nan(V3 v) { return std::isnan(v.x) || std::isnan(v.y) || std::isnan(v.z); }
inline V3 frame_c(V3 up) {                                                      // Utils.h:100-108
    if (fabsf(up.x) > fabsf(up.y)) {
        float li = 1.0f / sqrtf(up.x * up.x + up.z * up.z);
        return {up.z * li, 0.0f, -up.x * li};
    }
    float li = 1.0f / sqrtf(up.y * up.y + up.z * up.z);
    return {0.0f, up.z * li, -up.y * li};
}
inline V3 to_local(V3 dir, V3 up) {                                             // Utils.h:99-112
    V3 C = frame_c(up);
    V3 B = cross(C, up);
    return {dot(dir, B), dot(dir, C), dot(dir, up)};
}
inline V3 to_world(V3 dir, V3 up) {                                             // Utils.h:117-130
    V3 C = frame_c(up);
    V3 B = cross(C, up);
    return normalize(dir.x * B + dir.y * C + dir.z * up);
}

// ------------------------------------------------------------- RNG
// gdt::LCG<16> (random.h:55-89): TEA-16 seeding, then a 32-bit LCG.
struct Lcg {
    uint32_t state;
    void init(uint32_t v0, uint32_t v1) {
        uint32_t s0 = 0;
        for (int n = 0; n < 16; n++) {
            s0 += 0x9e3779b9u;
            v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
            v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
        }
        state = v0;
    }
    float operator()() {
        state = 1664525u * state + 1013904223u;
        return (float)(state & 0x00FFFFFFu) / (float)0x01000000;
    }
};

// ------------------------------------------------------------- Sampling.h
inline float power_heuristic(float p1, float p2) {                              // Sampling.h:5-10
    float a = powf(p1, 2.0f), b = powf(p2, 2.0f);
    return a / (a + b);
}
inline V3 plane_to_sphere(V2 uv) {                                              // Sampling.h:12-23
    float phi = kTwoPi * (uv.x - 0.5f);
    float theta = kPi * (uv.y - 0.5f);
    return normalize(V3(cosf(theta) * cosf(phi), sinf(theta), cosf(theta) * sinf(phi)));
}
inline V2 sphere_to_plane(V3 v) {                                               // Sampling.h:25-32
    V2 uv{atan2f(v.z, v.x), asinf(v.y)};
    uv.x /= kTwoPi;
    uv.y /= kPi;
    uv.x += 0.5f;
    uv.y += 0.5f;
    return uv;
}
inline V3 uniform_sample_cone(V2 s, float cos_angle) {                          // Sampling.h:35-47
    float phi = kTwoPi * s.x;
    float ct = mixf(cos_angle, 1.0f, s.y);
    float st = sqrtf(1.0f - ct * ct);
    return normalize(V3(st * cosf(phi), st * sinf(phi), ct));
}
inline float uniform_pdf_cone(float c) { return 1.0f / (kTwoPi * (1.0f - c)); } // Sampling.h:49-51
inline V3 cosine_sample_hemisphere(V2 s) {                                      // Sampling.h:55-68
    V3 p;
    float r = sqrtf(s.x);
    float phi = kTwoPi * s.y;
    p.x = r * cosf(phi);
    p.y = r * sinf(phi);
    p.z = sqrtf(fmaxf(0.0f, 1.0f - p.x * p.x - p.y * p.y));
    return normalize(p);
}
inline float cosine_pdf(float c) { return c * kInvPi; }                         // Sampling.h:70-72
inline V3 uniform_sample_sphere(V2 s) {                                         // Sampling.h:76-88
    float phi = kTwoPi * s.x;
    float ct = 1.0f - 2.0f * s.y;
    float st = sqrtf(1.0f - ct * ct);
    return normalize(V3(st * cosf(phi), st * sinf(phi), ct));
}

// ------------------------------------------------------------- Filter.h
inline V2 filter_jitter(V2 s, int type) {                                       // Filter.h:5-57
    if (type == 0) return {s.x - 0.5f, s.y - 0.5f};
    if (type == 1) {
        V2 j{s.x * 2.0f, s.y * 2.0f};
        j.x = j.x < 1.0f ? sqrtf(j.x) - 1.0f : 1.0f - sqrtf(2.0f - j.x);
        j.y = j.y < 1.0f ? sqrtf(j.y) - 1.0f : 1.0f - sqrtf(2.0f - j.y);
        return {0.5f + j.x, 0.5f + j.y};
    }
    if (type == 2) {
        float u1 = s.x, u2 = s.y;
        if (u2 > u1) { u1 *= 0.5f; u2 -= u1; } else { u2 *= 0.5f; u1 -= u2; }
        return {0.5f + u1, 0.5f + u2};
    }
    if (type == 3) {
        float r1 = fmaxf(FLT_MIN, s.x);
        float r = sqrtf(-2.0f * logf(r1));
        float th = kTwoPi * s.y;
        V2 uv{r * cosf(th), r * sinf(th)};
        return {0.5f + 0.375f * uv.x, 0.5f + 0.375f * uv.y};
    }
    return s;
}

// ------------------------------------------------------------- textures
// FP32 software bilinear with wrap on normalised coordinates (stands for
// tex2D with cudaFilterModeLinear + cudaAddressModeWrap, Scene.cpp:340-351).
inline int wrap_index(int i, int n) { int m = i % n; return m < 0 ? m + n : m; }
// texel(i, c): channel c of texel i (row-major)
template <typename Texel>
inline void bilinear_fetch(int w, int h, float u, float v, Texel texel, float* out) {
    if (!(fabsf(u) <= FLT_MAX)) u = 0.0f;  // NaN or infinite coordinates read at 0
    if (!(fabsf(v) <= FLT_MAX)) v = 0.0f;
    u = u - floorf(u);
    v = v - floorf(v);
    float x = u * (float)w - 0.5f;
    float y = v * (float)h - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    float ax = x - fx, ay = y - fy;
    int x0 = wrap_index((int)fx, w), x1 = wrap_index((int)fx + 1, w);
    int y0 = wrap_index((int)fy, h), y1 = wrap_index((int)fy + 1, h);
    const size_t i00 = (size_t)y0 * w + x0, i10 = (size_t)y0 * w + x1;
    const size_t i01 = (size_t)y1 * w + x0, i11 = (size_t)y1 * w + x1;
    for (int c = 0; c < 4; c++) {
        float top = texel(i00, c) * (1.0f - ax) + texel(i10, c) * ax;
        float bot = texel(i01, c) * (1.0f - ax) + texel(i11, c) * ax;
        out[c] = top * (1.0f - ay) + bot * ay;
    }
}
inline void bilinear(const float* img, int w, int h, float u, float v, float* out) {
    bilinear_fetch(w, h, u, v, [&](size_t i, int c) { return img[4 * i + c]; }, out);
}
// Material textures: RGBA8 with cudaReadModeNormalizedFloat (byte / 255),
// linear filter, wrap, normalised coordinates (Renderer.cpp:95-144).
inline void bilinear_rgba8(const uint8_t* img, int w, int h, float u, float v, float* out) {
    bilinear_fetch(w, h, u, v, [&](size_t i, int c) { return (float)img[4 * i + c] / 255.0f; }, out);
}

// ------------------------------------------------------------- Microfacet.h
inline V3 fresnel_schlick(V3 f0, float VdotH) {                                  // Microfacet.h:14-20
    float tmp = 1.0f - clampf(VdotH, 0.0f, 1.0f);
    float tmp2 = tmp * tmp;
    float Fc = tmp2 * tmp2 * tmp;
    return f0 + (1.0f - f0) * Fc;
}
inline V3 fresnel_conductor(V3 V, V3 H, V3 er, V3 ei) {                          // Microfacet.h:22-46
    V3 N = H;
    float c = dot(V, N), c2 = c * c, s2 = 1.0f - c2, s4 = s2 * s2;
    V3 t1 = er * er - ei * ei - s2;
    V3 a2pb2 = t1 * t1 + 4.0f * ei * ei * er * er;
    for (int i = 0; i < 3; i++) a2pb2[i] = sqrtf(fmaxf(0.0f, a2pb2[i]));
    V3 a = 0.5f * (a2pb2 + t1);
    for (int i = 0; i < 3; i++) a[i] = sqrtf(fmaxf(0.0f, a[i]));
    V3 term1 = a2pb2 + s2, term2 = 2.0f * c * a, term3 = a2pb2 * c2 + s4, term4 = term2 * s2;
    V3 rs = (term1 - term2) / (term1 + term2);
    V3 rp = rs * (term3 - term4) / (term3 + term4);
    return 0.5f * (rs + rp);
}
inline V3 average_fresnel_conductor(V3 eta, V3 k) {                             // Microfacet.h:48-70
    V3 r, e;
    for (int i = 0; i < 3; i++) {
        r[i] = (sqr(eta[i] - 1.0f) + sqr(k[i])) / (sqr(eta[i] + 1.0f) + sqr(k[i]));
        float t1 = 1.0f + sqrtf(r[i]);
        float t2 = 1.0f - sqrtf(r[i]);
        float t3 = (1.0f - r[i]) / (1.0f + r[i]);
        e[i] = (t1 - eta[i] * t2) / (t1 - t3 * t2);
    }
    return V3(0.087237f) + 0.0230685f * e - 0.0864902f * e * e + 0.0774594f * e * e * e +
           0.782654f * r - 0.136432f * r * r + 0.278708f * r * r * r + 0.19744f * e * r +
           0.0360605f * e * e * r - 0.2586f * e * r * r;
}
inline float fresnel_dielectric(V3 V, V3 H, float eta_inv) {                     // Microfacet.h:72-85
    float ci = fabsf(dot(V, H));
    float ct2 = 1.0f - sqr(eta_inv) * (1.0f - sqr(ci));
    if (ct2 <= 0.0f) return 1.0f;
    float ct = sqrtf(ct2);
    float rs = (eta_inv * ci - ct) / (eta_inv * ci + ct);
    float rp = (ci - eta_inv * ct) / (ci + eta_inv * ct);
    return (rs * rs + rp * rp) / 2.0f;
}
inline float average_fresnel_dielectric(float eta) {                             // Microfacet.h:87-113
    if (eta < 1.0f) return -1.4399f * (eta * eta) + 0.7099f * eta + 0.6681f + 0.0636f / eta;
    float i1 = 1.0f / eta, i2 = i1 * i1, i3 = i2 * i1, i4 = i3 * i1, i5 = i4 * i1;
    return 0.919317f - 3.4793f * i1 + 6.75335f * i2 - 7.80989f * i3 + 4.98554f * i4 - 1.36881f * i5;
}
inline float smith_g1(V3 V, V3 H, V3 N, float au, float av) {                    // Microfacet.h:117-142
    float c = dot(V, N);
    if (c * dot(V, H) <= 0.0f) return 0.0f;
    if (fabsf(c - 1.0f) < kEps) return 1.0f;
    if (au == av) {
        float c2 = sqr(c), t2 = (1.0f - c2) / c2, a2 = au * au;
        return 2.0f / (1.0f + sqrtf(1.0f + a2 * t2));
    }
    V3 d = to_local(V, N);
    float xy = sqr(au * d.x) + sqr(av * d.y), ta2 = xy / sqr(d.z);
    return 2.0f / (1.0f + sqrtf(1.0f + ta2));
}
inline float ggx_d(V3 H, V3 N, float au, float av) {                              // Microfacet.h:144-160
    float c = dot(H, N);
    if (c <= 0.0f) return 0.0f;
    float c2 = sqr(c), t2 = (1.0f - c2) / c2, a2 = au * av;
    if (au == av) return a2 / (kPi * powf(c, 2.0f) * sqr(a2 + t2));
    V3 d = to_local(H, N);
    return 1.0f / (kPi * a2 * sqr(sqr(d.x / au) + sqr(d.y / av) + sqr(d.z)));
}
inline float ggx_dv(V3 V, V3 H, V3 N, float au, float av) {                       // Microfacet.h:162-164
    return smith_g1(V, H, N, au, av) * dot(V, H) * ggx_d(H, N, au, av) / dot(N, V);
}
inline V3 sample_ggx(float au, float av, V2 s) {                                 // Microfacet.h:166-186
    float sp, cp, a2;
    if (au == av) {
        float phi = kTwoPi * s.y;
        cp = cosf(phi);
        sp = sinf(phi);
        a2 = au * au;
    } else {
        float phi = atanf(av / au * tanf(kPi + kTwoPi * s.y)) + kPi * floorf(2.0f * s.y + 0.5f);
        cp = cosf(phi);
        sp = sinf(phi);
        a2 = 1.0f / (sqr(cp / au) + sqr(sp / av));
    }
    float t2 = a2 * s.x / (1.0f - s.x), ct = 1.0f / sqrtf(1.0f + t2), st = sqrtf(1.0f - ct * ct);
    return normalize(V3(st * cp, st * sp, ct));
}
inline V3 sample_vndf(V3 N, V3 Ve, float au, float av, V2 s) {                   // Microfacet.h:188-214
    V3 V = to_local(Ve, N);
    V3 Vh = normalize(V3(au * V.x, av * V.y, V.z));
    float len2 = sqr(Vh.x) + sqr(Vh.y);
    V3 T1 = len2 > 0.0f ? V3(-Vh.y, Vh.x, 0.0f) * sqrtf(1.0f / len2) : V3(1.0f, 0.0f, 0.0f);
    V3 T2 = cross(Vh, T1);
    float r = sqrtf(s.x), phi = s.y * kTwoPi;
    float t1 = r * cosf(phi), t2 = r * sinf(phi);
    float sw = 0.5f * (1.0f + Vh.z);
    t2 = (1.0f - sw) * sqrtf(1.0f - sqr(t1)) + sw * t2;
    V3 Nh = t1 * T1 + t2 * T2 + sqrtf(fmaxf(0.0f, 1.0f - sqr(t1) - sqr(t2))) * Vh;
    return normalize(V3(au * Nh.x, av * Nh.y, fmaxf(0.0f, Nh.z)));
}
inline V2 hammersley(uint32_t i, uint32_t n) {                                   // Microfacet.h:220-229
    uint32_t b = (i << 16u) | (i >> 16u);
    b = ((b & 0x55555555u) << 1u) | ((b & 0xAAAAAAAAu) >> 1u);
    b = ((b & 0x33333333u) << 2u) | ((b & 0xCCCCCCCCu) >> 2u);
    b = ((b & 0x0F0F0F0Fu) << 4u) | ((b & 0xF0F0F0F0u) >> 4u);
    b = ((b & 0x00FF00FFu) << 8u) | ((b & 0xFF00FF00u) >> 8u);
    return {(float)i / (float)n, (float)b * 2.3283064365386963e-10f};
}
void kulla_conty(float* bsdf, float* albedo) {                                    // Microfacet.h:231-291
    const V3 N(0.0f, 0.0f, 1.0f);
    auto integrate_bsdf = [&](V3 V, float r) {
        const float step = 1.0f / 128;
        float acc = 0.0f;
        for (uint32_t i = 0; i < 128; ++i) {
            V3 H = sample_ggx(r, r, hammersley(i + 1, 129));
            V3 L = reflect(-V, H);
            float G = smith_g1(L, H, N, r, r) * smith_g1(V, H, N, r, r);
            float NdotV = dot(N, V), NdotL = dot(N, L), NdotH = dot(N, H), HdotV = dot(H, V);
            if (NdotL > 0.0f && NdotH > 0.0f && HdotV > 0.0f) acc += (HdotV * G) / (NdotV * NdotH);
        }
        return std::min(acc * step, 1.0f);
    };
    auto integrate_albedo = [&](V3 V, float r, float b) {
        const float step = 1.0f / 128;
        float acc = 0.0f;
        for (uint32_t i = 0; i < 128; ++i) {
            V3 H = sample_ggx(r, r, hammersley(i + 1, 129));
            V3 L = reflect(-V, H);
            float HdotV = dot(H, V), NdotL = dot(N, L), NdotH = dot(N, H);
            if (NdotL > 0.0f && NdotH > 0.0f && HdotV > 0.0f) acc += b * NdotL;
        }
        return acc * 2.0f * step;
    };
    float step = 1.0f / kLut;
    for (int i = kLut - 1; i >= 0; --i) {
        float acc = 0.0f;
        float r = step * ((float)i + 0.5f);
        for (int j = kLut - 1; j >= 0; --j) {
            float NdotV = step * ((float)j + 0.5f);
            V3 V(sqrtf(1.0f - NdotV * NdotV), 0.0f, NdotV);
            float b = integrate_bsdf(V, r);
            bsdf[i * kLut + j] = b;
            acc += integrate_albedo(V, r, b);
        }
        albedo[i] = acc * step;
    }
}
inline int lut_index(float x) { return (x == x) ? (int)x : 0; }
inline float bsdf_avg(float c, float r, const float* t) {                        // Microfacet.h:293-324
    float o1 = sqr(r) * kLut, o2 = c * kLut;
    int i1 = lut_index(o1), i2 = lut_index(o2);
    if (i1 >= kLut - 1) {
        if (i2 >= kLut - 1) return t[(kLut - 1) * kLut + kLut - 1];
        return mixf(t[(kLut - 1) * kLut + i2], t[(kLut - 1) * kLut + i2 + 1], o2 - i2);
    }
    if (i2 >= kLut - 1)
        return mixf(t[i1 * kLut + kLut - 1], t[(i1 + 1) * kLut + kLut - 1], o1 - i1);
    return mixf(mixf(t[i1 * kLut + i2], t[(i1 + 1) * kLut + i2], o1 - i1),
                mixf(t[i1 * kLut + i2 + 1], t[(i1 + 1) * kLut + i2 + 1], o1 - i1), o2 - i2);
}
inline float albedo_avg(float r, const float* t) {                               // Microfacet.h:326-335
    float o = sqr(r) * kLut;
    int i = lut_index(o);
    if (i >= kLut - 1) return t[kLut - 1];
    return mixf(t[i], t[i + 1], o - i);
}

struct Lut { float bsdf[kLut * kLut]; float albedo[kLut]; };
const Lut& lut() {
    static Lut l = [] { Lut t; kulla_conty(t.bsdf, t.albedo); return t; }();
    return l;
}
inline V3 multiple_scatter(float NdotL, float NdotV, float r, V3 F_avg) {         // Microfacet.h:337-346
    const Lut& L = lut();
    float bL = bsdf_avg(NdotL, r, L.bsdf), bV = bsdf_avg(NdotV, r, L.bsdf), aa = albedo_avg(r, L.albedo);
    float f_ms = (1.0f - bL) * (1.0f - bV) / (kPi * (1.0f - aa));
    V3 f_add = F_avg * F_avg * aa / (1.0f - F_avg * (1.0f - aa));
    return f_ms * f_add;
}

// ------------------------------------------------------------- Material.h
struct Mat {
    int type;
    V3 albedo;
    float roughness, anisotropy;
    V3 eta, k;
    float int_ior, ext_ior;
    bool nonlinear;
    V3 specular;
    float metallic, coat_weight, coat_ru, coat_rv;
    int in_medium, out_medium;
    int tex[8];  // albedo roughness anisotropy specular metallic coat_u coat_v normal
};
Mat to_mat(const orc_material& m) {
    Mat r;
    r.type = m.type;
    r.albedo = load3(m.albedo);
    r.roughness = m.roughness;
    r.anisotropy = m.anisotropy;
    r.eta = load3(m.eta);
    r.k = load3(m.k);
    r.int_ior = m.int_ior;
    r.ext_ior = m.ext_ior;
    r.nonlinear = m.nonlinear != 0;
    r.specular = load3(m.specular);
    r.metallic = m.metallic;
    r.coat_weight = m.coat_weight;
    r.coat_ru = m.coat_roughness_u;
    r.coat_rv = m.coat_roughness_v;
    r.in_medium = m.in_medium;
    r.out_medium = m.out_medium;
    for (int i = 0; i < 8; i++) r.tex[i] = m.tex[i];
    return r;
}
struct Isect {                                                                   // Utils.h:62-74
    float distance;
    V3 position;
    V3 shadeNormal;
    bool frontFace;
    const Mat* mat;
    int in_medium, out_medium;                                                   // Utils.h:45-60
    int get_medium(bool ff) const { return ff ? out_medium : in_medium; }
    void set_face_normal(V3 dir, V3 n) {
        frontFace = dot(dir, n) < 0.0f;
        shadeNormal = frontFace ? n : -n;
    }
};

V3 eval_diffuse(const Isect& is, V3 V, V3 L, float& pdf) {                        // Material.h:7-38
    float r = is.mat->roughness;
    V3 albedo = is.mat->albedo;
    V3 lL = to_local(L, is.shadeNormal);
    V3 N = is.shadeNormal;
    V3 H = normalize(V + L);
    float NdotL = lL.z, NdotV = dot(N, V), VdotH = dot(V, H);
    if (NdotL <= 0.0f || NdotV <= 0.0f) return V3(0.0f);
    float s = r * r, s2 = s * s;
    float VdotL = 2.0f * VdotH * VdotH - 1.0f;
    float cosri = VdotL - NdotV * NdotL;
    float C1 = 1.0f - 0.5f * s2 / (s2 + 0.33f);
    float C2 = 0.45f * s2 / (s2 + 0.09f) * cosri * (cosri >= 0.0f ? fmaxf(NdotL, NdotV) : 1.0f);
    V3 brdf = albedo * kInvPi * (C1 + C2) * (1.0f + r * 0.5f);
    pdf = cosine_pdf(NdotL);
    return brdf;
}
V3 sample_diffuse(const Isect& is, Lcg& rng, V3 V, V3& Lw, float& pdf) {         // Material.h:40-72
    float r = is.mat->roughness;
    V3 albedo = is.mat->albedo;
    float u0 = rng(), u1 = rng();
    V3 lL = cosine_sample_hemisphere({u0, u1});
    Lw = to_world(lL, is.shadeNormal);
    V3 L = Lw, N = is.shadeNormal;
    V3 H = normalize(V + L);
    float NdotL = lL.z, NdotV = dot(N, V), VdotH = dot(V, H);
    if (NdotL <= 0.0f || NdotV <= 0.0f) return V3(0.0f);
    float s = r * r, s2 = s * s;
    float VdotL = 2.0f * VdotH * VdotH - 1.0f;
    float cosri = VdotL - NdotV * NdotL;
    float C1 = 1.0f - 0.5f * s2 / (s2 + 0.33f);
    float C2 = 0.45f * s2 / (s2 + 0.09f) * cosri * (cosri >= 0.0f ? fmaxf(NdotL, NdotV) : 1.0f);
    V3 brdf = albedo * kInvPi * (C1 + C2) * (1.0f + r * 0.5f);
    pdf = cosine_pdf(NdotL);
    return brdf;
}
V3 eval_conductor(const Isect& is, V3 V, V3 L, float& pdf) {                      // Material.h:76-109
    const Mat& m = *is.mat;
    float au = sqr(m.roughness) * (1.0f + m.anisotropy), av = sqr(m.roughness) * (1.0f - m.anisotropy);
    V3 F_avg = average_fresnel_conductor(m.eta, m.k);
    V3 N = is.shadeNormal;
    V3 H = normalize(V + L);
    float Dv = ggx_dv(V, H, N, au, av);
    pdf = Dv * fabsf(1.0f / (4.0f * dot(V, H)));
    float NdotV = dot(N, V), NdotL = dot(N, L);
    if (NdotV <= 0.0f || NdotL <= 0.0f) return V3(0.0f);
    V3 F = fresnel_conductor(V, H, m.eta, m.k);
    float G = smith_g1(V, H, N, au, av) * smith_g1(L, H, N, au, av);
    float D = ggx_d(H, N, au, av);
    return m.albedo * (F * D * G / (4.0f * NdotV * NdotL) + multiple_scatter(NdotL, NdotV, m.roughness, F_avg));
}
V3 sample_conductor(const Isect& is, Lcg& rng, V3 V, V3& Lw, float& pdf) {       // Material.h:111-146
    const Mat& m = *is.mat;
    float au = sqr(m.roughness) * (1.0f + m.anisotropy), av = sqr(m.roughness) * (1.0f - m.anisotropy);
    V3 F_avg = average_fresnel_conductor(m.eta, m.k);
    V3 N = is.shadeNormal;
    float u0 = rng(), u1 = rng();
    V3 H = sample_vndf(N, V, au, av, {u0, u1});
    H = to_world(H, N);
    Lw = reflect(-V, H);
    V3 L = Lw;
    float Dv = ggx_dv(V, H, N, au, av);
    pdf = Dv * fabsf(1.0f / (4.0f * dot(V, H)));
    float NdotV = dot(N, V), NdotL = dot(N, L);
    if (NdotV <= 0.0f || NdotL <= 0.0f) return V3(0.0f);
    V3 F = fresnel_conductor(V, H, m.eta, m.k);
    float G = smith_g1(V, H, N, au, av) * smith_g1(L, H, N, au, av);
    float D = ggx_d(H, N, au, av);
    return m.albedo * (F * D * G / (4.0f * NdotV * NdotL) + multiple_scatter(NdotL, NdotV, m.roughness, F_avg));
}
V3 eval_dielectric(const Isect& is, V3 V, V3 L, float& pdf) {                     // Material.h:150-206
    const Mat& m = *is.mat;
    float au = sqr(m.roughness) * (1.0f + m.anisotropy), av = sqr(m.roughness) * (1.0f - m.anisotropy);
    float eta = m.int_ior / m.ext_ior;
    float eio = is.frontFace ? (1.0f / eta) : eta;
    V3 N = is.shadeNormal, H;
    bool refl = dot(N, L) * dot(N, V) > 0.0f;
    if (refl) {
        H = normalize(V + L);
    } else {
        H = -normalize(eio * V + L);
        if (dot(N, H) < 0.0f) H = -H;
    }
    float Dv = ggx_dv(V, H, N, au, av);
    pdf = Dv * fabsf(1.0f / (4.0f * dot(V, H)));
    float NdotV = fabsf(dot(N, V)), NdotL = fabsf(dot(N, L));
    float F = fresnel_dielectric(V, H, eio);
    float G = smith_g1(V, H, N, au, av) * smith_g1(L, H, N, au, av);
    float D = ggx_d(H, N, au, av);
    V3 f(0.0f);
    if (refl) {
        float dwh = fabsf(1.0f / (4.0f * dot(V, H)));
        pdf = F * Dv * dwh;
        f = m.albedo * F * D * G / (4.0f * NdotV * NdotL);
    } else {
        float HdotV = dot(H, V), HdotL = dot(H, L);
        float sd = eio * HdotV + HdotL;
        float factor = fabsf(HdotL * HdotV / (NdotL * NdotV));
        float dwh = fabsf(HdotL) / sqr(sd);
        pdf = (1.0f - F) * Dv * dwh;
        f = m.albedo * (1.0f - F) * D * G * factor / sqr(sd);
        f = f * sqr(1.0f / eio);
    }
    return f;
}
V3 sample_dielectric(const Isect& is, Lcg& rng, V3 V, V3& Lw, float& pdf) {      // Material.h:208-269
    const Mat& m = *is.mat;
    float au = sqr(m.roughness) * (1.0f + m.anisotropy), av = sqr(m.roughness) * (1.0f - m.anisotropy);
    float eta = m.int_ior / m.ext_ior;
    float eio = is.frontFace ? (1.0f / eta) : eta;
    V3 N = is.shadeNormal;
    float u0 = rng(), u1 = rng();
    V3 H = sample_vndf(N, V, au, av, {u0, u1});
    H = to_world(H, N);
    float Dv = ggx_dv(V, H, N, au, av);
    float F = fresnel_dielectric(V, H, eio);
    float D = ggx_d(H, N, au, av);
    if (rng() < F) {
        Lw = reflect(-V, H);
        V3 L = Lw;
        if (dot(N, L) <= 0.0f) return V3(0.0f);
        float dwh = fabsf(1.0f / (4.0f * dot(V, H)));
        pdf = F * Dv * dwh;
        float NdotV = fabsf(dot(N, V)), NdotL = fabsf(dot(N, L));
        float G = smith_g1(V, H, N, au, av) * smith_g1(L, H, N, au, av);
        return m.albedo * F * D * G / (4.0f * NdotV * NdotL);
    }
    Lw = refract(-V, H, eio);
    V3 L = Lw;
    if (dot(N, L) * dot(N, V) >= 0.0f) return V3(0.0f);
    float NdotV = fabsf(dot(N, V)), NdotL = fabsf(dot(N, L));
    float G = smith_g1(V, H, N, au, av) * smith_g1(L, H, N, au, av);
    float HdotV = dot(H, V), HdotL = dot(H, L);
    float sd = eio * HdotV + HdotL;
    float factor = fabsf(HdotL * HdotV / (NdotL * NdotV));
    float dwh = fabsf(HdotL) / sqr(sd);
    pdf = (1.0f - F) * Dv * dwh;
    V3 f = m.albedo * (1.0f - F) * D * G * factor / sqr(sd);
    return f * sqr(1.0f / eio);
}
V3 eval_plastic(const Isect& is, V3 V, V3 L, float& pdf) {                        // Material.h:273-325
    const Mat& m = *is.mat;
    V3 kd = m.albedo, ks = m.specular;
    float d_sum = kd.x + kd.y + kd.z, s_sum = ks.x + ks.y + ks.z;
    float au = sqr(m.roughness) * (1.0f + m.anisotropy), av = sqr(m.roughness) * (1.0f - m.anisotropy);
    float eta = m.int_ior / m.ext_ior;
    float F_avg = average_fresnel_dielectric(eta);
    V3 N = is.shadeNormal;
    V3 H = normalize(V + L);
    float NdotV = dot(N, V), NdotL = dot(N, L);
    if (NdotV <= 0.0f || NdotL <= 0.0f) return V3(0.0f);
    float Dv = ggx_dv(V, H, N, au, av);
    float Fo = fresnel_dielectric(V, N, 1.0f / eta), Fi = fresnel_dielectric(L, N, 1.0f / eta);
    float ssw = s_sum / (s_sum + d_sum);
    float ps = Fi * ssw, pd = (1.0f - Fi) * (1.0f - ssw);
    ps = ps / (ps + pd);
    V3 F(fresnel_dielectric(L, H, 1.0f / eta));
    float D = ggx_d(H, N, au, av);
    float G = smith_g1(V, H, N, au, av) * smith_g1(L, H, N, au, av);
    V3 brdf = m.nonlinear ? kd / (1.0f - kd * F_avg) : kd / (1.0f - F_avg);
    brdf = brdf * ((1.0f - Fi) * (1.0f - Fo) * kInvPi);
    brdf = brdf + ks * (F * D * G / (4.0f * NdotL * NdotV) + multiple_scatter(NdotL, NdotV, m.roughness, V3(F_avg)));
    pdf = ps * Dv * fabsf(1.0f / (4.0f * dot(V, H))) + (1.0f - ps) * cosine_pdf(NdotL);
    return brdf;
}
V3 sample_plastic(const Isect& is, Lcg& rng, V3 V, V3& Lw, float& pdf) {         // Material.h:327-403
    const Mat& m = *is.mat;
    V3 kd = m.albedo, ks = m.specular;
    float d_sum = kd.x + kd.y + kd.z, s_sum = ks.x + ks.y + ks.z;
    float au = sqr(m.roughness) * (1.0f + m.anisotropy), av = sqr(m.roughness) * (1.0f - m.anisotropy);
    float eta = m.int_ior / m.ext_ior;
    float F_avg = average_fresnel_dielectric(eta);
    V3 N = is.shadeNormal;
    float NdotV = dot(N, V);
    if (NdotV <= 0.0f) return V3(0.0f);
    float Fo = fresnel_dielectric(V, N, 1.0f / eta), Fi = Fo;
    float ssw = s_sum / (s_sum + d_sum);
    float ps = Fi * ssw, pd = (1.0f - Fi) * (1.0f - ssw);
    ps = ps / (ps + pd);
    V3 L, H;
    float NdotL;
    if (rng() < ps) {
        float u0 = rng(), u1 = rng();
        H = sample_vndf(N, V, au, av, {u0, u1});
        H = to_world(H, N);
        Lw = reflect(-V, H);
        L = Lw;
        NdotL = dot(N, L);
        if (NdotL <= 0.0f) return V3(0.0f);
    } else {
        float u0 = rng(), u1 = rng();
        V3 lL = cosine_sample_hemisphere({u0, u1});
        Lw = to_world(lL, N);
        L = Lw;
        H = normalize(V + L);
        Fi = fresnel_dielectric(L, N, 1.0f / eta);
        NdotL = dot(N, L);
        if (NdotL <= 0.0f) return V3(0.0f);
    }
    float Dv = ggx_dv(V, H, N, au, av);
    float G = smith_g1(V, H, N, au, av) * smith_g1(L, H, N, au, av);
    float D = ggx_d(H, N, au, av);
    V3 F(fresnel_dielectric(L, H, 1.0f / eta));
    V3 brdf = m.nonlinear ? kd / (1.0f - kd * F_avg) : kd / (1.0f - F_avg);
    brdf = brdf * ((1.0f - Fi) * (1.0f - Fo) * kInvPi);
    brdf = brdf + ks * (F * D * G / (4.0f * NdotL * NdotV) + multiple_scatter(NdotL, NdotV, m.roughness, V3(F_avg)));
    pdf = ps * Dv * fabsf(1.0f / (4.0f * dot(V, H))) + (1.0f - ps) * cosine_pdf(NdotL);
    return brdf;
}
V3 eval_metal_workflow(const Isect& is, V3 V, V3 L, float& pdf) {                 // Material.h:407-449
    const Mat& m = *is.mat;
    float au = sqr(m.roughness) * (1.0f + m.anisotropy), av = sqr(m.roughness) * (1.0f - m.anisotropy);
    V3 N = is.shadeNormal;
    V3 H = normalize(L + V);
    float NdotV = dot(N, V), NdotL = dot(N, L);
    if (NdotV <= 0.0f || NdotL <= 0.0f) return V3(0.0f);
    float diel = 1.0f - m.metallic;
    float diffuse = diel, specular = m.metallic + diel;
    float p_diffuse = diffuse / (diffuse + specular);
    float Dv = ggx_dv(V, H, N, au, av);
    float G = smith_g1(V, H, N, au, av) * smith_g1(L, H, N, au, av);
    float D = ggx_d(H, N, au, av);
    V3 F0 = V3(0.04f) * (1.0f - m.metallic) + m.albedo * m.metallic;
    V3 F = fresnel_schlick(F0, dot(V, H));
    V3 spec = D * F * G / (4.0f * NdotL * NdotV);
    V3 diff = m.albedo * kInvPi;
    V3 brdf = p_diffuse * diff + (1.0f - p_diffuse) * spec;
    pdf = (1.0f - p_diffuse) * Dv * fabsf(1.0f / (4.0f * dot(V, H))) + p_diffuse * cosine_pdf(NdotL);
    return brdf;
}
V3 sample_metal_workflow(const Isect& is, Lcg& rng, V3 V, V3& Lw, float& pdf) {  // Material.h:451-516
    const Mat& m = *is.mat;
    float au = sqr(m.roughness) * (1.0f + m.anisotropy), av = sqr(m.roughness) * (1.0f - m.anisotropy);
    V3 N = is.shadeNormal;
    float NdotV = dot(N, V);
    if (NdotV <= 0.0f) return V3(0.0f);
    float diel = 1.0f - m.metallic;
    float diffuse = diel, specular = m.metallic + diel;
    float p_diffuse = diffuse / (diffuse + specular);
    V3 L, H;
    float NdotL;
    if (rng() < p_diffuse) {
        float u0 = rng(), u1 = rng();
        V3 lL = cosine_sample_hemisphere({u0, u1});
        Lw = to_world(lL, N);
        L = Lw;
        H = normalize(V + L);
        NdotL = dot(N, L);
        if (NdotL <= 0.0f) return V3(0.0f);
    } else {
        float u0 = rng(), u1 = rng();
        H = sample_vndf(N, V, au, av, {u0, u1});
        H = to_world(H, N);
        Lw = reflect(-V, H);
        L = Lw;
        NdotL = dot(N, L);
        if (NdotL <= 0.0f) return V3(0.0f);
    }
    float Dv = ggx_dv(V, H, N, au, av);
    float G = smith_g1(V, H, N, au, av) * smith_g1(L, H, N, au, av);
    float D = ggx_d(H, N, au, av);
    V3 F0 = V3(0.04f) * (1.0f - m.metallic) + m.albedo * m.metallic;
    V3 F = fresnel_schlick(F0, dot(V, H));
    V3 spec = D * F * G / (4.0f * NdotL * NdotV);
    V3 diff = m.albedo * kInvPi;
    V3 brdf = p_diffuse * diff + (1.0f - p_diffuse) * spec;
    pdf = (1.0f - p_diffuse) * Dv * fabsf(1.0f / (4.0f * dot(V, H))) + p_diffuse * cosine_pdf(NdotL);
    return brdf;
}
V3 eval_thin_dielectric(const Isect& is, V3 V, V3 L, float& pdf) {                // Material.h:520-575
    const Mat& m = *is.mat;
    float eta = m.int_ior / m.ext_ior;
    float a = sqr(m.roughness);
    V3 N = is.shadeNormal, H;
    bool refl = dot(N, L) * dot(N, V) > 0.0f;
    if (refl) {
        H = normalize(V + L);
    } else {
        V3 Vr = to_world(to_local(V, -N), N);
        H = normalize(Vr + L);
        if (dot(N, H) < 0.0f) H = -H;
    }
    float Dv = ggx_dv(V, H, N, a, a);
    float F = fresnel_dielectric(V, H, 1.0f / eta);
    if (F < 1.0f) F *= 2.0f / (1.0f + F);
    float D = ggx_d(H, N, a, a);
    float dwh = fabsf(1.0f / (4.0f * dot(V, H)));
    float NdotV = fabsf(dot(N, V)), NdotL = fabsf(dot(N, L));
    float G = smith_g1(V, H, N, a, a) * smith_g1(L, H, N, a, a);
    if (refl) {
        if (dot(N, L) <= 0.0f) return V3(0.0f);
        pdf = F * Dv * dwh;
        return m.albedo * F * D * G / (4.0f * NdotV * NdotL);
    }
    if (dot(N, L) * dot(N, V) >= 0.0f) return V3(0.0f);
    pdf = (1.0f - F) * Dv * dwh;
    return m.albedo * (1.0f - F) * D * G / (4.0f * NdotV * NdotL);
}
V3 sample_thin_dielectric(const Isect& is, Lcg& rng, V3 V, V3& Lw, float& pdf) { // Material.h:577-630
    const Mat& m = *is.mat;
    float eta = m.int_ior / m.ext_ior;
    float a = sqr(m.roughness);
    V3 N = is.shadeNormal;
    float u0 = rng(), u1 = rng();
    V3 H = sample_vndf(N, V, a, a, {u0, u1});
    H = to_world(H, N);
    float Dv = ggx_dv(V, H, N, a, a);
    float F = fresnel_dielectric(V, H, 1.0f / eta);
    if (F < 1.0f) F *= 2.0f / (1.0f + F);
    float D = ggx_d(H, N, a, a);
    if (rng() < F) {
        Lw = reflect(-V, H);
        V3 L = Lw;
        if (dot(N, L) <= 0.0f) return V3(0.0f);
        float dwh = fabsf(1.0f / (4.0f * dot(V, H)));
        float NdotV = fabsf(dot(N, V)), NdotL = fabsf(dot(N, L));
        float G = smith_g1(V, H, N, a, a) * smith_g1(L, H, N, a, a);
        pdf = F * Dv * dwh;
        return m.albedo * F * D * G / (4.0f * NdotV * NdotL);
    }
    Lw = -V;
    V3 L = Lw;
    if (dot(N, L) * dot(N, V) >= 0.0f) return V3(0.0f);
    float NdotV = fabsf(dot(N, V)), NdotL = fabsf(dot(N, L));
    float G = smith_g1(V, H, N, a, a) * smith_g1(L, H, N, a, a);
    float dwh = fabsf(1.0f / (4.0f * dot(V, H)));
    pdf = (1.0f - F) * Dv * dwh;
    return m.albedo * (1.0f - F) * D * G / (4.0f * NdotV * NdotL);
}
V3 eval_clearcoat(const Isect& is, V3 V, V3 L, float& pdf) {                      // Material.h:634-667
    const Mat& m = *is.mat;
    float au = sqr(m.coat_ru), av = sqr(m.coat_rv);
    V3 N = is.shadeNormal;
    V3 H = normalize(V + L);
    float Dv = ggx_dv(V, H, N, au, av);
    float NdotV = dot(N, V), NdotL = dot(N, L);
    if (NdotV <= 0.0f || NdotL <= 0.0f) return V3(0.0f);
    float F = fresnel_dielectric(V, H, 1.0f / 1.5f);
    float cw = m.coat_weight * F;
    float G = smith_g1(V, H, N, au, av) * smith_g1(L, H, N, au, av);
    float D = ggx_d(H, N, au, av);
    float cond_pdf = 0.0f;
    float coat_pdf = Dv * fabsf(1.0f / (4.0f * dot(V, H)));
    V3 cond = eval_conductor(is, V, L, cond_pdf);
    V3 coat(D * G / (4.0f * NdotV * NdotL));
    V3 brdf = cw * coat + (1.0f - cw) * cond;
    pdf = cw * coat_pdf + (1.0f - cw) * cond_pdf;
    return brdf;
}
V3 sample_clearcoat(const Isect& is, Lcg& rng, V3 V, V3& Lw, float& pdf) {       // Material.h:669-734
    const Mat& m = *is.mat;
    float au = sqr(m.coat_ru), av = sqr(m.coat_rv);
    V3 N = is.shadeNormal, H;
    float F = fresnel_dielectric(V, N, 1.0f / 1.5f);
    float cw = m.coat_weight * F;
    if (rng() < cw) {
        float u0 = rng(), u1 = rng();
        H = sample_vndf(N, V, au, av, {u0, u1});
        H = to_world(H, N);
        Lw = reflect(-V, H);
        V3 L = Lw;
        float NdotV = dot(N, V), NdotL = dot(N, L);
        if (NdotV <= 0.0f || NdotL <= 0.0f) return V3(0.0f);
        float Dv = ggx_dv(V, H, N, au, av);
        F = fresnel_dielectric(V, H, 1.0f / 1.5f);
        cw = m.coat_weight * F;
        float G = smith_g1(V, H, N, au, av) * smith_g1(L, H, N, au, av);
        float D = ggx_d(H, N, au, av);
        float coat_pdf = Dv * fabsf(1.0f / (4.0f * dot(V, H)));
        float cond_pdf = 0.0f;
        V3 cond = eval_conductor(is, V, Lw, cond_pdf);
        V3 coat(D * G / (4.0f * NdotV * NdotL));
        pdf = cw * coat_pdf + (1.0f - cw) * cond_pdf;
        return cw * coat + (1.0f - cw) * cond;
    }
    float cond_pdf = 0.0f;
    V3 cond = sample_conductor(is, rng, V, Lw, cond_pdf);
    V3 L = Lw;
    H = normalize(V + L);
    float NdotV = dot(N, V), NdotL = dot(N, L);
    if (NdotV <= 0.0f || NdotL <= 0.0f) return V3(0.0f);
    float Dv = ggx_dv(V, H, N, au, av);
    F = fresnel_dielectric(V, H, 1.0f / 1.5f);
    cw = m.coat_weight * F;
    float G = smith_g1(V, H, N, au, av) * smith_g1(L, H, N, au, av);
    float D = ggx_d(H, N, au, av);
    float coat_pdf = Dv * fabsf(1.0f / (4.0f * dot(V, H)));
    V3 coat(D * G / (4.0f * NdotV * NdotL));
    pdf = cw * coat_pdf + (1.0f - cw) * cond_pdf;
    return cw * coat + (1.0f - cw) * cond;
}
V3 eval_diffuse_transmitter(const Isect& is, V3 V, V3 L, float& pdf) {            // Material.h:738-757
    V3 N = -is.shadeNormal;
    V3 lL = to_local(L, N);
    float NdotL = lL.z, NdotV = dot(N, V);
    if (NdotL * NdotV >= 0.0f) return V3(0.0f);
    pdf = cosine_pdf(fabsf(NdotL));
    return is.mat->albedo * kInvPi;
}
V3 sample_diffuse_transmitter(const Isect& is, Lcg& rng, V3 V, V3& Lw, float& pdf) { // Material.h:759-779
    float u0 = rng(), u1 = rng();
    V3 lL = cosine_sample_hemisphere({u0, u1});
    V3 N = -is.shadeNormal;
    Lw = to_world(lL, N);
    float NdotL = lL.z, NdotV = dot(N, V);
    if (NdotL * NdotV >= 0.0f) return V3(0.0f);
    pdf = cosine_pdf(fabsf(NdotL));
    return is.mat->albedo * kInvPi;
}
V3 eval_material(const Isect& is, V3 V, V3 L, float& pdf) {                       // Material.h:783-813
    pdf = 0.0f;
    switch (is.mat->type) {
        case 0: return eval_diffuse(is, V, L, pdf);
        case 1: return eval_conductor(is, V, L, pdf);
        case 2: return eval_dielectric(is, V, L, pdf);
        case 3: return eval_plastic(is, V, L, pdf);
        case 4: return eval_metal_workflow(is, V, L, pdf);
        case 5: return eval_thin_dielectric(is, V, L, pdf);
        case 6: return eval_clearcoat(is, V, L, pdf);
        case 7: return eval_diffuse_transmitter(is, V, L, pdf);
    }
    return V3(0.0f);
}
V3 sample_material(const Isect& is, Lcg& rng, V3 V, V3& L, float& pdf) {         // Material.h:815-845
    pdf = 0.0f;
    switch (is.mat->type) {
        case 0: return sample_diffuse(is, rng, V, L, pdf);
        case 1: return sample_conductor(is, rng, V, L, pdf);
        case 2: return sample_dielectric(is, rng, V, L, pdf);
        case 3: return sample_plastic(is, rng, V, L, pdf);
        case 4: return sample_metal_workflow(is, rng, V, L, pdf);
        case 5: return sample_thin_dielectric(is, rng, V, L, pdf);
        case 6: return sample_clearcoat(is, rng, V, L, pdf);
        case 7: return sample_diffuse_transmitter(is, rng, V, L, pdf);
    }
    return V3(0.0f);
}

// ------------------------------------------------------------- Medium.h
struct Med {
    int type, phase;
    V3 g, sigma_s, sigma_a;
    float scale;
};
Med to_med(const orc_medium& m) {
    return {m.type, m.phase_type, load3(m.g), load3(m.sigma_s), load3(m.sigma_a), m.scale};
}
V3 eval_phase(const Med& m, V3 V, V3 L, float& pdf) {                             // Medium.h:7-43, 83-94
    pdf = 0.0f;
    if (m.phase == 0) {
        pdf = 0.25f * kInvPi;
        return V3(0.25f * kInvPi);
    }
    if (m.phase == 1) {
        float c = dot(L, V);
        V3 att(0.0f);
        float p = 0.0f;
        for (int d = 0; d < 3; ++d) {
            float t = 1.0f + m.g[d] * m.g[d] + 2.0f * m.g[d] * c;
            att[d] = (0.25f * kInvPi) * (1.0f - m.g[d] * m.g[d]) / (t * sqrtf(t));
            p += att[d];
        }
        p *= (1.0f / 3.0f);
        pdf = p;
        if (p <= 0.0f) return V3(0.0f);
        return att;
    }
    return V3(0.0f);
}
V3 sample_phase(const Med& m, Lcg& rng, V3 V, V3& L, float& pdf) {               // Medium.h:14-22, 45-79, 96-108
    pdf = 0.0f;
    if (m.phase == 0) {
        float u0 = rng(), u1 = rng();
        L = uniform_sample_sphere({u0, u1});
        pdf = 0.25f * kInvPi;
        return V3(0.25f * kInvPi);
    }
    if (m.phase == 1) {
        int ch = std::min((int)(rng() * 3), 2);
        float g = m.g[ch];
        float c;
        if (fabsf(g) < kEps) {
            c = 1.0f - 2.0f * rng();
        } else {
            float st = (1.0f - g * g) / (1.0f - g + 2.0f * g * rng());
            c = (1.0f + g * g - st * st) / (2.0f * g);
        }
        float p = 0.0f;
        V3 att(0.0f);
        for (int d = 0; d < 3; ++d) {
            float t = 1.0f + m.g[d] * m.g[d] + 2.0f * m.g[d] * c;
            att[d] = (0.25f * kInvPi) * (1.0f - m.g[d] * m.g[d]) / (t * sqrtf(t));
            p += att[d];
        }
        p *= (1.0f / 3.0f);
        pdf = p;
        if (p <= 0.0f) return V3(0.0f);
        float s = sqrtf(fmaxf(0.0f, 1.0f - c * c));
        float phi = kTwoPi * rng();
        V3 lL = normalize(V3(s * cosf(phi), s * sinf(phi), c));
        L = to_world(lL, V);
        return att;
    }
    return V3(0.0f);
}
inline float medium_weight(V3 ss, V3 st) {                                        // Medium.h:116-125
    V3 alb = ss / st;
    float w = 0.0f;
    for (int d = 0; d < 3; ++d)
        if (alb[d] > w && st[d] != 0.0f) w = alb[d];
    if (w > 0.0f) w = fmaxf(w, 0.5f);
    return w;
}
V3 eval_medium_distance(const Med& m, bool scattered, float dist, float& tpdf) {  // Medium.h:112-158, 222-231
    tpdf = 0.0f;
    if (m.type != 0) return V3(0.0f);
    V3 ss = m.sigma_s * m.scale, sa = m.sigma_a * m.scale, st = sa + ss;
    float w = medium_weight(ss, st);
    V3 att(0.0f);
    float pdf = 0.0f;
    bool valid = false;
    for (int d = 0; d < 3; ++d) {
        att[d] = expf(-st[d] * dist);
        if (att[d] > 0.0f) valid = true;
    }
    if (scattered) {
        for (int d = 0; d < 3; ++d) pdf += st[d] * att[d];
        pdf *= w * (1.0f / 3.0f);
        att = att * ss;
    } else {
        for (int d = 0; d < 3; ++d) pdf += att[d];
        pdf = w * (1.0f / 3.0f) * pdf + (1.0f - w);
    }
    tpdf = pdf;
    if (!valid) att = V3(0.0f);
    return att;
}
bool sample_medium_distance(const Med& m, float maxd, float& dist, float& tpdf, V3& T, Lcg& rng) { // Medium.h:160-218, 233-244
    tpdf = 0.0f;
    T = V3(0.0f);
    if (m.type != 0) return false;
    V3 ss = m.sigma_s * m.scale, sa = m.sigma_a * m.scale, st = sa + ss;
    float w = medium_weight(ss, st);
    bool scattered = false;
    float xi = rng();
    if (xi < w) {
        xi /= w;
        int ch = std::min((int)(rng() * 3), 2);
        dist = -logf(1.0f - xi) / st[ch];
        if (dist < maxd) {
            tpdf = 0.0f;
            for (int d = 0; d < 3; ++d) tpdf += st[d] * expf(-st[d] * dist);
            tpdf *= w * (1.0f / 3.0f);
            scattered = true;
        }
    }
    if (!scattered) {
        dist = maxd;
        tpdf = 0.0f;
        for (int d = 0; d < 3; ++d) tpdf += expf(-st[d] * dist);
        tpdf = w * (1.0f / 3.0f) * tpdf + (1.0f - w);
    }
    bool valid = false;
    for (int d = 0; d < 3; ++d) {
        T[d] = expf(-st[d] * dist);
        if (T[d] > 0.0f) valid = true;
    }
    if (scattered) T = T * ss;
    if (!valid) T = V3(0.0f);
    return scattered;
}

// ------------------------------------------------------------- Light.h
struct Lgt {
    int type;
    bool double_side;
    V3 radiance, position, u, v;
    float radius;
    int medium;
};
Lgt to_lgt(const orc_light& l) {
    return {l.type, l.double_side != 0, load3(l.radiance), load3(l.position), load3(l.u), load3(l.v),
            l.radius, l.medium};
}
struct Ray { V3 o, d; };
float intersect_quad(const Ray& r, V3 corner, V3 _u, V3 _v) {                     // Light.h:6-26
    V3 n = normalize(cross(_u, _v));
    V3 u = _u * (1.0f / dot(_u, _u));
    V3 v = _v * (1.0f / dot(_v, _v));
    float dt = dot(r.d, n);
    float t = (dot(n, corner) - dot(n, r.o)) / dt;
    if (t > kEps) {
        V3 p = r.o + r.d * t;
        V3 vi = p - corner;
        float a1 = dot(u, vi);
        if (a1 >= 0.0f && a1 <= 1.0f) {
            float a2 = dot(v, vi);
            if (a2 >= 0.0f && a2 <= 1.0f) return t;
        }
    }
    return FLT_MAX;
}
float intersect_sphere(const Ray& r, V3 c, float radius) {                        // Light.h:78-98
    V3 co = c - r.o;
    float b = dot(co, r.d);
    float det = b * b - dot(co, co) + radius * radius;
    if (det < 0.0f) return FLT_MAX;
    det = sqrtf(det);
    float t1 = b - det;
    if (t1 > kEps) return t1;
    float t2 = b + det;
    if (t2 > kEps) return t2;
    return FLT_MAX;
}
V3 eval_light(const Lgt& l, const Ray& r, float dist, float& pdf, float& ldist) {  // Light.h:28-52, 100-116, 145-156
    if (l.type == 0) {
        V3 n = normalize(cross(l.u, l.v));
        float c = dot(r.d, n);
        if (!l.double_side && c < 0.0f) { pdf = 0.0f; return V3(0.0f); }
        ldist = intersect_quad(r, l.position, l.u, l.v);
        float area = length(l.u) * length(l.v);
        if (ldist < dist) {
            pdf = 1.0f / area;
            pdf *= sqr(ldist) / fabsf(c);
            return l.radiance;
        }
        pdf = 0.0f;
        return V3(0.0f);
    }
    if (l.type == 1) {
        ldist = intersect_sphere(r, l.position, l.radius);
        if (ldist < dist) {
            V3 dir = l.position - r.o;
            float d2 = dot(dir, dir);
            float s2 = l.radius * l.radius / d2;
            float c = sqrtf(1.0f - s2);
            pdf = uniform_pdf_cone(c);
            return l.radiance;
        }
        pdf = 0.0f;
        return V3(0.0f);
    }
    pdf = 0.0f;
    return V3(0.0f);
}
V3 sample_light(const Lgt& l, V3 p, V2 s, V3& L, float& dist, float& pdf) {       // Light.h:54-74, 118-141, 158-169
    if (l.type == 0) {
        V3 n = normalize(cross(l.u, l.v));
        V3 pos = l.position + l.u * s.x + l.v * s.y;
        L = pos - p;
        float d2 = dot(L, L);
        dist = sqrtf(d2);
        L = L / dist;
        float c = dot(L, n);
        if (!l.double_side && c < 0.0f) { pdf = 0.0f; return V3(0.0f); }
        float area = length(l.u) * length(l.v);
        pdf = 1.0f / area;
        pdf *= d2 / fabsf(c);
        return l.radiance;
    }
    if (l.type == 1) {
        V3 dir = l.position - p;
        float d2 = dot(dir, dir);
        float inv = 1.0f / sqrtf(d2);
        dir = dir * inv;
        dist = d2 * inv;
        float st = l.radius * inv;
        if (st < 1.0f) {
            float ct = sqrtf(1.0f - st * st);
            V3 lL = uniform_sample_cone(s, ct);
            float ci = lL.z;
            L = to_world(lL, dir);
            pdf = uniform_pdf_cone(ct);
            dist = ci * dist - sqrtf(fmaxf(0.0f, l.radius * l.radius - (1.0f - ci * ci) * d2));
            return l.radiance;
        }
        pdf = 0.0f;
        return V3(0.0f);
    }
    pdf = 0.0f;
    return V3(0.0f);
}

// ------------------------------------------------------------- env cache
// HdrTexture::CalculateHdrCache (Scene.cpp:217-313), same summation order.
void env_cache(int W, int H, const float* hdr, float* cache) {
    float lumSum = 0.0f;
    std::vector<float> pdf((size_t)W * H);
    for (int i = 0; i < H; i++)
        for (int j = 0; j < W; j++) {
            const float* t = hdr + 4 * ((size_t)i * W + j);
            float lum = 0.299f * t[0] + 0.587f * t[1] + 0.114f * t[2];   // Utils.h:76-78
            pdf[(size_t)i * W + j] = lum;
            lumSum += lum;
        }
    for (size_t k = 0; k < pdf.size(); k++) pdf[k] /= lumSum;
    std::vector<float> marg(W, 0.0f);
    for (int j = 0; j < W; j++)
        for (int i = 0; i < H; i++) marg[j] += pdf[(size_t)i * W + j];
    std::vector<float> cdfx = marg;
    for (int j = 1; j < W; j++) cdfx[j] += cdfx[j - 1];
    // conditional CDF per column, stored column-major [j][i]
    std::vector<float> cdfy((size_t)W * H);
    for (int j = 0; j < W; j++) {
        float acc = 0.0f;
        for (int i = 0; i < H; i++) {
            float c = pdf[(size_t)i * W + j] / marg[j];
            acc = (i == 0) ? c : acc + c;
            cdfy[(size_t)j * H + i] = acc;
        }
    }
    for (int j = 0; j < W; j++)
        for (int i = 0; i < H; i++) {
            float xi1 = (float)i / H, xi2 = (float)j / W;
            int x = (int)(std::lower_bound(cdfx.begin(), cdfx.end(), xi1) - cdfx.begin());
            const float* col = cdfy.data() + (size_t)x * H;   // x can equal W: see below
            int y;
            if (x < W) {
                y = (int)(std::lower_bound(col, col + H, xi2) - col);
            } else {
                y = 0;  // the reference reads past the end here; never reached when cdfx[W-1] >= xi1
            }
            float* c = cache + 4 * ((size_t)i * W + j);
            c[0] = (float)x / W;
            c[1] = (float)y / H;
            c[2] = pdf[(size_t)i * W + j];
            c[3] = 0.0f;
        }
}

// ------------------------------------------------------------- LBVH
inline uint32_t expand10(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
inline uint32_t quantize(float c, float lo, float ext) {
    float n = ext > 0.0f ? (c - lo) / ext : 0.0f;
    float q = fminf(fmaxf(n * 1024.0f, 0.0f), 1023.0f);
    return (uint32_t)q;
}
inline V3 centroid(const float* p) {
    V3 a = load3(p), b = load3(p + 3), c = load3(p + 6);
    return (a + b + c) * (1.0f / 3.0f);
}
void morton(int64_t n, const float* pos, uint32_t* codes, float* cb) {
    V3 lo(FLT_MAX), hi(-FLT_MAX);
    for (int64_t i = 0; i < n; i++) {
        V3 c = centroid(pos + 9 * i);
        for (int a = 0; a < 3; a++) { lo[a] = fminf(lo[a], c[a]); hi[a] = fmaxf(hi[a], c[a]); }
    }
    V3 ext = hi - lo;
    for (int64_t i = 0; i < n; i++) {
        V3 c = centroid(pos + 9 * i);
        codes[i] = (expand10(quantize(c.x, lo.x, ext.x)) << 2) | (expand10(quantize(c.y, lo.y, ext.y)) << 1) |
                   expand10(quantize(c.z, lo.z, ext.z));
    }
    if (cb) { cb[0] = lo.x; cb[1] = lo.y; cb[2] = lo.z; cb[3] = hi.x; cb[4] = hi.y; cb[5] = hi.z; }
}
constexpr uint32_t kLeaf = 0x80000000u;
inline int clz32(uint32_t x) { return x ? __builtin_clz(x) : 32; }
struct Bvh {
    int64_t n = 0;
    std::vector<uint32_t> codes, order, left, right, parent;
    std::vector<float> lo, hi;          // (n-1) internal then n leaves, 3 floats each
    std::vector<float> tri;             // sorted triangles, 9 floats
    std::vector<uint32_t> prim;         // sorted position -> triangle id
    std::vector<uint32_t> qnode;        // kQWords per internal node: the traversal's compressed records
};

// Compressed node records.  The traversal reads, per internal node, 32
// bytes: the node's own lower corner (the origin, 3 floats), its children
// (the Karras split gamma: left child gamma, right child gamma + 1, with a
// leaf flag each), one power-of-two scale per axis, and each child's box as
// 8-bit multiples of that scale above the origin, rounded outwards by at
// least half a step (a lower bound of 0 is the origin itself):
//   word 0..2  origin.xyz
//   word 3     gamma | left leaf << 30 | right leaf << 31
//   word 4     e_x, e_y, e_z as signed bytes (scale 2^e)
//   word 5..7  child 0 lo.xyz, hi.xyz, child 1 lo.xyz, hi.xyz (one byte each)
// A slab bound at q steps is at distance fma(q, 2^e * inv, (origin - o) * inv).
constexpr int kQWords = 8;
constexpr int kQuantEmin = -64, kQuantEmax = 63;
constexpr uint32_t kQLeft = 1u << 30, kQRight = 1u << 31, kQGamma = kQLeft - 1;

double pow2(int e) {                     // exact 2^e for e in [-1022, 1023]
    uint64_t bits = (uint64_t)(e + 1023) << 52;
    double x;
    std::memcpy(&x, &bits, 8);
    return x;
}

// Smallest scale exponent whose 253 steps reach the extent.
int quant_exponent(double extent) {
    if (!(extent > 0.0)) return kQuantEmin;  // also NaN
    uint64_t bits;
    std::memcpy(&bits, &extent, 8);
    const int k = (int)((bits >> 52) & 0x7ff) - 1022;  // extent < 2^k for a normal extent
    int e = std::max(kQuantEmin, k - 9);
    while (e < kQuantEmax && extent > 253.0 * pow2(e)) e++;
    return e;
}

// The node record of internal node v (Karras layout: lo/hi of internal v at
// v, of leaf k at n - 1 + k), computed in double so that the rounding is
// outwards whatever the magnitudes.
void quantize_node(int64_t ni, const uint32_t* left, const uint32_t* right, const float* lo, const float* hi,
                   int64_t v, uint32_t* out) {
    const uint32_t cl = left[v], cr = right[v];
    const int64_t il = (cl & kLeaf) ? ni + (cl & ~kLeaf) : cl;
    const int64_t ir = (cr & kLeaf) ? ni + (cr & ~kLeaf) : cr;
    int e[3];
    for (int a = 0; a < 3; a++) {
        std::memcpy(&out[a], &lo[3 * v + a], 4);
        e[a] = quant_exponent((double)hi[3 * v + a] - (double)lo[3 * v + a]);
    }
    out[3] = ((cl & ~kLeaf) & kQGamma) | ((cl & kLeaf) ? kQLeft : 0u) | ((cr & kLeaf) ? kQRight : 0u);
    out[4] = (uint32_t)(uint8_t)(int8_t)e[0] | (uint32_t)(uint8_t)(int8_t)e[1] << 8 |
             (uint32_t)(uint8_t)(int8_t)e[2] << 16;
    uint8_t q[12];
    const int64_t ch[2] = {il, ir};
    for (int c = 0; c < 2; c++)
        for (int a = 0; a < 3; a++) {
            const double inv_step = pow2(-e[a]);
            const double xl = ((double)lo[3 * ch[c] + a] - (double)lo[3 * v + a]) * inv_step;
            const double xh = ((double)hi[3 * ch[c] + a] - (double)lo[3 * v + a]) * inv_step;
            q[6 * c + a] = (uint8_t)(xl - 0.5 >= 0.0 ? std::floor(xl - 0.5) : 0.0);
            q[6 * c + 3 + a] = (uint8_t)(xh + 0.5 < 254.0 ? std::floor(xh + 0.5) + 1.0 : 255.0);
        }
    std::memcpy(&out[5], q, 12);
}
// ---- treelet restructuring (Karras & Aila 2013, "Fast Parallel Construction
// of High-Quality Bounding Volume Hierarchies"), then Karras relabelling ----
// Each round visits the internal nodes bottom up; a node with at least
// 7 << round leaves below it roots a treelet: starting from its two children,
// the treelet leaf (internal) of largest surface area is replaced by its two
// children until there are 7.  Over the subsets of those treelet leaves a
// dynamic program finds the binary tree of least SAH cost
// (kCostNode * area per internal node, plus the treelet leaves' own subtree
// costs); if it is cheaper than the current one, the treelet's internal
// nodes are rewired to it (the root keeps its id, the others are reused in
// the order they were expanded).  The hierarchy is then relabelled as
// Karras's layout of its own leaf order (left to right): every node's leaves
// are consecutive, and internal node ids are the same endpoint rule (the
// left child of a node takes the last leaf position of its range, the right
// child the first), so that the two children of a node are again gamma and
// gamma + 1.
constexpr int kTreeletLeaves = 7;
constexpr int kTreeletRounds = 1;
constexpr float kCostNode = 1.2f, kCostTri = 1.0f;

inline float box_area(const float* lo, const float* hi) {
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return 2.0f * ((dx * dy + dy * dz) + dz * dx);
}

struct TreeletState {
    int64_t ni;
    uint32_t *left, *right, *parent;
    float *lo, *hi, *cost;
    uint32_t* nleaves;
    int64_t index(uint32_t ref) const { return (ref & kLeaf) ? ni + (ref & ~kLeaf) : ref; }
    // internal node v from its two children: box, leaf count and SAH cost
    void join(int64_t v) {
        const int64_t a = index(left[v]), c = index(right[v]);
        for (int k = 0; k < 3; k++) {
            lo[3 * v + k] = fminf(lo[3 * a + k], lo[3 * c + k]);
            hi[3 * v + k] = fmaxf(hi[3 * a + k], hi[3 * c + k]);
        }
        nleaves[v] = nleaves[a] + nleaves[c];
        cost[v] = kCostNode * box_area(&lo[3 * v], &hi[3 * v]) + cost[a] + cost[c];
    }
};

// Optimises the treelet rooted at internal node v (its subtrees are final).
void treelet(TreeletState& T, int64_t v) {
    uint32_t ref[kTreeletLeaves];
    uint32_t inner[kTreeletLeaves];
    int m = 2, ninner = 0;
    ref[0] = T.left[v];
    ref[1] = T.right[v];
    while (m < kTreeletLeaves) {
        int best = -1;
        float best_area = -1.0f;
        for (int i = 0; i < m; i++) {
            if (ref[i] & kLeaf) continue;
            const float a = box_area(&T.lo[3 * (size_t)ref[i]], &T.hi[3 * (size_t)ref[i]]);
            if (a > best_area) { best_area = a; best = i; }
        }
        if (best < 0) break;
        const uint32_t c = ref[best];
        inner[ninner++] = c;
        ref[best] = T.left[c];
        ref[m++] = T.right[c];
    }
    if (m < 3) return;
    const int full = (1 << m) - 1;
    float slo[1 << kTreeletLeaves][3], shi[1 << kTreeletLeaves][3], scost[1 << kTreeletLeaves];
    int part[1 << kTreeletLeaves];
    for (int s = 1; s <= full; s++) {
        if ((s & (s - 1)) == 0) {
            const int64_t x = T.index(ref[__builtin_ctz(s)]);
            for (int k = 0; k < 3; k++) { slo[s][k] = T.lo[3 * x + k]; shi[s][k] = T.hi[3 * x + k]; }
            scost[s] = T.cost[x];
            continue;
        }
        const int top = 31 - __builtin_clz((unsigned)s), rest = s & ~(1 << top);
        const int64_t x = T.index(ref[top]);
        for (int k = 0; k < 3; k++) {
            slo[s][k] = fminf(slo[rest][k], T.lo[3 * x + k]);
            shi[s][k] = fmaxf(shi[rest][k], T.hi[3 * x + k]);
        }
        // best split: p holds the lowest leaf of s; the first strictly cheapest from p = s - 1 down
        const int low = s & -s;
        float best = INFINITY;
        int bp = 0;
        for (int q = (s - 1) & s; q > 0; q = (q - 1) & s) {
            if (!(q & low)) continue;
            const float c = scost[q] + scost[s ^ q];
            if (c < best) { best = c; bp = q; }
        }
        scost[s] = kCostNode * box_area(slo[s], shi[s]) + best;
        part[s] = bp;
    }
    if (!(scost[full] < T.cost[v])) return;
    // rewire: subsets of two or more leaves become internal nodes, v first,
    // then the expanded nodes in order, depth first, left part first
    int next = 0;
    struct Frame { int s; int64_t node; int stage; };
    Frame st[kTreeletLeaves];
    int sp = 0;
    st[sp++] = {full, v, 0};
    auto child_ref = [&](int sub) -> uint32_t {
        if ((sub & (sub - 1)) == 0) return ref[__builtin_ctz(sub)];
        return inner[next++];
    };
    // Pre-order assignment of node ids, then joins in post-order.
    int64_t order_nodes[kTreeletLeaves];
    int order_subsets[kTreeletLeaves];
    int nn = 0;
    while (sp > 0) {
        Frame f = st[--sp];
        const int pl = part[f.s], pr = f.s ^ pl;
        const uint32_t rl = child_ref(pl), rr = child_ref(pr);
        T.left[f.node] = rl;
        T.right[f.node] = rr;
        T.parent[T.index(rl)] = (uint32_t)f.node;
        T.parent[T.index(rr)] = (uint32_t)f.node;
        order_nodes[nn] = f.node;
        order_subsets[nn++] = f.s;
        if ((pr & (pr - 1)) != 0) st[sp++] = {pr, (int64_t)rr, 0};
        if ((pl & (pl - 1)) != 0) st[sp++] = {pl, (int64_t)rl, 0};
    }
    (void)order_subsets;
    for (int i = nn - 1; i >= 0; i--) T.join(order_nodes[i]);
}

void restructure(Bvh& b, int rounds) {
    const int64_t n = b.n, ni = n - 1;
    if (ni < 2) return;
    std::vector<float> cost(2 * n - 1);
    std::vector<uint32_t> nl(2 * n - 1);
    TreeletState T{ni, b.left.data(), b.right.data(), b.parent.data(), b.lo.data(), b.hi.data(), cost.data(), nl.data()};
    for (int64_t k = 0; k < n; k++) {
        nl[ni + k] = 1;
        cost[ni + k] = kCostTri * box_area(&b.lo[3 * (ni + k)], &b.hi[3 * (ni + k)]);
    }
    auto postorder = [&](auto&& visit) {
        std::vector<std::pair<int64_t, int>> st = {{0, 0}};
        while (!st.empty()) {
            auto [v, state] = st.back();
            st.pop_back();
            if (state == 0) {
                st.push_back({v, 1});
                for (uint32_t c : {b.left[v], b.right[v]})
                    if (!(c & kLeaf)) st.push_back({(int64_t)c, 0});
            } else {
                visit(v);
            }
        }
    };
    // each round recomputes a node from its (already optimised) children
    // before its own treelet is considered
    for (int r = 0; r < rounds; r++) {
        const uint32_t gamma = (uint32_t)kTreeletLeaves << r;
        postorder([&](int64_t v) {
            T.join(v);
            if (nl[v] >= gamma) treelet(T, v);
        });
    }
    // relabel: leaf order left to right, Karras ids by the endpoint rule
    std::vector<int64_t> first(2 * n - 1, 0), newid(ni, 0);
    {
        std::vector<int64_t> st = {0};
        while (!st.empty()) {
            const int64_t v = st.back();
            st.pop_back();
            const int64_t a = T.index(b.left[v]), c = T.index(b.right[v]);
            first[a] = first[v];
            first[c] = first[v] + nl[a];
            if (!(b.left[v] & kLeaf)) { newid[a] = first[a] + nl[a] - 1; st.push_back(a); }
            if (!(b.right[v] & kLeaf)) { newid[c] = first[c]; st.push_back(c); }
        }
    }
    auto remap = [&](uint32_t ref) -> uint32_t {
        return (ref & kLeaf) ? (kLeaf | (uint32_t)first[ni + (ref & ~kLeaf)]) : (uint32_t)newid[ref];
    };
    std::vector<uint32_t> L(ni), R(ni), P(2 * n - 1, 0xffffffffu), order(n), prim(n);
    std::vector<float> lo(3 * (2 * n - 1)), hi(3 * (2 * n - 1)), tri(9 * n);
    for (int64_t v = 0; v < ni; v++) {
        const int64_t w = newid[v];
        L[w] = remap(b.left[v]);
        R[w] = remap(b.right[v]);
        std::memcpy(&lo[3 * w], &b.lo[3 * v], 12);
        std::memcpy(&hi[3 * w], &b.hi[3 * v], 12);
    }
    for (int64_t k = 0; k < n; k++) {
        const int64_t q = first[ni + k];
        order[q] = b.order[k];
        prim[q] = b.prim[k];
        std::memcpy(&tri[9 * q], &b.tri[9 * k], 36);
        std::memcpy(&lo[3 * (ni + q)], &b.lo[3 * (ni + k)], 12);
        std::memcpy(&hi[3 * (ni + q)], &b.hi[3 * (ni + k)], 12);
    }
    for (int64_t w = 0; w < ni; w++)
        for (uint32_t c : {L[w], R[w]}) P[(c & kLeaf) ? ni + (c & ~kLeaf) : c] = (uint32_t)w;
    b.left.swap(L);
    b.right.swap(R);
    b.parent.swap(P);
    b.lo.swap(lo);
    b.hi.swap(hi);
    b.order.swap(order);
    b.prim.swap(prim);
    b.tri.swap(tri);
}

void build_bvh(Bvh& b, int64_t n, const float* pos, int rounds = kTreeletRounds) {
    b.n = n;
    b.codes.resize(n);
    b.order.resize(n);
    if (n == 0) return;
    std::vector<uint32_t> raw(n);
    morton(n, pos, raw.data(), nullptr);
    for (int64_t i = 0; i < n; i++) b.order[i] = (uint32_t)i;
    std::stable_sort(b.order.begin(), b.order.end(), [&](uint32_t a, uint32_t c) { return raw[a] < raw[c]; });
    for (int64_t i = 0; i < n; i++) b.codes[i] = raw[b.order[i]];
    const int64_t ni = n - 1;
    b.left.assign(ni > 0 ? ni : 0, 0);
    b.right.assign(ni > 0 ? ni : 0, 0);
    b.parent.assign(2 * n - 1, 0xffffffffu);   // internal ids, then leaves at ni + k
    auto delta = [&](int64_t i, int64_t j) -> int {
        if (j < 0 || j >= n) return -1;
        uint32_t ci = b.codes[i], cj = b.codes[j];
        if (ci != cj) return clz32(ci ^ cj);
        return 32 + clz32((uint32_t)i ^ (uint32_t)j);
    };
    // Karras 2012, Figure 4.
    for (int64_t i = 0; i < ni; i++) {
        int d = (delta(i, i + 1) - delta(i, i - 1)) >= 0 ? 1 : -1;
        int dmin = delta(i, i - d);
        int64_t lmax = 2;
        while (delta(i, i + lmax * d) > dmin) lmax *= 2;
        int64_t l = 0;
        for (int64_t t = lmax / 2; t >= 1; t /= 2)
            if (delta(i, i + (l + t) * d) > dmin) l += t;
        int64_t j = i + l * d;
        int dnode = delta(i, j);
        int64_t s = 0, t = l;
        do {
            t = (t + 1) >> 1;
            if (delta(i, i + (s + t) * d) > dnode) s += t;
        } while (t > 1);
        int64_t gamma = i + s * d + std::min(d, 0);
        uint32_t L = (std::min(i, j) == gamma) ? (kLeaf | (uint32_t)gamma) : (uint32_t)gamma;
        uint32_t R = (std::max(i, j) == gamma + 1) ? (kLeaf | (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
        b.left[i] = L;
        b.right[i] = R;
        b.parent[(L & kLeaf) ? ni + (L & ~kLeaf) : L] = (uint32_t)i;
        b.parent[(R & kLeaf) ? ni + (R & ~kLeaf) : R] = (uint32_t)i;
    }
    b.lo.assign(3 * (2 * n - 1), 0.0f);
    b.hi.assign(3 * (2 * n - 1), 0.0f);
    b.tri.resize(9 * n);
    b.prim.resize(n);
    for (int64_t k = 0; k < n; k++) {
        const float* p = pos + 9 * (int64_t)b.order[k];
        std::memcpy(&b.tri[9 * k], p, 9 * sizeof(float));
        b.prim[k] = b.order[k];
        float* l = &b.lo[3 * (ni + k)];
        float* h = &b.hi[3 * (ni + k)];
        for (int a = 0; a < 3; a++) {
            l[a] = fminf(fminf(p[a], p[3 + a]), p[6 + a]);
            h[a] = fmaxf(fmaxf(p[a], p[3 + a]), p[6 + a]);
        }
    }
    // refit internal nodes bottom-up: children always have a larger
    // internal id than ... not guaranteed, so use an explicit post-order.
    if (ni > 0) {
        std::vector<uint32_t> stack;
        std::vector<uint8_t> done(ni, 0);
        stack.push_back(0);
        while (!stack.empty()) {
            uint32_t v = stack.back();
            bool ready = true;
            for (uint32_t c : {b.left[v], b.right[v]})
                if (!(c & kLeaf) && !done[c]) { stack.push_back(c); ready = false; }
            if (!ready) continue;
            stack.pop_back();
            uint32_t cl = b.left[v], cr = b.right[v];
            int64_t il = (cl & kLeaf) ? ni + (cl & ~kLeaf) : cl;
            int64_t ir = (cr & kLeaf) ? ni + (cr & ~kLeaf) : cr;
            for (int a = 0; a < 3; a++) {
                b.lo[3 * v + a] = fminf(b.lo[3 * il + a], b.lo[3 * ir + a]);
                b.hi[3 * v + a] = fmaxf(b.hi[3 * il + a], b.hi[3 * ir + a]);
            }
            done[v] = 1;
        }
    }
    if (rounds > 0) restructure(b, rounds);
    b.qnode.assign((size_t)kQWords * (ni > 0 ? ni : 0), 0u);
    for (int64_t v = 0; v < ni; v++)
        quantize_node(ni, b.left.data(), b.right.data(), b.lo.data(), b.hi.data(), v, &b.qnode[kQWords * v]);
}

struct RayPre {
    V3 o, d, inv;
    int kx, ky, kz;
    float Sx, Sy, Sz;
};
RayPre prep_ray(V3 o, V3 d) {
    RayPre r;
    r.o = o;
    r.d = d;
    // reciprocal direction for the slab test; a zero component becomes
    // +-1e-20 so that (bound - origin) * inv is never 0 * inf = NaN
    auto safe = [](float x) { return fabsf(x) > 1e-20f ? x : copysignf(1e-20f, x); };
    r.inv = V3(1.0f / safe(d.x), 1.0f / safe(d.y), 1.0f / safe(d.z));
    int kz = 0;
    if (fabsf(d.y) > fabsf(d[kz])) kz = 1;
    if (fabsf(d.z) > fabsf(d[kz])) kz = 2;
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    if (d[kz] < 0.0f) std::swap(kx, ky);
    r.kx = kx; r.ky = ky; r.kz = kz;
    r.Sx = d[kx] / d[kz];
    r.Sy = d[ky] / d[kz];
    r.Sz = 1.0f / d[kz];
    return r;
}
// Watertight ray/triangle (Woop, Benthin, Wald, JCGT 2013), FP32 only.
inline bool tri_test(const RayPre& r, const float* p, float tmin, float& t, float& bu, float& bv) {
    V3 A = load3(p) - r.o, B = load3(p + 3) - r.o, C = load3(p + 6) - r.o;
    float Ax = A[r.kx] - r.Sx * A[r.kz], Ay = A[r.ky] - r.Sy * A[r.kz];
    float Bx = B[r.kx] - r.Sx * B[r.kz], By = B[r.ky] - r.Sy * B[r.kz];
    float Cx = C[r.kx] - r.Sx * C[r.kz], Cy = C[r.ky] - r.Sy * C[r.kz];
    float U = Cx * By - Cy * Bx, Vv = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if ((U < 0.0f || Vv < 0.0f || W < 0.0f) && (U > 0.0f || Vv > 0.0f || W > 0.0f)) return false;
    float det = U + Vv + W;
    if (det == 0.0f) return false;
    float Az = r.Sz * A[r.kz], Bz = r.Sz * B[r.kz], Cz = r.Sz * C[r.kz];
    float T = U * Az + Vv * Bz + W * Cz;
    t = T / det;
    if (!(t > tmin)) return false;
    bu = Vv / det;
    bv = W / det;
    return true;
}
// Slab test of a box given as its six slab distances (x lo/hi, y, z); the
// far distance is widened by 2 ulp.
inline bool slab_test(float tx0, float tx1, float ty0, float ty1, float tz0, float tz1, float tmin, float tbest,
                      float& tnear) {
    float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fminf(tz0, tz1));
    float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1)) * 1.00000024f;
    tnear = tn;
    return tn <= tf && tf >= tmin && tn <= tbest;
}
// The two child boxes of a compressed node record against the ray: per axis
// the origin's distance A = (origin - o) * inv and the step's B = 2^e * inv,
// a bound q steps above the origin at fma(q, B, A).
inline void qnode_test(const RayPre& r, const uint32_t* w, float tmin, float tbest, bool hit[2], float tn[2]) {
    float org[3];
    std::memcpy(org, w, 12);
    const float oc[3] = {r.o.x, r.o.y, r.o.z}, ic[3] = {r.inv.x, r.inv.y, r.inv.z};
    float A[3], B[3];
    for (int a = 0; a < 3; a++) {
        A[a] = (org[a] - oc[a]) * ic[a];
        // ic * 2^e, correctly rounded like ldexpf (2^e, |e| <= 64, is a normal float)
        const uint32_t pbits = (uint32_t)((int)(int8_t)((w[4] >> (8 * a)) & 0xffu) + 127) << 23;
        float p2;
        std::memcpy(&p2, &pbits, 4);
        B[a] = ic[a] * p2;
    }
    uint8_t q[12];
    std::memcpy(q, &w[5], 12);
    for (int c = 0; c < 2; c++) {
        float t[6];
        for (int a = 0; a < 3; a++) {
            t[2 * a] = std::fma((float)q[6 * c + a], B[a], A[a]);
            t[2 * a + 1] = std::fma((float)q[6 * c + 3 + a], B[a], A[a]);
        }
        hit[c] = slab_test(t[0], t[1], t[2], t[3], t[4], t[5], tmin, tbest, tn[c]);
    }
}
struct Hit { float t; uint32_t prim; float u, v; };
constexpr uint32_t kNoHit = 0xffffffffu;

// Canonical ordered traversal: both children of a visited node are fetched
// (2 records), leaves are tested at once in (left, right) order, then the
// nearer internal child is descended and the farther pushed (any hit: the
// farther descended first, which finds the occluder of a ray leaving a
// surface sooner; the answer is the same).
template <bool ANY>
Hit traverse(const Bvh& b, V3 o, V3 d, float tmin, float tmax, uint32_t* cnt) {
    Hit h{tmax, kNoHit, 0.0f, 0.0f};
    uint32_t nn = 0, nt = 0;
    if (b.n == 0) { if (cnt) { cnt[0] = 0; cnt[1] = 0; } return h; }
    RayPre r = prep_ray(o, d);
    const int64_t ni = b.n - 1;
    auto try_leaf = [&](uint32_t k) -> bool {
        float t, u, v;
        nt++;
        if (!tri_test(r, &b.tri[9 * (size_t)k], tmin, t, u, v)) return false;
        uint32_t pid = b.prim[k];
        bool take = ANY ? (t < tmax) : (t < h.t || (t == h.t && h.prim != kNoHit && pid < h.prim));
        if (take) { h.t = t; h.prim = pid; h.u = u; h.v = v; return true; }
        return false;
    };
    if (ni == 0) {
        try_leaf(0);
        if (cnt) { cnt[0] = nn; cnt[1] = nt; }
        return h;
    }
    uint32_t stack[128];
    int sp = 0;
    uint32_t node = 0;
    for (;;) {
        const uint32_t* w = &b.qnode[kQWords * (size_t)node];
        const uint32_t g = w[3] & kQGamma;
        uint32_t c[2] = {(w[3] & kQLeft) ? (kLeaf | g) : g, (w[3] & kQRight) ? (kLeaf | (g + 1)) : g + 1};
        nn += 2;
        float tn[2];
        bool hit[2];
        qnode_test(r, w, tmin, h.t, hit, tn);
        bool done = false;
        for (int k = 0; k < 2 && !done; k++) {   // any hit stops at the first occluder
            if (hit[k] && (c[k] & kLeaf)) {
                hit[k] = false;
                if (try_leaf(c[k] & ~kLeaf) && ANY) done = true;
            }
        }
        if (done) break;
        if (hit[0] && hit[1]) {
            int nearer = (tn[1] < tn[0]) ? 1 : 0;
            if (ANY) nearer = 1 - nearer;  // any hit descends the farther child first
            stack[sp++] = c[1 - nearer];
            node = c[nearer];
        } else if (hit[0]) {
            node = c[0];
        } else if (hit[1]) {
            node = c[1];
        } else {
            if (sp == 0) break;
            node = stack[--sp];
        }
    }
    if (cnt) { cnt[0] = nn; cnt[1] = nt; }
    return h;
}

// ------------------------------------------------------------- integrator
struct Scene {
    Bvh bvh;
    const float* normals = nullptr;
    const float* positions = nullptr;
    const uint32_t* mat_ids = nullptr;
    std::vector<Mat> mats;
    std::vector<Lgt> lights;
    std::vector<Med> media;
    int env_w = 0, env_h = 0;
    const float* env = nullptr;
    std::vector<float> env_cache;
    const float* uvs = nullptr;
    std::vector<orc_texture> textures;
    bool textured = false;
};

// NormalFromTangentToWorld (Utils.h:85-94).
V3 normal_from_tangent(V3 n, V3 t) {
    t = normalize(t * 2.0f - 1.0f);
    V3 up = fabsf(n.z) < 0.9f ? V3(0.0f, 0.0f, 1.0f) : V3(1.0f, 0.0f, 0.0f);
    V3 tx = normalize(cross(up, n));
    V3 ty = normalize(cross(n, tx));
    return normalize(tx * t.x + ty * t.y + n * t.z);
}

// Texture overrides of __closesthit__radiance (devicePrograms.cu:86-115):
// interpolated texture coordinate, then each texture the material names
// replaces its parameter; the normal map bends Ns.
void apply_textures(const Scene& S, uint32_t pid, float w, float u, float v, Mat& m, V3& Ns) {
    float tu, tv;
    if (S.uvs) {
        const float* t = S.uvs + 6 * (size_t)pid;
        tu = w * t[0] + u * t[2] + v * t[4];
        tv = w * t[1] + u * t[3] + v * t[5];
    } else {  // Scene.cpp:160-165: (0,0), (1,0), (0,1)
        tu = w * 0.0f + u * 1.0f + v * 0.0f;
        tv = w * 0.0f + u * 0.0f + v * 1.0f;
    }
    float c[4];
    auto tex = [&](int id) {
        const orc_texture& t = S.textures[id];
        bilinear_rgba8(t.rgba8, t.width, t.height, tu, tv, c);
    };
    if (m.tex[0] >= 0) { tex(m.tex[0]); m.albedo = V3(c[0], c[1], c[2]); }
    if (m.tex[1] >= 0) { tex(m.tex[1]); m.roughness = c[0]; }
    if (m.tex[2] >= 0) { tex(m.tex[2]); m.anisotropy = c[0]; }
    if (m.tex[3] >= 0) { tex(m.tex[3]); m.specular = V3(c[0], c[1], c[2]); }
    if (m.tex[4] >= 0) { tex(m.tex[4]); m.metallic = c[0]; }
    if (m.tex[5] >= 0) { tex(m.tex[5]); m.coat_ru = c[0]; }
    if (m.tex[6] >= 0) { tex(m.tex[6]); m.coat_rv = c[0]; }
    if (m.tex[7] >= 0) { tex(m.tex[7]); Ns = normal_from_tangent(Ns, V3(c[0], c[1], c[2])); }
}
struct Cam {                                                                     // Renderer.cpp:796-813
    V3 position, direction, horizontal, vertical;
    int medium;
};
Cam make_camera(const orc_camera& c, int W, int H) {
    Cam k;
    k.medium = c.medium;
    k.position = load3(c.from);
    k.direction = normalize(load3(c.at) - load3(c.from));
    float cf = c.cos_fovy;
    float aspect = (float)W / (float)H;
    k.horizontal = cf * aspect * normalize(cross(k.direction, load3(c.up)));
    k.vertical = cf * normalize(cross(k.horizontal, k.direction));
    return k;
}
struct Stats { uint64_t ext = 0, shadow = 0, nodes = 0, tris = 0, ext_nodes = 0, ext_tris = 0; };

struct PathOut { V3 radiance, albedo, normal; float depth; };

// __raygen__renderFrame (devicePrograms.cu:170-532) for one (pixel, frame).
PathOut trace_path(const Scene& S, const Cam& cam, const orc_render_params& P, int ix, int iy, int frame,
                   Stats* st) {
    const int W = P.width, Hh = P.height;
    Lcg rng;
    rng.init((uint32_t)(ix + W * iy), (uint32_t)frame);
    const int nL = (int)S.lights.size();

    auto closest = [&](V3 o, V3 d, float tmin, float tmax) {
        uint32_t c[2];
        Hit h = traverse<false>(S.bvh, o, d, tmin, tmax, c);
        if (st) { st->ext++; st->nodes += c[0]; st->tris += c[1]; st->ext_nodes += c[0]; st->ext_tris += c[1]; }
        return h;
    };
    auto occluded = [&](V3 o, V3 d, float tmin, float tmax) {
        uint32_t c[2];
        Hit h = traverse<true>(S.bvh, o, d, tmin, tmax, c);
        if (st) { st->shadow++; st->nodes += c[0]; st->tris += c[1]; }
        return h.prim != kNoHit;
    };
    // lightTrace lambda (devicePrograms.cu:192-211), quirk of light_pdf kept.
    auto light_trace = [&](const Ray& ray, int& mi_in, int& mi_out, V3& Lr, float& lpdf, float& closest_d) {
        bool hitL = false;
        for (int i = 0; i < nL; i++) {
            const Lgt& l = S.lights[i];
            float ld = 0.0f;
            V3 Li = eval_light(l, ray, closest_d, lpdf, ld);
            if (lpdf == 0.0f) continue;
            hitL = true;
            Lr = Li;
            closest_d = ld;
            mi_in = mi_out = l.medium;
        }
        return hitL;
    };
    auto env_fetch = [&](const float* img, V2 uv, float* out) { bilinear(img, S.env_w, S.env_h, uv.x, uv.y, out); };

    PathOut out;
    float j0 = rng(), j1 = rng();
    V2 jit = filter_jitter({j0, j1}, P.filter);
    V2 screen{((float)ix + 0.5f + jit.x) / (float)W, ((float)iy + 0.5f + jit.y) / (float)Hh};
    V3 rd = normalize(cam.direction + (screen.x - 0.5f) * cam.horizontal + (screen.y - 0.5f) * cam.vertical);
    Ray ray{cam.position, rd};
    V3 radiance(0.0f), history(1.0f), V = -ray.d, L = ray.d;
    float light_pdf = 0.0f, bsdf_pdf = 0.0f, phase_pdf = 0.0f;
    V3 bsdf(0.0f), attenuation(0.0f), light_radiance(0.0f);
    float closest_distance = FLT_MAX;
    V3 pre_position = cam.position;
    float pre_pdf = 0.0f, mult_trans_pdf = 1.0f, trans_pdf = 1.0f;
    V3 transmittance(0.0f);
    Isect is;
    is.distance = FLT_MAX;
    is.position = ray.o;
    is.shadeNormal = V3(0.0f);
    is.frontFace = true;
    is.mat = nullptr;
    is.in_medium = is.out_medium = cam.medium;
    bool first = true;
    Mat hit_mat;  // the hit's material after its textures (is.mat points here then)
    V3 pix_albedo(0.0f), pix_normal(0.0f);
    float pix_depth = 1.0f;
    const bool has_env = S.env_w > 0;

    for (int bounce = 0; bounce < P.max_depth; bounce++) {
        // optixTrace radiance + __closesthit__radiance / __miss__radiance
        Hit h = closest(ray.o, ray.d, kEps, FLT_MAX);
        if (h.prim != kNoHit) {
            uint32_t pid = h.prim;
            const float* n = S.normals + 9 * (size_t)pid;
            const float* p = S.positions + 9 * (size_t)pid;
            float w = 1.0f - h.u - h.v;
            V3 Ns = normalize(w * load3(n) + h.u * load3(n + 3) + h.v * load3(n + 6));
            V3 P0 = w * load3(p) + h.u * load3(p + 3) + h.v * load3(p + 6);
            const Mat* m = &S.mats[S.mat_ids[pid]];
            if (S.textured) {
                hit_mat = *m;
                apply_textures(S, pid, w, h.u, h.v, hit_mat, Ns);
                m = &hit_mat;
            }
            is.set_face_normal(ray.d, Ns);
            is.distance = length(P0 - is.position);
            is.position = P0;
            is.mat = m;
            is.in_medium = m->in_medium;
            is.out_medium = m->out_medium;
            if (first) { pix_normal = Ns; pix_albedo = m->albedo; first = false; }
        } else {
            is.distance = FLT_MAX;
            is.in_medium = is.out_medium = cam.medium;
            is.frontFace = true;
            if (first) { pix_albedo = V3(0.0f); pix_normal = V3(0.0f); first = false; }
        }
        closest_distance = is.distance;
        bool hitLight = light_trace(ray, is.in_medium, is.out_medium, light_radiance, light_pdf, closest_distance);
        // depth AOV (no reference counterpart; a denoiser input): first surface
        // or light along the camera ray, as d / (1 + d), 1 for a miss
        if (bounce == 0)
            pix_depth = closest_distance < FLT_MAX ? closest_distance / (1.0f + closest_distance) : 1.0f;

        bool scattered = false;
        int medium_index = -1;
        if (!S.media.empty()) {
            medium_index = is.get_medium(hitLight ? true : is.frontFace);
            float distance = 0.0f;
            if (medium_index != -1) {
                const Med& med = S.media[medium_index];
                scattered = sample_medium_distance(med, closest_distance, distance, trans_pdf, transmittance, rng);
                history = history * (transmittance / trans_pdf);
                mult_trans_pdf *= trans_pdf;
                if (scattered) {
                    is.position = pre_position + distance * L;
                    is.distance = distance;
                    if (nL != 0) {
                        int idx = std::min(std::max((int)((float)nL * rng()), 0), nL - 1);
                        const Lgt& l = S.lights[idx];
                        float ld = 0.0f;
                        Ray sr;
                        sr.o = is.position;
                        float u0 = rng(), u1 = rng();
                        light_radiance = sample_light(l, sr.o, {u0, u1}, sr.d, ld, light_pdf);
                        light_pdf *= (1.0f / (float)nL);
                        V3 tv; float tp; int a, b2;
                        closest_distance = ld - kEps;
                        if (!light_trace(sr, a, b2, tv, tp, closest_distance) &&
                            !occluded(sr.o, sr.d, kEps, ld - kEps)) {
                            transmittance = eval_medium_distance(med, false, closest_distance, trans_pdf);
                            attenuation = eval_phase(med, V, sr.d, phase_pdf);
                            float mw = power_heuristic(light_pdf, phase_pdf * trans_pdf);
                            radiance = radiance + mw * (transmittance / trans_pdf) * light_radiance * attenuation *
                                                      history / light_pdf;
                        }
                    }
                    attenuation = sample_phase(med, rng, V, L, phase_pdf);
                    history = history * (attenuation / phase_pdf);
                    pre_pdf = phase_pdf;
                }
            }
        }

        if (!scattered) {
            // A ray inside a medium that leaves the scene unscattered would
            // shade stale surface data (no surface was hit); it ends here.
            if (!hitLight && is.mat == nullptr && closest_distance == FLT_MAX && medium_index != -1) break;
            if (hitLight) {
                float mw = 1.0f;
                if (bounce != 0) {
                    pre_pdf *= mult_trans_pdf;
                    mw = power_heuristic(pre_pdf, light_pdf);
                }
                radiance = radiance + mw * history * light_radiance;
                break;
            }
            bool notHit = closest_distance == FLT_MAX;
            if (notHit && medium_index == -1) {
                if (has_env) {
                    V2 uv = sphere_to_plane(ray.d);                              // Environment.h:5-11
                    light_pdf = (float)(S.env_w * S.env_h) / (2.0f * sqr(kPi));
                    float e[4], c[4];
                    env_fetch(S.env, uv, e);
                    light_radiance = V3(e[0], e[1], e[2]);
                    env_fetch(S.env_cache.data(), uv, c);
                    light_pdf *= c[2];
                    float mw = 1.0f;
                    if (bounce != 0) {
                        pre_pdf *= mult_trans_pdf;
                        mw = power_heuristic(pre_pdf, light_pdf);
                    }
                    radiance = radiance + mw * history * light_radiance;
                }
                break;
            }
            if (has_env) {
                float x0 = rng(), x1 = rng();
                float uvt[4];
                env_fetch(S.env_cache.data(), {x0, x1}, uvt);
                V2 uv{uvt[0], uvt[1]};
                Ray sr;
                sr.o = is.position;
                sr.d = plane_to_sphere(uv);                                       // Environment.h:13-17
                light_pdf = (float)(S.env_w * S.env_h) / (2.0f * sqr(kPi));
                float c[4], e[4];
                env_fetch(S.env_cache.data(), uv, c);
                light_pdf *= c[2];
                env_fetch(S.env, uv, e);
                light_radiance = V3(e[0], e[1], e[2]);
                V3 tv; float tp; int a, b2;
                closest_distance = FLT_MAX;
                if (!light_trace(sr, a, b2, tv, tp, closest_distance) && !occluded(sr.o, sr.d, kEps, FLT_MAX)) {
                    bsdf = eval_material(is, V, sr.d, bsdf_pdf);
                    float ct = fabsf(dot(is.shadeNormal, sr.d));
                    float mw = power_heuristic(light_pdf, bsdf_pdf);
                    V3 th(1.0f);
                    if (medium_index != -1) {
                        transmittance = eval_medium_distance(S.media[medium_index], false, closest_distance, trans_pdf);
                        th = transmittance / trans_pdf;
                        mw = power_heuristic(light_pdf, bsdf_pdf * trans_pdf);
                    }
                    radiance = radiance + mw * th * light_radiance * bsdf * ct * history / light_pdf;
                }
            }
            if (nL != 0) {
                int idx = std::min(std::max((int)((float)nL * rng()), 0), nL - 1);
                const Lgt& l = S.lights[idx];
                float ld = 0.0f;
                Ray sr;
                sr.o = is.position;
                float u0 = rng(), u1 = rng();
                light_radiance = sample_light(l, sr.o, {u0, u1}, sr.d, ld, light_pdf);
                light_pdf *= (1.0f / (float)nL);
                V3 tv; float tp; int a, b2;
                closest_distance = ld - kEps;
                if (!light_trace(sr, a, b2, tv, tp, closest_distance) && !occluded(sr.o, sr.d, kEps, ld - kEps)) {
                    bsdf = eval_material(is, V, sr.d, bsdf_pdf);
                    float ct = fabsf(dot(is.shadeNormal, sr.d));
                    float mw = power_heuristic(light_pdf, bsdf_pdf);
                    V3 th(1.0f);
                    if (medium_index != -1) {
                        transmittance = eval_medium_distance(S.media[medium_index], false, closest_distance, trans_pdf);
                        th = transmittance / trans_pdf;
                        mw = power_heuristic(light_pdf, bsdf_pdf * trans_pdf);
                    }
                    radiance = radiance + mw * th * light_radiance * bsdf * ct * history / light_pdf;
                }
            }
            bsdf = sample_material(is, rng, V, L, bsdf_pdf);
            float ct = fabsf(dot(is.shadeNormal, L));
            history = history * (bsdf * ct / bsdf_pdf);
            pre_pdf = bsdf_pdf;
        }

        if (bounce >= P.rr_start) {                                              // devicePrograms.cu:481-485
            float prr = fminf((history.x + history.y + history.z) / 3.0f, 0.95f);
            if (rng() > prr) break;
            history = history / prr;
        }
        if (is_nan(history)) break;
        V = -L;
        ray.o = is.position;
        ray.d = L;
        mult_trans_pdf = 1.0f;
        pre_position = is.position;
    }
    out.radiance = radiance;
    out.albedo = pix_albedo;
    out.normal = pix_normal;
    out.depth = pix_depth;
    return out;
}

void load_scene(Scene& S, const orc_scene* sc) {
    build_bvh(S.bvh, sc->n_tris, sc->positions);
    S.positions = sc->positions;
    S.normals = sc->normals;
    S.mat_ids = sc->mat_ids;
    for (int i = 0; i < sc->n_materials; i++) S.mats.push_back(to_mat(sc->materials[i]));
    for (int i = 0; i < sc->n_lights; i++) S.lights.push_back(to_lgt(sc->lights[i]));
    for (int i = 0; i < sc->n_media; i++) S.media.push_back(to_med(sc->media[i]));
    S.env_w = sc->env_w;
    S.env_h = sc->env_h;
    S.env = sc->env_rgba;
    S.uvs = sc->uvs;
    for (int i = 0; i < sc->n_textures; i++) S.textures.push_back(sc->textures[i]);
    for (const Mat& m : S.mats)
        for (int i = 0; i < 8; i++) S.textured = S.textured || m.tex[i] >= 0;
    if (S.env_w > 0 && S.env_h > 0) {
        S.env_cache.resize((size_t)4 * S.env_w * S.env_h);
        env_cache(S.env_w, S.env_h, S.env, S.env_cache.data());
    } else {
        S.env_w = S.env_h = 0;
    }
}

template <typename F>
void parallel_rows(int rows, int threads, F&& f) {
    if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
    if (threads < 1) threads = 1;
    std::atomic<int> next(0);
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++)
        pool.emplace_back([&, t] {
            for (;;) {
                int r = next.fetch_add(1);
                if (r >= rows) break;
                f(r, t);
            }
        });
    for (auto& th : pool) th.join();
}

}  // namespace

// ====================================================================== C API
extern "C" {

uint32_t orc_lcg_init(uint32_t v0, uint32_t v1) { Lcg r; r.init(v0, v1); return r.state; }
float orc_lcg_next(uint32_t* state) { Lcg r; r.state = *state; float f = r(); *state = r.state; return f; }

void orc_morton_codes(int64_t n, const float* positions, uint32_t* codes, float* cbounds6) {
    morton(n, positions, codes, cbounds6);
}

int orc_lbvh_build(int64_t n, const float* positions, int32_t rounds, uint32_t* codes, uint32_t* order,
                   uint32_t* left, uint32_t* right, uint32_t* parent, float* lo, float* hi, uint32_t* qnode) {
    Bvh b;
    build_bvh(b, n, positions, rounds < 0 ? kTreeletRounds : rounds);
    if (n == 0) return 0;
    std::memcpy(codes, b.codes.data(), n * 4);
    std::memcpy(order, b.order.data(), n * 4);
    if (n > 1) {
        std::memcpy(left, b.left.data(), (n - 1) * 4);
        std::memcpy(right, b.right.data(), (n - 1) * 4);
    }
    if (parent) std::memcpy(parent, b.parent.data(), (2 * n - 1) * 4);
    std::memcpy(lo, b.lo.data(), (2 * n - 1) * 12);
    std::memcpy(hi, b.hi.data(), (2 * n - 1) * 12);
    if (qnode && n > 1) std::memcpy(qnode, b.qnode.data(), b.qnode.size() * 4);
    return 0;
}

void* orc_bvh_create(int64_t n, const float* positions) {
    Bvh* b = new Bvh;
    build_bvh(*b, n, positions);
    return b;
}
void orc_bvh_destroy(void* bvh) { delete static_cast<Bvh*>(bvh); }

void orc_bvh_closest(const void* bvh, int64_t n_rays, const float* rays, float* hits, uint32_t* counts) {
    const Bvh& b = *static_cast<const Bvh*>(bvh);
    for (int64_t i = 0; i < n_rays; i++) {
        const float* r = rays + 8 * i;
        Hit h = traverse<false>(b, load3(r), load3(r + 4), r[3], r[7], counts ? counts + 2 * i : nullptr);
        hits[4 * i] = h.t;
        std::memcpy(&hits[4 * i + 1], &h.prim, 4);
        hits[4 * i + 2] = h.u;
        hits[4 * i + 3] = h.v;
    }
}
void orc_bvh_any(const void* bvh, int64_t n_rays, const float* rays, uint8_t* occ, uint32_t* counts) {
    const Bvh& b = *static_cast<const Bvh*>(bvh);
    for (int64_t i = 0; i < n_rays; i++) {
        const float* r = rays + 8 * i;
        Hit h = traverse<true>(b, load3(r), load3(r + 4), r[3], r[7], counts ? counts + 2 * i : nullptr);
        occ[i] = h.prim != kNoHit;
    }
}
void orc_brute_closest(int64_t n, const float* pos, int64_t n_rays, const float* rays, float* hits) {
    for (int64_t i = 0; i < n_rays; i++) {
        const float* r = rays + 8 * i;
        RayPre rp = prep_ray(load3(r), load3(r + 4));
        Hit h{r[7], kNoHit, 0.0f, 0.0f};
        for (int64_t k = 0; k < n; k++) {
            float t, u, v;
            if (!tri_test(rp, pos + 9 * k, r[3], t, u, v)) continue;
            uint32_t pid = (uint32_t)k;
            if (t < h.t || (t == h.t && h.prim != kNoHit && pid < h.prim)) { h.t = t; h.prim = pid; h.u = u; h.v = v; }
        }
        hits[4 * i] = h.t;
        std::memcpy(&hits[4 * i + 1], &h.prim, 4);
        hits[4 * i + 2] = h.u;
        hits[4 * i + 3] = h.v;
    }
}

void orc_kulla_conty(float* bsdf_avg, float* albedo) { kulla_conty(bsdf_avg, albedo); }

void orc_sample_ggx(float au, float av, float u0, float u1, float* h3) {
    V3 h = sample_ggx(au, av, V2{u0, u1});
    h3[0] = h.x; h3[1] = h.y; h3[2] = h.z;
}
void orc_fresnel_conductor(const float* V, const float* H, const float* eta, const float* k, float* out3) {
    V3 r = fresnel_conductor(load3(V), load3(H), load3(eta), load3(k));
    out3[0] = r.x; out3[1] = r.y; out3[2] = r.z;
}
float orc_fresnel_dielectric(const float* V, const float* H, float eta_inv) {
    return fresnel_dielectric(load3(V), load3(H), eta_inv);
}
void orc_average_fresnel_conductor(const float* eta, const float* k, float* out3) {
    V3 r = average_fresnel_conductor(load3(eta), load3(k));
    out3[0] = r.x; out3[1] = r.y; out3[2] = r.z;
}
float orc_average_fresnel_dielectric(float eta) { return average_fresnel_dielectric(eta); }
float orc_ggx_d(const float* H, float au, float av) { return ggx_d(load3(H), V3(0.0f, 0.0f, 1.0f), au, av); }
float orc_smith_g1(const float* V, const float* H, float au, float av) {
    return smith_g1(load3(V), load3(H), V3(0.0f, 0.0f, 1.0f), au, av);
}

void orc_material_eval(const orc_material* m, const float* n, int front_face, const float* V, const float* L,
                       float* f, float* pdf) {
    Mat mm = to_mat(*m);
    Isect is;
    is.shadeNormal = load3(n);
    is.frontFace = front_face != 0;
    is.mat = &mm;
    float p = 0.0f;
    V3 r = eval_material(is, load3(V), load3(L), p);
    f[0] = r.x; f[1] = r.y; f[2] = r.z;
    *pdf = p;
}
void orc_material_sample(const orc_material* m, const float* n, int front_face, const float* V,
                         uint32_t* rng_state, float* f, float* pdf, float* L) {
    Mat mm = to_mat(*m);
    Isect is;
    is.shadeNormal = load3(n);
    is.frontFace = front_face != 0;
    is.mat = &mm;
    Lcg rng;
    rng.state = *rng_state;
    float p = 0.0f;
    V3 Lw(0.0f);
    V3 r = sample_material(is, rng, load3(V), Lw, p);
    *rng_state = rng.state;
    f[0] = r.x; f[1] = r.y; f[2] = r.z;
    *pdf = p;
    L[0] = Lw.x; L[1] = Lw.y; L[2] = Lw.z;
}
void orc_medium_sample_distance(const orc_medium* m, float maxd, uint32_t* rng_state, float* out) {
    Med md = to_med(*m);
    Lcg rng;
    rng.state = *rng_state;
    float d = 0.0f, tp = 0.0f;
    V3 T;
    bool s = sample_medium_distance(md, maxd, d, tp, T, rng);
    *rng_state = rng.state;
    out[0] = d; out[1] = tp; out[2] = T.x; out[3] = T.y; out[4] = T.z; out[5] = s ? 1.0f : 0.0f;
}
void orc_medium_eval_distance(const orc_medium* m, int scattered, float distance, float* out) {
    Med md = to_med(*m);
    float tp = 0.0f;
    V3 T = eval_medium_distance(md, scattered != 0, distance, tp);
    out[0] = tp; out[1] = T.x; out[2] = T.y; out[3] = T.z;
}
void orc_phase_eval(const orc_medium* m, const float* V, const float* L, float* out) {
    Med md = to_med(*m);
    float p = 0.0f;
    V3 f = eval_phase(md, load3(V), load3(L), p);
    out[0] = f.x; out[1] = f.y; out[2] = f.z; out[3] = p;
}
void orc_phase_sample(const orc_medium* m, const float* V, uint32_t* rng_state, float* out) {
    Med md = to_med(*m);
    Lcg rng;
    rng.state = *rng_state;
    float p = 0.0f;
    V3 L(0.0f);
    V3 f = sample_phase(md, rng, load3(V), L, p);
    *rng_state = rng.state;
    out[0] = f.x; out[1] = f.y; out[2] = f.z; out[3] = p; out[4] = L.x; out[5] = L.y; out[6] = L.z;
}
void orc_filter_jitter(int type, float u0, float u1, float* out2) {
    V2 j = filter_jitter({u0, u1}, type);
    out2[0] = j.x;
    out2[1] = j.y;
}
void orc_env_cache(int w, int h, const float* rgba, float* cache) { env_cache(w, h, rgba, cache); }
void orc_bilinear(int w, int h, const float* rgba, float u, float v, float* out4) { bilinear(rgba, w, h, u, v, out4); }
void orc_texture_sample(int w, int h, const uint8_t* rgba8, float u, float v, float* out4) {
    bilinear_rgba8(rgba8, w, h, u, v, out4);
}
void orc_normal_from_tangent(const float* n, const float* t, float* out3) {
    V3 r = normal_from_tangent(load3(n), load3(t));
    out3[0] = r.x; out3[1] = r.y; out3[2] = r.z;
}

// PostProcessing.cu:3-48
void orc_tonemap(int64_t n, const float* rgba, uint32_t* out) {
    for (int64_t i = 0; i < n; i++) {
        const float* c = rgba + 4 * i;
        float lum = 0.299f * c[0] + 0.587f * c[1] + 0.114f * c[2];
        float factor = 1.0f / (1.0f + lum / 1.5f);
        float f[3] = {c[0] * factor, c[1] * factor, c[2] * factor};
        uint32_t px = 0;
        for (int k = 0; k < 3; k++) {
            float g = powf(f[k], 1.0f / 2.2f);
            g = fminf(1.0f, fmaxf(0.0f, g));
            px |= (uint32_t)(g * 255.9f) << (8 * k);
        }
        px |= 255u << 24;
        out[i] = px;
    }
}

int orc_render(const orc_scene* sc, const orc_camera* cam_in, const orc_render_params* p, float* color,
               float* albedo, float* normal, float* depth, float* path_radiance) {
    if (p->width <= 0 || p->height <= 0 || p->spp < 0) return -1;
    Scene S;
    load_scene(S, sc);
    Cam cam = make_camera(*cam_in, p->width, p->height);
    const int W = p->width, H = p->height;
    parallel_rows(H, p->threads, [&](int iy, int) {
        for (int ix = 0; ix < W; ix++) {
            size_t pix = (size_t)iy * W + ix;
            double dc[4], da[4], dn[4], dd = depth ? depth[pix] : 0.0;
            for (int k = 0; k < 4; k++) { dc[k] = color[4 * pix + k]; da[k] = albedo[4 * pix + k]; dn[k] = normal[4 * pix + k]; }
            for (int s = 0; s < p->spp; s++) {
                int frame = p->frame_base + s;
                PathOut o = trace_path(S, cam, *p, ix, iy, frame, nullptr);
                if (path_radiance) {
                    float* pr = path_radiance + 3 * ((size_t)s * W * H + pix);
                    pr[0] = o.radiance.x; pr[1] = o.radiance.y; pr[2] = o.radiance.z;
                }
                // devicePrograms.cu:499-531 with numPixelSamples = 1
                float c[4] = {o.radiance.x, o.radiance.y, o.radiance.z, 1.0f};
                float a[4] = {o.albedo.x, o.albedo.y, o.albedo.z, 1.0f};
                float nn[4] = {o.normal.x, o.normal.y, o.normal.z, 1.0f};
                float z = o.depth;
                if (is_nan(o.radiance)) {
                    c[0] = c[1] = c[2] = 0.0f;
                    a[0] = a[1] = a[2] = 0.0f;
                    nn[0] = nn[1] = nn[2] = 0.0f;
                    z = 0.0f;
                }
                if (p->fp64_accum) {
                    dd = frame > 0 ? (z + (double)frame * dd) / (frame + 1.0) : z;
                    for (int k = 0; k < 4; k++) {
                        if (frame > 0) {
                            dc[k] = (c[k] + (double)frame * dc[k]) / (frame + 1.0);
                            da[k] = (a[k] + (double)frame * da[k]) / (frame + 1.0);
                            dn[k] = (nn[k] + (double)frame * dn[k]) / (frame + 1.0);
                        } else {
                            dc[k] = c[k]; da[k] = a[k]; dn[k] = nn[k];
                        }
                    }
                } else {
                    float* pc = color + 4 * pix;
                    float* pa = albedo + 4 * pix;
                    float* pn = normal + 4 * pix;
                    if (depth) {
                        if (frame > 0) {
                            z += (float)frame * depth[pix];
                            z /= ((float)frame + 1.0f);
                        }
                        depth[pix] = z;
                    }
                    for (int k = 0; k < 4; k++) {
                        if (frame > 0) {
                            c[k] += (float)frame * pc[k];
                            c[k] /= ((float)frame + 1.0f);
                            a[k] += (float)frame * pa[k];
                            a[k] /= ((float)frame + 1.0f);
                            nn[k] += (float)frame * pn[k];
                            nn[k] /= ((float)frame + 1.0f);
                        }
                        pc[k] = c[k]; pa[k] = a[k]; pn[k] = nn[k];
                    }
                }
            }
            if (p->fp64_accum && depth) depth[pix] = (float)dd;
            if (p->fp64_accum)
                for (int k = 0; k < 4; k++) {
                    color[4 * pix + k] = (float)dc[k];
                    albedo[4 * pix + k] = (float)da[k];
                    normal[4 * pix + k] = (float)dn[k];
                }
        }
    });
    return 0;
}

int orc_render_stats(const orc_scene* sc, const orc_camera* cam_in, const orc_render_params* p, uint64_t* out4) {
    if (p->width <= 0 || p->height <= 0) return -1;
    Scene S;
    load_scene(S, sc);
    Cam cam = make_camera(*cam_in, p->width, p->height);
    int threads = p->threads <= 0 ? (int)std::thread::hardware_concurrency() : p->threads;
    if (threads < 1) threads = 1;
    std::vector<Stats> per(threads);
    const int step = p->row_step > 1 ? p->row_step : 1;
    const int rows = (p->height + step - 1) / step;
    auto t0 = std::chrono::steady_clock::now();
    parallel_rows(rows, threads, [&](int r, int t) {
        const int iy = r * step;
        for (int ix = 0; ix < p->width; ix++)
            for (int s = 0; s < p->spp; s++) trace_path(S, cam, *p, ix, iy, p->frame_base + s, &per[t]);
    });
    auto t1 = std::chrono::steady_clock::now();
    Stats tot;
    for (auto& s : per) {
        tot.ext += s.ext; tot.shadow += s.shadow; tot.nodes += s.nodes; tot.tris += s.tris;
        tot.ext_nodes += s.ext_nodes; tot.ext_tris += s.ext_tris;
    }
    out4[0] = tot.ext; out4[1] = tot.shadow; out4[2] = tot.nodes; out4[3] = tot.tris;
    out4[4] = tot.ext_nodes; out4[5] = tot.ext_tris;
    out4[6] = (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count();
    out4[7] = (uint64_t)rows * p->width * p->spp;
    return 0;
}

}  // extern "C"
