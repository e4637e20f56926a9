// Device shading model: Fresnel/GGX/Kulla-Conty (core/cuda/Microfacet.h),
// the eight BSDFs (core/cuda/Material.h), homogeneous media and phase
// functions (core/cuda/Medium.h), analytic quad/sphere lights
// (core/cuda/Light.h).  Semantics follow the reference line by line; the
// data layout is the build's: one 128-byte oxr_material per surface slot in
// HBM, read through the read-only path, and a 64x64+64 Kulla-Conty table.
#pragma once
#include "dmath.h"
#include "../../include/oxr.h"

namespace oxr {

constexpr int kLut = 64;  // Microfacet.h:218

struct Lut {
    const float* bsdf;    // 64*64, index [roughness][cos]
    const float* albedo;  // 64
};

// The shading point as the reference's Interaction (Utils.h:62-74).
struct Surf {
    f3 n;            // shadeNormal
    bool front;      // frontFace
    const oxr_material* m;
};

// ---------------------------------------------------------------- Fresnel
OXR_D f3 fresnel_schlick(f3 f0, float VdotH) {                       // Microfacet.h:14-20
    float t = 1.0f - clampf(VdotH, 0.0f, 1.0f);
    float t2 = t * t;
    float fc = t2 * t2 * t;
    return f0 + (1.0f - f0) * fc;
}
OXR_D float fresnel_conductor_ch(float c, float c2, float s2, float s4, float er, float ei) {
    // one channel of FresnelConductor (Microfacet.h:22-46)
    float t1 = er * er - ei * ei - s2;
    float a2pb2 = t1 * t1 + 4.0f * ei * ei * er * er;
    a2pb2 = sqrtf(fmaxf(0.0f, a2pb2));
    float a = 0.5f * (a2pb2 + t1);
    a = sqrtf(fmaxf(0.0f, a));
    float term1 = a2pb2 + s2, term2 = 2.0f * c * a, term3 = a2pb2 * c2 + s4, term4 = term2 * s2;
    float rs = (term1 - term2) / (term1 + term2);
    float rp = rs * (term3 - term4) / (term3 + term4);
    return 0.5f * (rs + rp);
}
OXR_D f3 fresnel_conductor(f3 V, f3 H, f3 er, f3 ei) {
    float c = dot(V, H), c2 = c * c, s2 = 1.0f - c2, s4 = s2 * s2;
    return f3(fresnel_conductor_ch(c, c2, s2, s4, er.x, ei.x), fresnel_conductor_ch(c, c2, s2, s4, er.y, ei.y),
              fresnel_conductor_ch(c, c2, s2, s4, er.z, ei.z));
}
OXR_D float avg_fresnel_conductor_ch(float eta, float k) {              // Microfacet.h:48-70
    float r = (sqr(eta - 1.0f) + sqr(k)) / (sqr(eta + 1.0f) + sqr(k));
    float t1 = 1.0f + sqrtf(r), t2 = 1.0f - sqrtf(r), t3 = (1.0f - r) / (1.0f + r);
    float e = (t1 - eta * t2) / (t1 - t3 * t2);
    float v = 0.087237f;
    v = v + 0.0230685f * e;
    v = v - 0.0864902f * e * e;
    v = v + 0.0774594f * e * e * e;
    v = v + 0.782654f * r;
    v = v - 0.136432f * r * r;
    v = v + 0.278708f * r * r * r;
    v = v + 0.19744f * e * r;
    v = v + 0.0360605f * e * e * r;
    v = v - 0.2586f * e * r * r;
    return v;
}
OXR_D f3 avg_fresnel_conductor(f3 eta, f3 k) {
    return f3(avg_fresnel_conductor_ch(eta.x, k.x), avg_fresnel_conductor_ch(eta.y, k.y),
              avg_fresnel_conductor_ch(eta.z, k.z));
}
OXR_D float fresnel_dielectric(f3 V, f3 H, float eta_inv) {            // Microfacet.h:72-85
    float ci = fabsf(dot(V, H));
    float ct2 = 1.0f - sqr(eta_inv) * (1.0f - sqr(ci));
    if (ct2 <= 0.0f) return 1.0f;
    float ct = sqrtf(ct2);
    float rs = (eta_inv * ci - ct) / (eta_inv * ci + ct);
    float rp = (ci - eta_inv * ct) / (ci + eta_inv * ct);
    return (rs * rs + rp * rp) / 2.0f;
}
OXR_D float avg_fresnel_dielectric(float eta) {                         // Microfacet.h:87-113
    if (eta < 1.0f) return -1.4399f * (eta * eta) + 0.7099f * eta + 0.6681f + 0.0636f / eta;
    float i1 = 1.0f / eta, i2 = i1 * i1, i3 = i2 * i1, i4 = i3 * i1, i5 = i4 * i1;
    return 0.919317f - 3.4793f * i1 + 6.75335f * i2 - 7.80989f * i3 + 4.98554f * i4 - 1.36881f * i5;
}

// ---------------------------------------------------------------- GGX
OXR_D float smith_g1(f3 V, f3 H, f3 N, float au, float av) {           // Microfacet.h:117-142
    float c = dot(V, N);
    if (c * dot(V, H) <= 0.0f) return 0.0f;
    if (fabsf(c - 1.0f) < kEps) return 1.0f;
    if (au == av) {
        float c2 = sqr(c), t2 = (1.0f - c2) / c2, a2 = au * au;
        return 2.0f / (1.0f + sqrtf(1.0f + a2 * t2));
    }
    f3 d = to_local(V, N);
    float xy = sqr(au * d.x) + sqr(av * d.y), ta2 = xy / sqr(d.z);
    return 2.0f / (1.0f + sqrtf(1.0f + ta2));
}
OXR_D float ggx_d(f3 H, f3 N, float au, float av) {                    // Microfacet.h:144-160
    float c = dot(H, N);
    if (c <= 0.0f) return 0.0f;
    float c2 = sqr(c), t2 = (1.0f - c2) / c2, a2 = au * av;
    if (au == av) return a2 / (kPi * powf(c, 2.0f) * sqr(a2 + t2));
    f3 d = to_local(H, N);
    return 1.0f / (kPi * a2 * sqr(sqr(d.x / au) + sqr(d.y / av) + sqr(d.z)));
}
OXR_D float ggx_dv(f3 V, f3 H, f3 N, float au, float av) {            // Microfacet.h:162-164
    return smith_g1(V, H, N, au, av) * dot(V, H) * ggx_d(H, N, au, av) / dot(N, V);
}
OXR_D f3 sample_vndf(f3 N, f3 Ve, float au, float av, f2 s) {          // Microfacet.h:188-214
    f3 V = to_local(Ve, N);
    f3 Vh = normalize(f3(au * V.x, av * V.y, V.z));
    float len2 = sqr(Vh.x) + sqr(Vh.y);
    f3 T1 = len2 > 0.0f ? f3(-Vh.y, Vh.x, 0.0f) * sqrtf(1.0f / len2) : f3(1.0f, 0.0f, 0.0f);
    f3 T2 = cross(Vh, T1);
    float r = sqrtf(s.x), phi = s.y * kTwoPi;
    float t1 = r * cosf(phi), t2 = r * sinf(phi);
    float sw = 0.5f * (1.0f + Vh.z);
    t2 = (1.0f - sw) * sqrtf(1.0f - sqr(t1)) + sw * t2;
    f3 Nh = t1 * T1 + t2 * T2 + sqrtf(fmaxf(0.0f, 1.0f - sqr(t1) - sqr(t2))) * Vh;
    return normalize(f3(au * Nh.x, av * Nh.y, fmaxf(0.0f, Nh.z)));
}

// Kulla-Conty lookups (Microfacet.h:293-346); the table is indexed at
// roughness^2 * 64 although built at roughness (i+0.5)/64.
OXR_D int lut_index(float x) { return (x == x) ? (int)x : 0; }
OXR_D float lut_bsdf(float c, float r, const float* t) {
    float o1 = sqr(r) * kLut, o2 = c * kLut;
    int i1 = lut_index(o1), i2 = lut_index(o2);
    if (i1 >= kLut - 1) {
        if (i2 >= kLut - 1) return t[(kLut - 1) * kLut + kLut - 1];
        return mixf(t[(kLut - 1) * kLut + i2], t[(kLut - 1) * kLut + i2 + 1], o2 - i2);
    }
    if (i2 >= kLut - 1) return mixf(t[i1 * kLut + kLut - 1], t[(i1 + 1) * kLut + kLut - 1], o1 - i1);
    return mixf(mixf(t[i1 * kLut + i2], t[(i1 + 1) * kLut + i2], o1 - i1),
                mixf(t[i1 * kLut + i2 + 1], t[(i1 + 1) * kLut + i2 + 1], o1 - i1), o2 - i2);
}
OXR_D float lut_albedo(float r, const float* t) {
    float o = sqr(r) * kLut;
    int i = lut_index(o);
    if (i >= kLut - 1) return t[kLut - 1];
    return mixf(t[i], t[i + 1], o - i);
}
OXR_D f3 multiple_scatter(const Lut& lut, float NdotL, float NdotV, float r, f3 Favg) {
    float bL = lut_bsdf(NdotL, r, lut.bsdf), bV = lut_bsdf(NdotV, r, lut.bsdf), aa = lut_albedo(r, lut.albedo);
    float fms = (1.0f - bL) * (1.0f - bV) / (kPi * (1.0f - aa));
    f3 fadd = Favg * Favg * aa / (1.0f - Favg * (1.0f - aa));
    return fms * fadd;
}

// ---------------------------------------------------------------- BSDFs
OXR_D void alphas(const oxr_material* m, float& au, float& av) {
    float r2 = sqr(m->roughness);
    au = r2 * (1.0f + m->anisotropy);
    av = r2 * (1.0f - m->anisotropy);
}
OXR_D f3 diffuse_lobe(const oxr_material* m, f3 V, f3 L, f3 N, float NdotL) {   // Material.h:16-33
    float r = m->roughness;
    f3 H = normalize(V + L);
    float NdotV = dot(N, V), VdotH = dot(V, H);
    float s = r * r, s2 = s * s;
    float VdotL = 2.0f * VdotH * VdotH - 1.0f;
    float cosri = VdotL - NdotV * NdotL;
    float C1 = 1.0f - 0.5f * s2 / (s2 + 0.33f);
    float C2 = 0.45f * s2 / (s2 + 0.09f) * cosri * (cosri >= 0.0f ? fmaxf(NdotL, NdotV) : 1.0f);
    return make3(m->albedo) * kInvPi * (C1 + C2) * (1.0f + r * 0.5f);
}
OXR_D f3 eval_diffuse(const Surf& s, f3 V, f3 L, float& pdf) {                  // Material.h:7-38
    float NdotL = to_local(L, s.n).z;
    if (NdotL <= 0.0f || dot(s.n, V) <= 0.0f) return f3(0.0f);
    f3 f = diffuse_lobe(s.m, V, L, s.n, NdotL);
    pdf = cosine_pdf(NdotL);
    return f;
}
OXR_D f3 sample_diffuse(const Surf& s, Rng& rng, f3 V, f3& Lw, float& pdf) {   // Material.h:40-72
    float u0 = rng(), u1 = rng();
    f3 lL = cosine_hemisphere(f2{u0, u1});
    Lw = to_world(lL, s.n);
    float NdotL = lL.z;
    if (NdotL <= 0.0f || dot(s.n, V) <= 0.0f) return f3(0.0f);
    f3 f = diffuse_lobe(s.m, V, Lw, s.n, NdotL);
    pdf = cosine_pdf(NdotL);
    return f;
}
OXR_D f3 conductor_value(const Surf& s, const Lut& lut, f3 V, f3 L, f3 H, float au, float av, float NdotV,
                         float NdotL) {                                          // Material.h:102-106
    const oxr_material* m = s.m;
    f3 eta = make3(m->eta), k = make3(m->k);
    f3 Favg = avg_fresnel_conductor(eta, k);
    f3 F = fresnel_conductor(V, H, eta, k);
    float G = smith_g1(V, H, s.n, au, av) * smith_g1(L, H, s.n, au, av);
    float D = ggx_d(H, s.n, au, av);
    return make3(m->albedo) * (F * D * G / (4.0f * NdotV * NdotL) + multiple_scatter(lut, NdotL, NdotV, m->roughness, Favg));
}
OXR_D f3 eval_conductor(const Surf& s, const Lut& lut, f3 V, f3 L, float& pdf) { // Material.h:76-109
    float au, av;
    alphas(s.m, au, av);
    f3 H = normalize(V + L);
    pdf = ggx_dv(V, H, s.n, au, av) * fabsf(1.0f / (4.0f * dot(V, H)));
    float NdotV = dot(s.n, V), NdotL = dot(s.n, L);
    if (NdotV <= 0.0f || NdotL <= 0.0f) return f3(0.0f);
    return conductor_value(s, lut, V, L, H, au, av, NdotV, NdotL);
}
OXR_D f3 sample_conductor(const Surf& s, const Lut& lut, Rng& rng, f3 V, f3& Lw, float& pdf) { // Material.h:111-146
    float au, av;
    alphas(s.m, au, av);
    float u0 = rng(), u1 = rng();
    f3 H = to_world(sample_vndf(s.n, V, au, av, f2{u0, u1}), s.n);
    Lw = reflect(-V, H);
    pdf = ggx_dv(V, H, s.n, au, av) * fabsf(1.0f / (4.0f * dot(V, H)));
    float NdotV = dot(s.n, V), NdotL = dot(s.n, Lw);
    if (NdotV <= 0.0f || NdotL <= 0.0f) return f3(0.0f);
    return conductor_value(s, lut, V, Lw, H, au, av, NdotV, NdotL);
}
OXR_D f3 eval_dielectric(const Surf& s, f3 V, f3 L, float& pdf) {               // Material.h:150-206
    const oxr_material* m = s.m;
    float au, av;
    alphas(m, au, av);
    float eta = m->int_ior / m->ext_ior;
    float eio = s.front ? (1.0f / eta) : eta;
    f3 N = s.n, H;
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
    f3 alb = make3(m->albedo);
    if (refl) {
        pdf = F * Dv * fabsf(1.0f / (4.0f * dot(V, H)));
        return alb * F * D * G / (4.0f * NdotV * NdotL);
    }
    float HdotV = dot(H, V), HdotL = dot(H, L);
    float sd = eio * HdotV + HdotL;
    float factor = fabsf(HdotL * HdotV / (NdotL * NdotV));
    pdf = (1.0f - F) * Dv * (fabsf(HdotL) / sqr(sd));
    f3 f = alb * (1.0f - F) * D * G * factor / sqr(sd);
    return f * sqr(1.0f / eio);
}
OXR_D f3 sample_dielectric(const Surf& s, Rng& rng, f3 V, f3& Lw, float& pdf) {  // Material.h:208-269
    const oxr_material* m = s.m;
    float au, av;
    alphas(m, au, av);
    float eta = m->int_ior / m->ext_ior;
    float eio = s.front ? (1.0f / eta) : eta;
    f3 N = s.n;
    float u0 = rng(), u1 = rng();
    f3 H = to_world(sample_vndf(N, V, au, av, f2{u0, u1}), N);
    float Dv = ggx_dv(V, H, N, au, av);
    float F = fresnel_dielectric(V, H, eio);
    float D = ggx_d(H, N, au, av);
    f3 alb = make3(m->albedo);
    if (rng() < F) {
        Lw = reflect(-V, H);
        if (dot(N, Lw) <= 0.0f) return f3(0.0f);
        pdf = F * Dv * fabsf(1.0f / (4.0f * dot(V, H)));
        float NdotV = fabsf(dot(N, V)), NdotL = fabsf(dot(N, Lw));
        float G = smith_g1(V, H, N, au, av) * smith_g1(Lw, H, N, au, av);
        return alb * F * D * G / (4.0f * NdotV * NdotL);
    }
    Lw = refract(-V, H, eio);
    if (dot(N, Lw) * dot(N, V) >= 0.0f) return f3(0.0f);
    float NdotV = fabsf(dot(N, V)), NdotL = fabsf(dot(N, Lw));
    float G = smith_g1(V, H, N, au, av) * smith_g1(Lw, H, N, au, av);
    float HdotV = dot(H, V), HdotL = dot(H, Lw);
    float sd = eio * HdotV + HdotL;
    float factor = fabsf(HdotL * HdotV / (NdotL * NdotV));
    pdf = (1.0f - F) * Dv * (fabsf(HdotL) / sqr(sd));
    f3 f = alb * (1.0f - F) * D * G * factor / sqr(sd);
    return f * sqr(1.0f / eio);
}
OXR_D f3 plastic_value(const Surf& s, const Lut& lut, f3 V, f3 L, f3 H, float au, float av, float NdotV,
                       float NdotL, float Fi, float Fo, float eta, float ps, float& pdf) {  // Material.h:299-324
    const oxr_material* m = s.m;
    f3 kd = make3(m->albedo), ks = make3(m->specular);
    float Favg = avg_fresnel_dielectric(eta);
    float Dv = ggx_dv(V, H, s.n, au, av);
    float G = smith_g1(V, H, s.n, au, av) * smith_g1(L, H, s.n, au, av);
    float D = ggx_d(H, s.n, au, av);
    f3 F(fresnel_dielectric(L, H, 1.0f / eta));
    f3 brdf = m->nonlinear ? kd / (1.0f - kd * Favg) : kd / f3(1.0f - Favg);
    brdf = brdf * ((1.0f - Fi) * (1.0f - Fo) * kInvPi);
    brdf = brdf + ks * (F * D * G / (4.0f * NdotL * NdotV) + multiple_scatter(lut, NdotL, NdotV, m->roughness, f3(Favg)));
    pdf = ps * Dv * fabsf(1.0f / (4.0f * dot(V, H))) + (1.0f - ps) * cosine_pdf(NdotL);
    return brdf;
}
OXR_D float plastic_spec_prob(const oxr_material* m, float Fi) {
    f3 kd = make3(m->albedo), ks = make3(m->specular);
    float d_sum = kd.x + kd.y + kd.z, s_sum = ks.x + ks.y + ks.z;
    float ssw = s_sum / (s_sum + d_sum);
    float ps = Fi * ssw, pd = (1.0f - Fi) * (1.0f - ssw);
    return ps / (ps + pd);
}
OXR_D f3 eval_plastic(const Surf& s, const Lut& lut, f3 V, f3 L, float& pdf) {   // Material.h:273-325
    float au, av;
    alphas(s.m, au, av);
    float eta = s.m->int_ior / s.m->ext_ior;
    f3 H = normalize(V + L);
    float NdotV = dot(s.n, V), NdotL = dot(s.n, L);
    if (NdotV <= 0.0f || NdotL <= 0.0f) return f3(0.0f);
    float Fo = fresnel_dielectric(V, s.n, 1.0f / eta), Fi = fresnel_dielectric(L, s.n, 1.0f / eta);
    float ps = plastic_spec_prob(s.m, Fi);
    return plastic_value(s, lut, V, L, H, au, av, NdotV, NdotL, Fi, Fo, eta, ps, pdf);
}
OXR_D f3 sample_plastic(const Surf& s, const Lut& lut, Rng& rng, f3 V, f3& Lw, float& pdf) { // Material.h:327-403
    float au, av;
    alphas(s.m, au, av);
    float eta = s.m->int_ior / s.m->ext_ior;
    float NdotV = dot(s.n, V);
    if (NdotV <= 0.0f) return f3(0.0f);
    float Fo = fresnel_dielectric(V, s.n, 1.0f / eta), Fi = Fo;
    float ps = plastic_spec_prob(s.m, Fi);
    f3 H;
    if (rng() < ps) {
        float u0 = rng(), u1 = rng();
        H = to_world(sample_vndf(s.n, V, au, av, f2{u0, u1}), s.n);
        Lw = reflect(-V, H);
        if (dot(s.n, Lw) <= 0.0f) return f3(0.0f);
    } else {
        float u0 = rng(), u1 = rng();
        Lw = to_world(cosine_hemisphere(f2{u0, u1}), s.n);
        H = normalize(V + Lw);
        Fi = fresnel_dielectric(Lw, s.n, 1.0f / eta);
        if (dot(s.n, Lw) <= 0.0f) return f3(0.0f);
    }
    float NdotL = dot(s.n, Lw);
    return plastic_value(s, lut, V, Lw, H, au, av, NdotV, NdotL, Fi, Fo, eta, ps, pdf);
}
OXR_D f3 metal_value(const Surf& s, f3 V, f3 L, f3 H, float au, float av, float NdotV, float NdotL, float pdiff,
                     float& pdf) {                                                 // Material.h:436-448
    const oxr_material* m = s.m;
    float Dv = ggx_dv(V, H, s.n, au, av);
    float G = smith_g1(V, H, s.n, au, av) * smith_g1(L, H, s.n, au, av);
    float D = ggx_d(H, s.n, au, av);
    f3 alb = make3(m->albedo);
    f3 F0 = f3(0.04f) * (1.0f - m->metallic) + alb * m->metallic;
    f3 F = fresnel_schlick(F0, dot(V, H));
    f3 spec = D * F * G / (4.0f * NdotL * NdotV);
    f3 diff = alb * kInvPi;
    pdf = (1.0f - pdiff) * Dv * fabsf(1.0f / (4.0f * dot(V, H))) + pdiff * cosine_pdf(NdotL);
    return pdiff * diff + (1.0f - pdiff) * spec;
}
OXR_D float metal_pdiff(const oxr_material* m) {
    float diel = 1.0f - m->metallic;
    float spec = m->metallic + diel;
    return diel / (diel + spec);
}
OXR_D f3 eval_metal(const Surf& s, f3 V, f3 L, float& pdf) {                       // Material.h:407-449
    float au, av;
    alphas(s.m, au, av);
    f3 H = normalize(L + V);
    float NdotV = dot(s.n, V), NdotL = dot(s.n, L);
    if (NdotV <= 0.0f || NdotL <= 0.0f) return f3(0.0f);
    return metal_value(s, V, L, H, au, av, NdotV, NdotL, metal_pdiff(s.m), pdf);
}
OXR_D f3 sample_metal(const Surf& s, Rng& rng, f3 V, f3& Lw, float& pdf) {        // Material.h:451-516
    float au, av;
    alphas(s.m, au, av);
    float NdotV = dot(s.n, V);
    if (NdotV <= 0.0f) return f3(0.0f);
    float pdiff = metal_pdiff(s.m);
    f3 H;
    if (rng() < pdiff) {
        float u0 = rng(), u1 = rng();
        Lw = to_world(cosine_hemisphere(f2{u0, u1}), s.n);
        H = normalize(V + Lw);
    } else {
        float u0 = rng(), u1 = rng();
        H = to_world(sample_vndf(s.n, V, au, av, f2{u0, u1}), s.n);
        Lw = reflect(-V, H);
    }
    float NdotL = dot(s.n, Lw);
    if (NdotL <= 0.0f) return f3(0.0f);
    return metal_value(s, V, Lw, H, au, av, NdotV, NdotL, pdiff, pdf);
}
OXR_D float thin_fresnel(f3 V, f3 H, float eta) {                                // Material.h:546-549
    float F = fresnel_dielectric(V, H, 1.0f / eta);
    if (F < 1.0f) F *= 2.0f / (1.0f + F);
    return F;
}
OXR_D f3 eval_thin(const Surf& s, f3 V, f3 L, float& pdf) {                       // Material.h:520-575
    const oxr_material* m = s.m;
    float eta = m->int_ior / m->ext_ior, a = sqr(m->roughness);
    f3 N = s.n, H;
    bool refl = dot(N, L) * dot(N, V) > 0.0f;
    if (refl) {
        H = normalize(V + L);
    } else {
        f3 Vr = to_world(to_local(V, -N), N);
        H = normalize(Vr + L);
        if (dot(N, H) < 0.0f) H = -H;
    }
    float Dv = ggx_dv(V, H, N, a, a);
    float F = thin_fresnel(V, H, eta);
    float D = ggx_d(H, N, a, a);
    float dwh = fabsf(1.0f / (4.0f * dot(V, H)));
    float NdotV = fabsf(dot(N, V)), NdotL = fabsf(dot(N, L));
    float G = smith_g1(V, H, N, a, a) * smith_g1(L, H, N, a, a);
    f3 alb = make3(m->albedo);
    if (refl) {
        if (dot(N, L) <= 0.0f) return f3(0.0f);
        pdf = F * Dv * dwh;
        return alb * F * D * G / (4.0f * NdotV * NdotL);
    }
    if (dot(N, L) * dot(N, V) >= 0.0f) return f3(0.0f);
    pdf = (1.0f - F) * Dv * dwh;
    return alb * (1.0f - F) * D * G / (4.0f * NdotV * NdotL);
}
OXR_D f3 sample_thin(const Surf& s, Rng& rng, f3 V, f3& Lw, float& pdf) {         // Material.h:577-630
    const oxr_material* m = s.m;
    float eta = m->int_ior / m->ext_ior, a = sqr(m->roughness);
    f3 N = s.n;
    float u0 = rng(), u1 = rng();
    f3 H = to_world(sample_vndf(N, V, a, a, f2{u0, u1}), N);
    float Dv = ggx_dv(V, H, N, a, a);
    float F = thin_fresnel(V, H, eta);
    float D = ggx_d(H, N, a, a);
    f3 alb = make3(m->albedo);
    bool refl = rng() < F;
    Lw = refl ? reflect(-V, H) : -V;
    if (refl) {
        if (dot(N, Lw) <= 0.0f) return f3(0.0f);
    } else {
        if (dot(N, Lw) * dot(N, V) >= 0.0f) return f3(0.0f);
    }
    float NdotV = fabsf(dot(N, V)), NdotL = fabsf(dot(N, Lw));
    float G = smith_g1(V, H, N, a, a) * smith_g1(Lw, H, N, a, a);
    float dwh = fabsf(1.0f / (4.0f * dot(V, H)));
    float w = refl ? F : (1.0f - F);
    pdf = w * Dv * dwh;
    return alb * w * D * G / (4.0f * NdotV * NdotL);
}
OXR_D f3 clearcoat_mix(const Surf& s, const Lut& lut, f3 V, f3 L, float au, float av, f3 cond, float cond_pdf,
                       float& pdf) {                                               // Material.h:693-705
    f3 N = s.n;
    f3 H = normalize(V + L);
    float NdotV = dot(N, V), NdotL = dot(N, L);
    float Dv = ggx_dv(V, H, N, au, av);
    float F = fresnel_dielectric(V, H, 1.0f / 1.5f);
    float cw = s.m->coat_weight * F;
    float G = smith_g1(V, H, N, au, av) * smith_g1(L, H, N, au, av);
    float D = ggx_d(H, N, au, av);
    float coat_pdf = Dv * fabsf(1.0f / (4.0f * dot(V, H)));
    f3 coat(D * G / (4.0f * NdotV * NdotL));
    pdf = cw * coat_pdf + (1.0f - cw) * cond_pdf;
    return cw * coat + (1.0f - cw) * cond;
}
OXR_D f3 eval_clearcoat(const Surf& s, const Lut& lut, f3 V, f3 L, float& pdf) {  // Material.h:634-667
    float au = sqr(s.m->coat_roughness_u), av = sqr(s.m->coat_roughness_v);
    float NdotV = dot(s.n, V), NdotL = dot(s.n, L);
    if (NdotV <= 0.0f || NdotL <= 0.0f) return f3(0.0f);
    float cond_pdf = 0.0f;
    f3 cond = eval_conductor(s, lut, V, L, cond_pdf);
    return clearcoat_mix(s, lut, V, L, au, av, cond, cond_pdf, pdf);
}
OXR_D f3 sample_clearcoat(const Surf& s, const Lut& lut, Rng& rng, f3 V, f3& Lw, float& pdf) { // Material.h:669-734
    float au = sqr(s.m->coat_roughness_u), av = sqr(s.m->coat_roughness_v);
    float F0 = fresnel_dielectric(V, s.n, 1.0f / 1.5f);
    float cw = s.m->coat_weight * F0;
    float cond_pdf = 0.0f;
    f3 cond;
    if (rng() < cw) {
        float u0 = rng(), u1 = rng();
        f3 H = to_world(sample_vndf(s.n, V, au, av, f2{u0, u1}), s.n);
        Lw = reflect(-V, H);
        float NdotV = dot(s.n, V), NdotL = dot(s.n, Lw);
        if (NdotV <= 0.0f || NdotL <= 0.0f) return f3(0.0f);
        // Material.h:693-705: the coat branch keeps the sampled H.
        float Dv = ggx_dv(V, H, s.n, au, av);
        float F = fresnel_dielectric(V, H, 1.0f / 1.5f);
        float c2 = s.m->coat_weight * F;
        float G = smith_g1(V, H, s.n, au, av) * smith_g1(Lw, H, s.n, au, av);
        float D = ggx_d(H, s.n, au, av);
        float coat_pdf = Dv * fabsf(1.0f / (4.0f * dot(V, H)));
        cond = eval_conductor(s, lut, V, Lw, cond_pdf);
        f3 coat(D * G / (4.0f * NdotV * NdotL));
        pdf = c2 * coat_pdf + (1.0f - c2) * cond_pdf;
        return c2 * coat + (1.0f - c2) * cond;
    }
    cond = sample_conductor(s, lut, rng, V, Lw, cond_pdf);
    float NdotV = dot(s.n, V), NdotL = dot(s.n, Lw);
    if (NdotV <= 0.0f || NdotL <= 0.0f) return f3(0.0f);
    return clearcoat_mix(s, lut, V, Lw, au, av, cond, cond_pdf, pdf);
}
OXR_D f3 eval_transmitter(const Surf& s, f3 V, f3 L, float& pdf) {               // Material.h:738-757
    f3 N = -s.n;
    float NdotL = to_local(L, N).z, NdotV = dot(N, V);
    if (NdotL * NdotV >= 0.0f) return f3(0.0f);
    pdf = cosine_pdf(fabsf(NdotL));
    return make3(s.m->albedo) * kInvPi;
}
OXR_D f3 sample_transmitter(const Surf& s, Rng& rng, f3 V, f3& Lw, float& pdf) { // Material.h:759-779
    float u0 = rng(), u1 = rng();
    f3 lL = cosine_hemisphere(f2{u0, u1});
    f3 N = -s.n;
    Lw = to_world(lL, N);
    float NdotL = lL.z, NdotV = dot(N, V);
    if (NdotL * NdotV >= 0.0f) return f3(0.0f);
    pdf = cosine_pdf(fabsf(NdotL));
    return make3(s.m->albedo) * kInvPi;
}
OXR_D f3 eval_material(const Surf& s, const Lut& lut, f3 V, f3 L, float& pdf) {   // Material.h:783-813
    pdf = 0.0f;
    switch (s.m->type) {
        case 0: return eval_diffuse(s, V, L, pdf);
        case 1: return eval_conductor(s, lut, V, L, pdf);
        case 2: return eval_dielectric(s, V, L, pdf);
        case 3: return eval_plastic(s, lut, V, L, pdf);
        case 4: return eval_metal(s, V, L, pdf);
        case 5: return eval_thin(s, V, L, pdf);
        case 6: return eval_clearcoat(s, lut, V, L, pdf);
        case 7: return eval_transmitter(s, V, L, pdf);
    }
    return f3(0.0f);
}
OXR_D f3 sample_material(const Surf& s, const Lut& lut, Rng& rng, f3 V, f3& L, float& pdf) { // Material.h:815-845
    pdf = 0.0f;
    switch (s.m->type) {
        case 0: return sample_diffuse(s, rng, V, L, pdf);
        case 1: return sample_conductor(s, lut, rng, V, L, pdf);
        case 2: return sample_dielectric(s, rng, V, L, pdf);
        case 3: return sample_plastic(s, lut, rng, V, L, pdf);
        case 4: return sample_metal(s, rng, V, L, pdf);
        case 5: return sample_thin(s, rng, V, L, pdf);
        case 6: return sample_clearcoat(s, lut, rng, V, L, pdf);
        case 7: return sample_transmitter(s, rng, V, L, pdf);
    }
    return f3(0.0f);
}

// The same dispatch fixed at compile time, for a shade kernel that only sees
// material type MT.  MT < 0 is the class of rays that left the scene, which
// never evaluate a BSDF: the calls compile to nothing.  MT == kMatAny keeps
// the run-time dispatch (a kernel that shades every class).
constexpr int kMatAny = 8;
template <int MT>
OXR_D f3 eval_material_t(const Surf& s, const Lut& lut, f3 V, f3 L, float& pdf) {
    pdf = 0.0f;
    if constexpr (MT == 0) return eval_diffuse(s, V, L, pdf);
    else if constexpr (MT == 1) return eval_conductor(s, lut, V, L, pdf);
    else if constexpr (MT == 2) return eval_dielectric(s, V, L, pdf);
    else if constexpr (MT == 3) return eval_plastic(s, lut, V, L, pdf);
    else if constexpr (MT == 4) return eval_metal(s, V, L, pdf);
    else if constexpr (MT == 5) return eval_thin(s, V, L, pdf);
    else if constexpr (MT == 6) return eval_clearcoat(s, lut, V, L, pdf);
    else if constexpr (MT == 7) return eval_transmitter(s, V, L, pdf);
    else if constexpr (MT == kMatAny) return eval_material(s, lut, V, L, pdf);
    else return f3(0.0f);
}
template <int MT>
OXR_D f3 sample_material_t(const Surf& s, const Lut& lut, Rng& rng, f3 V, f3& L, float& pdf) {
    pdf = 0.0f;
    if constexpr (MT == 0) return sample_diffuse(s, rng, V, L, pdf);
    else if constexpr (MT == 1) return sample_conductor(s, lut, rng, V, L, pdf);
    else if constexpr (MT == 2) return sample_dielectric(s, rng, V, L, pdf);
    else if constexpr (MT == 3) return sample_plastic(s, lut, rng, V, L, pdf);
    else if constexpr (MT == 4) return sample_metal(s, rng, V, L, pdf);
    else if constexpr (MT == 5) return sample_thin(s, rng, V, L, pdf);
    else if constexpr (MT == 6) return sample_clearcoat(s, lut, rng, V, L, pdf);
    else if constexpr (MT == 7) return sample_transmitter(s, rng, V, L, pdf);
    else if constexpr (MT == kMatAny) return sample_material(s, lut, rng, V, L, pdf);
    else return f3(0.0f);
}

// ---------------------------------------------------------------- media
OXR_D float hg_channel(float g, float c) {
    float t = 1.0f + g * g + 2.0f * g * c;
    return (0.25f * kInvPi) * (1.0f - g * g) / (t * sqrtf(t));
}
OXR_D f3 eval_phase(const oxr_medium* md, f3 V, f3 L, float& pdf) {                // Medium.h:7-43, 83-94
    pdf = 0.0f;
    if (md->phase_type == 0) {
        pdf = 0.25f * kInvPi;
        return f3(0.25f * kInvPi);
    }
    if (md->phase_type == 1) {
        float c = dot(L, V);
        f3 a(hg_channel(md->g[0], c), hg_channel(md->g[1], c), hg_channel(md->g[2], c));
        float p = ((0.0f + a.x) + a.y) + a.z;
        p *= (1.0f / 3.0f);
        pdf = p;
        if (p <= 0.0f) return f3(0.0f);
        return a;
    }
    return f3(0.0f);
}
OXR_D f3 sample_phase(const oxr_medium* md, Rng& rng, f3 V, f3& L, float& pdf) {  // Medium.h:14-22, 45-79, 96-108
    pdf = 0.0f;
    if (md->phase_type == 0) {
        float u0 = rng(), u1 = rng();
        L = uniform_sphere(f2{u0, u1});
        pdf = 0.25f * kInvPi;
        return f3(0.25f * kInvPi);
    }
    if (md->phase_type == 1) {
        int ch = min((int)(rng() * 3), 2);
        float g = md->g[ch];
        float c;
        if (fabsf(g) < kEps) {
            c = 1.0f - 2.0f * rng();
        } else {
            float st = (1.0f - g * g) / (1.0f - g + 2.0f * g * rng());
            c = (1.0f + g * g - st * st) / (2.0f * g);
        }
        f3 a(hg_channel(md->g[0], c), hg_channel(md->g[1], c), hg_channel(md->g[2], c));
        float p = ((0.0f + a.x) + a.y) + a.z;
        p *= (1.0f / 3.0f);
        pdf = p;
        if (p <= 0.0f) return f3(0.0f);
        float sn = sqrtf(fmaxf(0.0f, 1.0f - c * c));
        float phi = kTwoPi * rng();
        L = to_world(normalize(f3(sn * cosf(phi), sn * sinf(phi), c)), V);
        return a;
    }
    return f3(0.0f);
}
struct MediumCoef { f3 ss, st; float w; };
OXR_D MediumCoef medium_coef(const oxr_medium* md) {                                // Medium.h:113-125
    MediumCoef c;
    c.ss = make3(md->sigma_s) * md->scale;
    f3 sa = make3(md->sigma_a) * md->scale;
    c.st = sa + c.ss;
    f3 alb = c.ss / c.st;
    float w = 0.0f;
    if (alb.x > w && c.st.x != 0.0f) w = alb.x;
    if (alb.y > w && c.st.y != 0.0f) w = alb.y;
    if (alb.z > w && c.st.z != 0.0f) w = alb.z;
    if (w > 0.0f) w = fmaxf(w, 0.5f);
    c.w = w;
    return c;
}
OXR_D f3 eval_medium_distance(const oxr_medium* md, bool scattered, float dist, float& tpdf) { // Medium.h:112-158, 222-231
    tpdf = 0.0f;
    if (md->type != 0) return f3(0.0f);
    MediumCoef c = medium_coef(md);
    f3 att(expf(-c.st.x * dist), expf(-c.st.y * dist), expf(-c.st.z * dist));
    bool valid = att.x > 0.0f || att.y > 0.0f || att.z > 0.0f;
    float pdf;
    if (scattered) {
        pdf = ((0.0f + c.st.x * att.x) + c.st.y * att.y) + c.st.z * att.z;
        pdf *= c.w * (1.0f / 3.0f);
        att = att * c.ss;
    } else {
        pdf = ((0.0f + att.x) + att.y) + att.z;
        pdf = c.w * (1.0f / 3.0f) * pdf + (1.0f - c.w);
    }
    tpdf = pdf;
    return valid ? att : f3(0.0f);
}
OXR_D bool sample_medium_distance(const oxr_medium* md, float maxd, float& dist, float& tpdf, f3& T, Rng& rng) { // Medium.h:160-218, 233-244
    tpdf = 0.0f;
    T = f3(0.0f);
    if (md->type != 0) return false;
    MediumCoef c = medium_coef(md);
    bool scattered = false;
    float xi = rng();
    if (xi < c.w) {
        xi /= c.w;
        int ch = min((int)(rng() * 3), 2);
        dist = -logf(1.0f - xi) / c.st[ch];
        if (dist < maxd) {
            tpdf = ((0.0f + c.st.x * expf(-c.st.x * dist)) + c.st.y * expf(-c.st.y * dist)) + c.st.z * expf(-c.st.z * dist);
            tpdf *= c.w * (1.0f / 3.0f);
            scattered = true;
        }
    }
    if (!scattered) {
        dist = maxd;
        tpdf = ((0.0f + expf(-c.st.x * dist)) + expf(-c.st.y * dist)) + expf(-c.st.z * dist);
        tpdf = c.w * (1.0f / 3.0f) * tpdf + (1.0f - c.w);
    }
    T = f3(expf(-c.st.x * dist), expf(-c.st.y * dist), expf(-c.st.z * dist));
    bool valid = T.x > 0.0f || T.y > 0.0f || T.z > 0.0f;
    if (scattered) T = T * c.ss;
    if (!valid) T = f3(0.0f);
    return scattered;
}

// ---------------------------------------------------------------- lights
OXR_D float intersect_quad(f3 o, f3 d, f3 corner, f3 u_, f3 v_) {                 // Light.h:6-26
    f3 n = normalize(cross(u_, v_));
    f3 u = u_ * (1.0f / dot(u_, u_));
    f3 v = v_ * (1.0f / dot(v_, v_));
    float dt = dot(d, n);
    float t = (dot(n, corner) - dot(n, o)) / dt;
    if (t > kEps) {
        f3 vi = (o + d * t) - corner;
        float a1 = dot(u, vi);
        if (a1 >= 0.0f && a1 <= 1.0f) {
            float a2 = dot(v, vi);
            if (a2 >= 0.0f && a2 <= 1.0f) return t;
        }
    }
    return kFltMax;
}
OXR_D float intersect_sphere(f3 o, f3 d, f3 c, float r) {                         // Light.h:78-98
    f3 co = c - o;
    float b = dot(co, d);
    float det = b * b - dot(co, co) + r * r;
    if (det < 0.0f) return kFltMax;
    det = sqrtf(det);
    float t1 = b - det;
    if (t1 > kEps) return t1;
    float t2 = b + det;
    if (t2 > kEps) return t2;
    return kFltMax;
}
OXR_D f3 eval_light(const oxr_light* l, f3 o, f3 d, float dist, float& pdf, float& ldist) { // Light.h:28-52, 100-116, 145-156
    if (l->type == 0) {
        f3 lu = make3(l->u), lv = make3(l->v);
        f3 n = normalize(cross(lu, lv));
        float c = dot(d, n);
        if (!l->double_side && c < 0.0f) { pdf = 0.0f; return f3(0.0f); }
        ldist = intersect_quad(o, d, make3(l->position), lu, lv);
        float area = length(lu) * length(lv);
        if (ldist < dist) {
            pdf = 1.0f / area;
            pdf *= sqr(ldist) / fabsf(c);
            return make3(l->radiance);
        }
        pdf = 0.0f;
        return f3(0.0f);
    }
    if (l->type == 1) {
        f3 lp = make3(l->position);
        ldist = intersect_sphere(o, d, lp, l->radius);
        if (ldist < dist) {
            f3 dir = lp - o;
            float d2 = dot(dir, dir);
            float s2 = l->radius * l->radius / d2;
            pdf = cone_pdf(sqrtf(1.0f - s2));
            return make3(l->radiance);
        }
        pdf = 0.0f;
        return f3(0.0f);
    }
    pdf = 0.0f;
    return f3(0.0f);
}
OXR_D f3 sample_light(const oxr_light* l, f3 p, f2 s, f3& L, float& dist, float& pdf) { // Light.h:54-74, 118-141, 158-169
    if (l->type == 0) {
        f3 lu = make3(l->u), lv = make3(l->v);
        f3 n = normalize(cross(lu, lv));
        f3 pos = make3(l->position) + lu * s.x + lv * s.y;
        L = pos - p;
        float d2 = dot(L, L);
        dist = sqrtf(d2);
        L = L / dist;
        float c = dot(L, n);
        if (!l->double_side && c < 0.0f) { pdf = 0.0f; return f3(0.0f); }
        float area = length(lu) * length(lv);
        pdf = 1.0f / area;
        pdf *= d2 / fabsf(c);
        return make3(l->radiance);
    }
    if (l->type == 1) {
        f3 dir = make3(l->position) - p;
        float d2 = dot(dir, dir);
        float inv = 1.0f / sqrtf(d2);
        dir = dir * inv;
        dist = d2 * inv;
        float st = l->radius * inv;
        if (st < 1.0f) {
            float ct = sqrtf(1.0f - st * st);
            f3 lL = sample_cone(s, ct);
            float ci = lL.z;
            L = to_world(lL, dir);
            pdf = cone_pdf(ct);
            dist = ci * dist - sqrtf(fmaxf(0.0f, l->radius * l->radius - (1.0f - ci * ci) * d2));
            return make3(l->radiance);
        }
        pdf = 0.0f;
        return f3(0.0f);
    }
    pdf = 0.0f;
    return f3(0.0f);
}

}  // namespace oxr
