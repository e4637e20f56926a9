/*
 * CPU oracle for the path-tracing hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product (optixrender_amd/)
 * never links, includes or calls it.
 *
 * It is a scalar C++ restatement of the reference renderer's device code
 * (GraphicsEnthusiast/OptiXRender):
 *   core/cuda/devicePrograms.cu:170-532   raygen / bounce loop / MIS / RR
 *   core/cuda/devicePrograms.cu:71-165    closest-hit decode, miss programs
 *   core/cuda/Material.h                  8 BSDFs
 *   core/cuda/Microfacet.h                Fresnel, GGX, Kulla-Conty LUT
 *   core/cuda/Medium.h                    homogeneous medium, phase functions
 *   core/cuda/Light.h, Environment.h, Sampling.h, Filter.h, Utils.h
 *   core/cuda/PostProcessing.cu:3-48      tone map
 *   core/src/Scene.cpp:217-313            environment importance cache
 *   core/src/Renderer.cpp:796-813         pinhole camera
 *   common/gdt/gdt/random/random.h:55-89  LCG<16> (TEA seeding)
 * plus the build's own definitions for what OptiX / CUDA textures did in the
 * reference (LBVH over Morton codes after Karras 2012, watertight triangle
 * test after Woop/Benthin/Wald 2013, FP32 software bilinear with wrap).
 *
 * Parity status of each function is listed in oracle.cpp's header comment.
 * All structs below are plain 4-byte fields; their layouts equal the public
 * C-ABI structs in include/oxr.h (a CPU test checks sizes and offsets).
 */
#ifndef OXR_ORACLE_H
#define OXR_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_material {
    int32_t type;
    float albedo[3];
    float roughness;
    float anisotropy;
    float eta[3];
    float k[3];
    float int_ior, ext_ior;
    int32_t nonlinear;
    float specular[3];
    float metallic;
    float coat_weight;
    float coat_roughness_u, coat_roughness_v;
    int32_t in_medium, out_medium;
    int32_t tex[8];
} orc_material;

typedef struct orc_light {
    int32_t type;        /* 0 quad, 1 sphere */
    int32_t double_side;
    float radiance[3];
    float position[3];
    float u[3];
    float v[3];
    float radius;
    int32_t medium;
} orc_light;

typedef struct orc_medium {
    int32_t type;        /* 0 homogeneous */
    int32_t phase_type;  /* 0 isotropic, 1 Henyey-Greenstein */
    float g[3];
    float sigma_s[3];
    float sigma_a[3];
    float scale;
} orc_medium;

typedef struct orc_camera {
    float from[3], at[3], up[3];
    int32_t medium;
    float cos_fovy;
} orc_camera;

typedef struct orc_texture {
    int32_t width, height;
    const uint8_t* rgba8;       /* width*height*4 bytes, row 0 = v~0 */
} orc_texture;

typedef struct orc_scene {
    int64_t n_tris;
    const float* positions;     /* 9 floats per triangle (de-indexed) */
    const float* normals;       /* 9 floats per triangle */
    const uint32_t* mat_ids;    /* per triangle, index into materials */
    int32_t n_materials;
    const orc_material* materials;
    int32_t n_lights;
    const orc_light* lights;
    int32_t n_media;
    const orc_medium* media;
    int32_t env_w, env_h;       /* 0 = no environment */
    const float* env_rgba;      /* env_w*env_h*4 floats, row 0 = v~0 */
    const float* uvs;           /* 6 floats per triangle, or NULL: (0,0) (1,0) (0,1) */
    int32_t n_textures;
    const orc_texture* textures;
} orc_scene;

typedef struct orc_render_params {
    int32_t width, height;
    int32_t spp;            /* frames of one sample each */
    int32_t max_depth;      /* maxBounce */
    int32_t rr_start;       /* first bounce that plays Russian roulette */
    int32_t filter;         /* 0 box 1 tent 2 triangle 3 gaussian */
    int32_t frame_base;     /* frameID of the first frame */
    int32_t threads;        /* host threads, 0 = all */
    int32_t fp64_accum;     /* 1: accumulate the mean in double */
    int32_t row_step;       /* orc_render_stats: trace rows iy % row_step == 0 only (0 or 1: all) */
} orc_render_params;

/* ---- RNG (random.h:55-89) ---- */
uint32_t orc_lcg_init(uint32_t v0, uint32_t v1);
float orc_lcg_next(uint32_t* state);

/* ---- BVH ---- */
void orc_morton_codes(int64_t n, const float* positions, uint32_t* codes, float* cbounds6);
/* rounds: treelet restructuring rounds after the LBVH (0: the plain LBVH,
   < 0: the build's default, as oxr_scene_build_bvh does);
   codes: Morton codes in the radix sort's order;
   order: triangle id of each leaf position; left/right: child refs of internal nodes
   (n-1 each; bit 31 set = leaf, low bits = sorted leaf position);
   lo/hi: 3 floats per internal node then per leaf (2n-1 boxes);
   qnode (optional): the traversal's compressed record of each internal node,
   8 words: origin.xyz, gamma | left leaf << 30 | right leaf << 31, three
   signed scale exponents, twelve 8-bit child bounds. */
int orc_lbvh_build(int64_t n, const float* positions, int32_t rounds, uint32_t* codes, uint32_t* order,
                   uint32_t* left, uint32_t* right, uint32_t* parent, float* lo, float* hi, uint32_t* qnode);

/* Ray queries against an LBVH built internally over scene positions.
   rays: 8 floats each (ox oy oz tmin dx dy dz tmax).
   closest hits: 4 words each (t, prim id bits, u, v); prim = 0xffffffff on miss.
   counts (optional): 2 uint32 per ray (node records fetched, triangle tests). */
void* orc_bvh_create(int64_t n, const float* positions);
void orc_bvh_destroy(void* bvh);
void orc_bvh_closest(const void* bvh, int64_t n_rays, const float* rays, float* hits, uint32_t* counts);
void orc_bvh_any(const void* bvh, int64_t n_rays, const float* rays, uint8_t* occluded, uint32_t* counts);
void orc_brute_closest(int64_t n, const float* positions, int64_t n_rays, const float* rays, float* hits);

/* ---- shading building blocks ---- */
void orc_kulla_conty(float* bsdf_avg /*64*64*/, float* albedo_avg /*64*/);
/* Microfacet.h helpers in the local frame N = +z: SampleGGX (half vector),
   FresnelConductor (V and H unit, eta/k per channel), FresnelDielectric,
   AverageFresnelConductor, AverageFresnelDielectric, DistributionGGX,
   GeometrySmith_1. */
void orc_sample_ggx(float alpha_u, float alpha_v, float u0, float u1, float* h3);
void orc_fresnel_conductor(const float* V, const float* H, const float* eta, const float* k, float* out3);
float orc_fresnel_dielectric(const float* V, const float* H, float eta_inv);
void orc_average_fresnel_conductor(const float* eta, const float* k, float* out3);
float orc_average_fresnel_dielectric(float eta);
float orc_ggx_d(const float* H, float alpha_u, float alpha_v);
float orc_smith_g1(const float* V, const float* H, float alpha_u, float alpha_v);
/* Evaluate / sample a material at a shading frame.  n = shading normal
   (already flipped to the incident side), front_face as in SetFaceNormal.
   out: f[3], pdf, (sample) L[3]; rng_state advanced by the sample call. */
void orc_material_eval(const orc_material* m, const float* n, int front_face, const float* V,
                       const float* L, float* f, float* pdf);
void orc_material_sample(const orc_material* m, const float* n, int front_face, const float* V,
                         uint32_t* rng_state, float* f, float* pdf, float* L);
void orc_medium_sample_distance(const orc_medium* m, float max_distance, uint32_t* rng_state,
                                float* out /* distance, trans_pdf, T[3], scattered */);
void orc_medium_eval_distance(const orc_medium* m, int scattered, float distance,
                              float* out /* trans_pdf, T[3] */);
void orc_phase_eval(const orc_medium* m, const float* V, const float* L, float* out /* f[3], pdf */);
void orc_phase_sample(const orc_medium* m, const float* V, uint32_t* rng_state,
                      float* out /* f[3], pdf, L[3] */);
void orc_filter_jitter(int type, float u0, float u1, float* out2);
void orc_env_cache(int w, int h, const float* rgba, float* cache_rgba);
void orc_bilinear(int w, int h, const float* rgba, float u, float v, float* out4);
/* material texture lookup: RGBA8 texels read as byte/255, then as orc_bilinear */
void orc_texture_sample(int w, int h, const uint8_t* rgba8, float u, float v, float* out4);
/* NormalFromTangentToWorld: n (unit), t = texel RGB in [0,1] */
void orc_normal_from_tangent(const float* n, const float* t, float* out3);
void orc_tonemap(int64_t n, const float* rgba, uint32_t* out);

/* ---- whole render ---- */
/* color/albedo/normal: width*height float4 running means (in/out when
   frame_base > 0).  path_radiance (optional): spp*width*height*3 floats. */
int orc_render(const orc_scene* scene, const orc_camera* cam, const orc_render_params* p,
               float* color, float* albedo, float* normal, float* depth, float* path_radiance);
/* counts of extension/shadow rays and traversal work, for measurement:
   out[0] extension rays, out[1] shadow rays, out[2] node records fetched,
   out[3] triangle tests, out[4] / out[5] the same two for extension rays,
   out[6] nanoseconds spent tracing paths (after the scene's LBVH build and
   environment cache), out[7] paths traced. */
int orc_render_stats(const orc_scene* scene, const orc_camera* cam, const orc_render_params* p,
                     uint64_t* out8);

#ifdef __cplusplus
}
#endif
#endif
