/*
 * optixrender_amd -- C ABI of the MI355X path-tracing hot path.
 *
 * This is the drop-in boundary that replaces the reference renderer's OptiX
 * path (GraphicsEnthusiast/OptiXRender).  Plain C types only: opaque context,
 * plain pointers and sizes, int status codes (0 = OK, message through
 * oxr_last_error).  The C++ facade in optixrender_amd/cpp/ (Renderer, Scene,
 * SceneParser, CUDABuffer) is built on these entry points and keeps the
 * reference's class interfaces.
 *
 * Threading: one context is driven by one host thread at a time.
 */
#ifndef OXR_H
#define OXR_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OXR_OK 0
#define OXR_ERR_INVALID 1
#define OXR_ERR_HIP 2
#define OXR_ERR_OOM 3
#define OXR_ERR_STATE 4

typedef struct oxr_ctx oxr_ctx;

/* Material, one per surface slot (reference: Material, core/include/Scene.h:70-116,
   plus the mesh's in/out medium from TriangleMeshSBTData, LaunchParams.h:16-26).
   type: 0 Diffuse 1 Conductor 2 Dielectric 3 Plastic 4 MetalWorkflow
         5 ThinDielectric 6 ClearCoatedConductor 7 DiffuseTransmitter (Scene.h:59-68).
   tex[8]: texture ids (albedo, roughness, anisotropy, specular, metallic,
   coat_roughness_u, coat_roughness_v, normal); -1 = none. */
typedef struct oxr_material {
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
} oxr_material;

/* Analytic light (reference: Light, Scene.h:143-155). type 0 quad, 1 sphere. */
typedef struct oxr_light {
    int32_t type;
    int32_t double_side;
    float radiance[3];
    float position[3];
    float u[3];
    float v[3];
    float radius;
    int32_t medium;
} oxr_light;

/* Participating medium (reference: Medium/Phase, Scene.h:33-55). */
typedef struct oxr_medium {
    int32_t type;        /* 0 homogeneous */
    int32_t phase_type;  /* 0 isotropic, 1 Henyey-Greenstein */
    float g[3];
    float sigma_s[3];
    float sigma_a[3];
    float scale;
} oxr_medium;

/* Camera (reference: Camera, Scene.h:11-20; cosFovy fixed at 0.66 in
   Renderer::SetCamera, Renderer.cpp:803 -- here a field, 0 means 0.66). */
typedef struct oxr_camera {
    float from[3], at[3], up[3];
    int32_t medium;
    float cos_fovy;
} oxr_camera;

/* RGBA8 material texture (reference: Texture, Scene.cpp:181-195: stbi RGBA
   load flipped vertically, so row 0 is at v ~ 0).  Sampled as the reference's
   tex2D<float4> with linear filtering, wrap addressing and normalised-float
   reads (Renderer.cpp:95-144), in FP32 software bilinear. */
typedef struct oxr_texture {
    int32_t width, height;
    const uint8_t* rgba8;  /* width*height*4 bytes, row-major */
} oxr_texture;

/* Scene upload (replaces Scene + BuildAccel's per-mesh buffers,
   Renderer.cpp:146-285).  Triangles are de-indexed as TriangleMesh::LoadMesh
   leaves them (Scene.cpp:167-174): 9 floats of positions and 9 of normals per
   triangle, one material slot per triangle.  uvs (6 floats per triangle) may
   be NULL: the texture coordinates are then (0,0), (1,0), (0,1) per triangle
   (Scene.cpp:160-165).  env_rgba: env_w*env_h RGBA32F, row 0 at v ~ 0 (stbi
   flipped load, Scene.cpp:197-206).  textures: the images a material's tex[]
   ids index.  Host pointers are only read during the call. */
typedef struct oxr_scene_desc {
    int64_t n_tris;
    const float* positions;
    const float* normals;
    const float* uvs;
    const uint32_t* mat_ids;
    int32_t n_materials;
    const oxr_material* materials;
    int32_t n_lights;
    const oxr_light* lights;
    int32_t n_media;
    const oxr_medium* media;
    int32_t env_w, env_h;
    const float* env_rgba;
    int32_t n_textures;
    const oxr_texture* textures;
} oxr_scene_desc;

#define OXR_FLAG_KEEP_PATH_RADIANCE 1u  /* keep per-(frame, pixel) radiance for oxr_download_path_radiance */
#define OXR_FLAG_COUNT_TRAVERSAL 2u     /* count node fetches / triangle tests (slower; measurement only) */
#define OXR_FLAG_EAGER 4u               /* launch kernels one by one instead of replaying a captured hipGraph */
#define OXR_FLAG_STAGE_TIMING 8u        /* per-stage device times (HIP events between stages; implies eager) */
#define OXR_FLAG_NO_CAMERA_CULL 16u     /* trace every camera ray (results are identical either way) */
#define OXR_FLAG_PRESENT 32u            /* copy the finished RGBA8 frame to page-locked host memory on a copy
                                           stream, overlapped with what follows (oxr_present) */

/* One render call = spp progressive frames of one sample each (reference
   Render() with numPixelSamples = 1, Renderer.cpp:635-793). */
typedef struct oxr_render_params {
    int32_t spp;          /* frames rendered by this call (>= 1) */
    int32_t max_depth;    /* maxBounce (LaunchParams.h:30); 0 means 25 */
    int32_t rr_start;     /* first bounce that plays Russian roulette (reference: 0) */
    int32_t progressive;  /* 1: continue frameID; 0: frameID restarts at 0 (Renderer.cpp:642-644) */
    int32_t filter;       /* 0 box 1 tent 2 triangle 3 gaussian (reference default 3) */
    int32_t denoise;      /* 0 plain accumulation; 1 denoise the accumulated image before the tone map */
    uint32_t flags;       /* OXR_FLAG_* */
    int32_t batch_frames; /* frames per sample window, 0 = automatic (oxr_tuning.window_samples) */
} oxr_render_params;

typedef struct oxr_render_stats {
    double render_ms;        /* whole oxr_render wall time on the device stream */
    double trace_ms;         /* ray traversal: each iteration traces its extension rays and the previous
                                bounce's shadow rays in one launch */
    double shade_ms;
    double generate_ms;      /* window start and path regeneration */
    double accumulate_ms, tonemap_ms;
    double denoise_ms;       /* KPCNN denoiser (denoise = 1 only) */
    uint64_t extension_rays, shadow_rays;
    uint64_t node_fetches, tri_tests;   /* all traversals; only with OXR_FLAG_COUNT_TRAVERSAL */
    uint64_t ext_node_fetches, ext_tri_tests;  /* extension rays only; same condition */
    uint64_t paths;
    int32_t frame_id;                   /* next frameID */
    int32_t kernel_launches;
    int32_t iterations;                 /* wavefront iterations (trace .. shade) over all windows */
    int32_t windows;                    /* sample windows rendered */
    /* Camera rays known to leave the scene without a trace: their pixel's
       tile lies outside the image of every box of the BVH's top levels (the
       cover set, a few hundred boxes).  extension_rays counts the rays traced.
       With OXR_FLAG_COUNT_TRAVERSAL every ray is traced, and these are the
       rays that would have been skipped and the node fetches / triangle tests
       they took (included in ext_node_fetches / ext_tri_tests). */
    uint64_t camera_rays_resolved;
    uint64_t resolved_node_fetches, resolved_tri_tests;
    /* Rays traced by the tail kernel (oxr_tuning.tail_paths), included in
       extension_rays / shadow_rays, and (when counting) their node fetches /
       triangle tests, included in node_fetches / tri_tests. */
    uint64_t tail_extension_rays, tail_shadow_rays;
    uint64_t tail_node_fetches, tail_tri_tests;
    /* With OXR_FLAG_COUNT_TRAVERSAL, the SIMD use of the trace launches'
       traversal loop: internal-node visits made by the lanes, and 64 x the
       visits of the busiest lane of each 64-ray chunk (the loop iterations
       the waves ran).  Their ratio is the fraction of lanes doing work. */
    uint64_t trace_lane_visits, trace_wave_slots;
} oxr_render_stats;

/* ---- context and memory (CUDABuffer-style API, common/CUDABuffer.h:26-81) ---- */
int oxr_create(int device, oxr_ctx** out);
int oxr_destroy(oxr_ctx* ctx);
const char* oxr_last_error(const oxr_ctx* ctx);
int oxr_malloc(oxr_ctx* ctx, size_t bytes, void** dptr);
int oxr_free(oxr_ctx* ctx, void* dptr);
int oxr_memcpy_h2d(oxr_ctx* ctx, void* dst, const void* src, size_t bytes);
int oxr_memcpy_d2h(oxr_ctx* ctx, void* dst, const void* src, size_t bytes);
int oxr_memcpy_d2d(oxr_ctx* ctx, void* dst, const void* src, size_t bytes);
int oxr_synchronize(oxr_ctx* ctx);

/* ---- scenes (Scene / BuildAccel / CreateTextures / BuildSBT) ---- */
int oxr_scene_upload(oxr_ctx* ctx, const oxr_scene_desc* desc, uint32_t* scene_id);
/* BVH build on the device (replaces optixAccelBuild/Compact, Renderer.cpp:247-274):
   an LBVH (Morton codes, radix sort, Karras hierarchy, refit), one round of
   treelet restructuring (Karras & Aila 2013) relabelled as Karras's layout of
   its own leaf order, and the traversal's compressed node records. */
int oxr_scene_build_bvh(oxr_ctx* ctx, uint32_t scene_id, float* build_ms);
/* The same with `rounds` treelet rounds (0: the plain LBVH; -1: the default
   of oxr_scene_build_bvh; at most 4).  No reference counterpart: OptiX's
   build quality is not exposed. */
int oxr_scene_build_bvh_rounds(oxr_ctx* ctx, uint32_t scene_id, int32_t rounds, float* build_ms);
int oxr_scene_release(oxr_ctx* ctx, uint32_t scene_id);
int oxr_scene_select(oxr_ctx* ctx, uint32_t scene_id);
/* BVH of a built scene in Karras layout: codes/order: n; left/right: n-1
   (bit 31 = leaf, low bits = sorted position); lo/hi: 3*(2n-1) floats
   (internal nodes then leaves). Any pointer may be NULL. */
int oxr_scene_download_bvh(oxr_ctx* ctx, uint32_t scene_id, uint32_t* codes, uint32_t* order,
                           uint32_t* left, uint32_t* right, float* lo, float* hi);
/* The traversal's compressed node records of a built scene: 8 words per
   internal node (n-1): origin.xyz (floats), gamma | left leaf << 30 | right
   leaf << 31, three signed per-axis scale exponents, twelve 8-bit child
   bounds (child 0 lo.xyz hi.xyz, child 1 lo.xyz hi.xyz).  Replaces the
   acceleration structure OptiX builds and keeps opaque
   (optixAccelBuild/Compact, Renderer.cpp:247-274). */
int oxr_scene_download_nodes(oxr_ctx* ctx, uint32_t scene_id, uint32_t* out);
/* Environment importance cache built on the device (HdrTexture::CalculateHdrCache,
   Scene.cpp:217-313): env_w*env_h RGBA32F. */
int oxr_scene_download_env_cache(oxr_ctx* ctx, uint32_t scene_id, float* cache_rgba);

/* ---- ray queries against the selected scene (optixTrace, devicePrograms.cu:258-269,303-320) ----
   rays: n*8 floats (ox oy oz tmin dx dy dz tmax), host memory.
   closest: hits n*4 words (t, triangle id bits, u, v); id 0xffffffff = miss.
   any: occluded n bytes.  counts (optional): n*2 uint32 (node records, triangle tests). */
int oxr_trace_rays(oxr_ctx* ctx, const float* rays, int64_t n, int any_hit, float* hits,
                   uint8_t* occluded, uint32_t* counts);

/* Ray queries that allocate nothing.  No counterpart in the reference: for
   callers that cannot allocate device memory (code under HIP graph capture,
   tight serving loops, harnesses that forbid hipMalloc).  As oxr_trace_rays,
   but rays (n*8 floats), hits (n*4 words) and occluded (n bytes) are DEVICE
   pointers the caller owns, and no visit counts are returned.  The traversal
   stack spill and the overflow flag live in a fixed device array of 32 MiB
   per GPU, held until the process ends; n is at most 2^31 (OXR_ERR_INVALID
   beyond).  Every call on one GPU shares that array: calls that overlap on
   one GPU (two contexts or threads at once) race.  Blocks until the rays are
   traced, like oxr_trace_rays. */
int oxr_trace_rays_noalloc(oxr_ctx* ctx, const float* rays, int64_t n, int any_hit, float* hits,
                           uint8_t* occluded);

/* ---- renderer (Renderer::Resize/SetCamera/Render/DownloadPixels, Renderer.h:13-25) ---- */
int oxr_resize(oxr_ctx* ctx, int width, int height);
int oxr_set_camera(oxr_ctx* ctx, const oxr_camera* cam);
int oxr_render(oxr_ctx* ctx, const oxr_render_params* params);
int oxr_get_stats(oxr_ctx* ctx, oxr_render_stats* out);
int oxr_download_rgba8(oxr_ctx* ctx, uint32_t* out, size_t count);
/* The viewer's readback, streamed and double-buffered as the reference's
   comment asks for (Renderer.cpp:787-792, in place of DownloadPixels'
   synchronous copy, Renderer.cpp:878-880): the RGBA8 frame of the newest
   render made with OXR_FLAG_PRESENT, in page-locked host memory owned by the
   context (width*height pixels, row 0 at the bottom as oxr_download_rgba8).
   Waits only for that frame's copy.  The pointer stays valid through the next
   render with the flag and is overwritten by the one after it; oxr_resize and
   oxr_destroy free it.  frame_id: that render's stats.frame_id.
   OXR_ERR_STATE when no render with the flag ran since the last resize. */
int oxr_present(oxr_ctx* ctx, const uint32_t** pixels, int32_t* width, int32_t* height, int32_t* frame_id);
/* which: 0 colour, 1 albedo, 2 normal, 3 denoised colour (after a render
   with denoise = 1); count = width*height float4. */
int oxr_download_aov(oxr_ctx* ctx, int which, float* out_rgba, size_t count);
/* First-hit depth AOV, running mean of d / (1 + d) (1 for a miss): a
   denoiser input with no reference counterpart.  count = width*height. */
int oxr_download_depth(oxr_ctx* ctx, float* out, size_t count);
/* radiance of every path of the last render call: spp*width*height*3 floats,
   frame-major.  Needs OXR_FLAG_KEEP_PATH_RADIANCE. */
int oxr_download_path_radiance(oxr_ctx* ctx, float* out, size_t count);

/* ---- render server: continuous batching over resident scenes ----
   No counterpart in the reference (its viewer renders one frame per Render()
   call, main.cpp:66-128).  Runner threads, each with its own path pool and
   stream, admit requests FIFO while a slot is free and their buffers fit in
   the HBM budget; each runs the active requests of one (scene, max_depth,
   rr_start, camera medium) group that no other runner holds through a shared
   wavefront, handing out frames round-robin up to tick_paths paths per tick,
   so the ticks of different groups (other scenes, or more requests of a busy
   one) overlap on the device.  A request's image equals a single oxr_render
   of the same request (non-progressive, frames 0..spp-1). */
typedef struct oxr_server oxr_server;
typedef struct oxr_server_config {
    size_t hbm_budget_bytes;   /* wavefront buffers + all admitted requests' buffers */
    int32_t max_active;        /* admitted requests at a time (<= 256) */
    int32_t runners;           /* ticks in flight at once, each with its own path pool (0..8; 0 = 2) */
    int64_t tick_paths;        /* paths per shared wavefront tick */
    int64_t max_frame_pixels;  /* largest request frame accepted (width*height) */
} oxr_server_config;
typedef struct oxr_request {
    uint32_t scene_id;
    int32_t width, height, spp, max_depth, rr_start, filter;
    int32_t denoise;           /* 1: denoise the finished image with the context's weights (oxr_denoiser_load) */
    oxr_camera camera;
} oxr_request;
/* request states reported by oxr_poll */
#define OXR_REQ_PENDING 0
#define OXR_REQ_ACTIVE 1
#define OXR_REQ_DONE 2
#define OXR_REQ_CANCELLED 3
#define OXR_REQ_FAILED 4
typedef struct oxr_server_stats {
    uint64_t submitted, completed, cancelled, rejected, ticks, paths;
    double busy_ms, wall_ms;
    double p50_ms, p99_ms, mean_ms;  /* submit -> done latency of completed requests */
    int32_t active, pending;
    size_t bytes_in_use, budget_bytes;
    uint64_t overlapped_ticks;  /* ticks that began while another runner's tick was running */
    int32_t runners, max_running;  /* runner threads; most ticks in flight at once */
} oxr_server_stats;
int oxr_server_create(oxr_ctx* ctx, const oxr_server_config* cfg, oxr_server** out);
int oxr_server_destroy(oxr_server* s);
const char* oxr_server_last_error(const oxr_server* s);
int oxr_submit(oxr_server* s, const oxr_request* r, uint64_t* id);
int oxr_poll(oxr_server* s, uint64_t id, int* state);
/* blocks until done (0), failed/cancelled or timed out (non-zero); timeout_ms < 0 waits forever */
int oxr_wait(oxr_server* s, uint64_t id, int timeout_ms);
int oxr_cancel(oxr_server* s, uint64_t id);
int oxr_result_rgba8(oxr_server* s, uint64_t id, uint32_t* out, size_t count);
int oxr_result_color(oxr_server* s, uint64_t id, float* out_rgba, size_t count);
/* frees a finished request's buffers and forgets it */
int oxr_release_request(oxr_server* s, uint64_t id);
int oxr_server_get_stats(oxr_server* s, oxr_server_stats* out);
/* A finished request's submit -> admission time (waiting for a slot and HBM
   budget) and submit -> done latency; OXR_ERR_STATE until it is done. */
int oxr_request_times(oxr_server* s, uint64_t id, double* queued_ms, double* total_ms);

/* ---- launch tuning (no reference counterpart: OptiX chose these internally) ---- */
typedef struct oxr_tuning {
    int32_t trace_blocks_per_cu;  /* persistent grid of extend/connect per CU (1..16), capped at what fits at once */
    int32_t shade_blocks_per_cu;  /* persistent grid of shade/regenerate per CU (1..16), capped likewise */
    int32_t stack_lds;            /* traversal stack entries kept in LDS: 8, 16 or 32 */
    int32_t iters_per_poll;       /* wavefront iterations per graph replay between completion polls (even, 2..64) */
    int64_t pool_paths;           /* resident path slots (rounded up to a multiple of 2048) */
    int64_t window_samples;       /* samples whose radiance and AOVs are kept at once (48 B each) */
    int64_t tail_paths;           /* an iteration with at most this many paths carries each to its end in one
                                     launch (tail mode); 0 = never */
} oxr_tuning;
int oxr_get_tuning(oxr_ctx* ctx, oxr_tuning* out);
int oxr_set_tuning(oxr_ctx* ctx, const oxr_tuning* t);
/* Coordinate-descent search over oxr_tuning: renders `probe` on the selected
   scene for each candidate value of one knob at a time (LDS stack entries,
   trace and shade blocks per CU, pool slots, tail_paths, iterations per
   poll), each timed as the best of three renders after a warm-up, moves to
   a candidate only when it is at least 1% faster, applies the result and
   returns it with its render time.  Resets frameID. */
int oxr_autotune(oxr_ctx* ctx, const oxr_render_params* probe, oxr_tuning* best, double* best_ms);

/* ---- building blocks exposed for parity tests ---- */
/* ComputeFinalPixelColorsKernel (PostProcessing.cu:26-48) over n host float4. */
int oxr_tonemap_host(oxr_ctx* ctx, const float* rgba, int64_t n, uint32_t* out);

/* ---- denoiser (optixDenoiserCreate/Setup/Invoke, Renderer.cpp:660-785,
   817-874) ----  The OptiX AI denoiser is closed; this is a kernel-predicting
   CNN of our own: 5 layers of 3x3 convolutions 10 -> 32 -> 32 -> 32 -> 32 ->
   25 over log(1 + colour), albedo, normal and depth, whose softmaxed outputs
   form a 5x5 filter applied to the colour.  BF16 MFMA with FP32 accumulation.
   weights: per layer l, w[co][ci][3][3] then b[co] as floats (37,881 in all). */
#define OXR_DENOISER_PARAMS 37881
int oxr_denoiser_load(oxr_ctx* ctx, const float* weights, int64_t count);
/* Denoises host AOVs of one w x h image (float4 colour/albedo/normal, float
   depth) with the loaded weights into out (float4). */
int oxr_denoise_host(oxr_ctx* ctx, int w, int h, const float* color, const float* albedo, const float* normal,
                     const float* depth, float* out);
/* Kulla-Conty tables as uploaded (ComputeKullaConty, Microfacet.h:231-291). */
int oxr_kulla_conty(float* bsdf_avg /* 64*64 */, float* albedo_avg /* 64 */);
/* Batched EvaluateMaterial / SampleMaterial (Material.h:783-845) on the device.
   in: n*12 floats (N xyz, V xyz, L xyz, front_face, rng_state bits, unused).
   out: n*8 floats (f xyz, pdf, L xyz (sample), rng_state bits). mode 0 eval, 1 sample. */
int oxr_material_batch(oxr_ctx* ctx, const oxr_material* m, int mode, const float* in, int64_t n, float* out);

#ifdef __cplusplus
}
#endif
#endif
