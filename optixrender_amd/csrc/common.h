// Shared host/device declarations of the HIP library: device scene layout,
// launch helpers and the kernel launchers called by the C ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/oxr.h"

namespace oxr {

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kNoHit = 0xffffffffu;
constexpr int kStackTotal = 64;   // traversal stack entries per lane (LBVH depth <= 30 + log2 n < 64)
constexpr int kStackLdsMin = 8;   // fewest entries an instance keeps in LDS; the rest spill to HBM
constexpr int kTraceBlock = 256;

// BVH in traversal layout.  qnodes: one 32-byte record per internal node i
// at qnodes[2i], qnodes[2i + 1] (words 0..7):
//   0..2  origin: the node's lower corner (floats)
//   3     gamma | left leaf << 30 | right leaf << 31: the children are gamma
//         and gamma + 1 (Karras split), each an internal id or, flagged, a
//         sorted leaf
//   4     e_x, e_y, e_z as signed bytes: per-axis scale 2^e
//   5..7  12 bytes: child 0 lo.xyz, hi.xyz, child 1 lo.xyz, hi.xyz, each an
//         8-bit multiple of the scale above the origin, rounded outwards
// A child's slab bound at q steps lies at ray distance
// fma(q, 2^e * inv, (origin - o) * inv).
// tris: 3 float4 per sorted leaf: (v0, triangle id bits), (v1, 0), (v2, 0).
constexpr uint32_t kQLeft = 1u << 30, kQRight = 1u << 31, kQGamma = kQLeft - 1;
constexpr int kQuantEmin = -64, kQuantEmax = 63;
struct BvhView {
    const uint4* qnodes;
    const float4* tris;
    int64_t n;
};

struct BvhBuffers {
    // Karras layout kept for parity checks and refits
    uint32_t* codes = nullptr;     // sorted codes, n
    uint32_t* order = nullptr;     // sorted position -> triangle id, n
    uint32_t* left = nullptr;      // n-1
    uint32_t* right = nullptr;     // n-1
    uint32_t* parent = nullptr;    // 2n-1 (internal, then leaves)
    float* lo = nullptr;           // 3*(2n-1)
    float* hi = nullptr;           // 3*(2n-1)
    uint4* qnodes = nullptr;       // 2*(n-1): compressed records for the traversal
    float4* tris = nullptr;        // 3n
};

struct Counters64 {
    unsigned long long nodes, tris, rays;
};

// ---- launchers (bvh.hip) ----
// Builds all BvhBuffers from de-indexed positions (device, 9 floats/triangle).
// scratch must hold bvh_scratch_bytes(n).
size_t bvh_scratch_bytes(int64_t n);
// Device word set non-zero by the build if the hierarchy came out malformed.
uint32_t* bvh_error_word(void* d_scratch, int64_t n);
// rounds: treelet restructuring rounds after the LBVH (0: the plain LBVH).
constexpr int kTreeletRoundsDefault = 1;
hipError_t bvh_build(const float* d_positions, int64_t n, BvhBuffers& out, void* d_scratch, hipStream_t s,
                     int rounds);

// ---- launchers (render.hip) ----
// gstack: (kStackTotal - kStackLdsMin) * grid * kTraceBlock words; overflow:
// one int set non-zero if a traversal ever needed more than kStackTotal entries.
hipError_t trace_rays(const BvhView& b, const float* d_rays, int64_t n, int any, float4* d_hits, uint8_t* d_occ,
                      uint32_t* d_counts, uint32_t* gstack, int grid, int* overflow, hipStream_t s);
inline size_t gstack_words(int grid) { return (size_t)(kStackTotal - kStackLdsMin) * grid * kTraceBlock; }
// Ray queries that allocate nothing (oxr_trace_rays_noalloc): k_trace_rays
// on at most kNoAllocTraceGrid blocks, its stack spill and overflow word in a
// fixed device array (render.hip), 32 MiB + 256 B per device.
constexpr int kMaxDevices = 64;
constexpr int kNoAllocTraceGrid = 1024;
constexpr size_t kNoAllocSpillWords = (size_t)(kStackTotal - 32) * kNoAllocTraceGrid * kTraceBlock;
constexpr size_t kNoAllocScratchWords = kNoAllocSpillWords + 64;
constexpr int64_t kNoAllocMaxRays = (int64_t)1 << 31;  // rays per call
// The device address of the fixed scratch on `device` (looked up once per
// device with that device made current; the caller restores its device).
hipError_t noalloc_scratch(int device, uint32_t** out);
hipError_t trace_rays_noalloc(const BvhView& b, const float* d_rays, int64_t n, int any, float4* d_hits,
                              uint8_t* d_occ, uint32_t* scratch_ptr, hipStream_t s);
// lum_sum: the sequential float sum of the texel luminances in row-major
// order (the reference's host loop), computed by env_luminance_sum.
hipError_t env_cache_build(const float4* d_env, int w, int h, float lum_sum, float4* d_cache, float4* d_env_pdf,
                           float* d_scratch,
                           hipStream_t s);
float env_luminance_sum(const float* rgba, int64_t texels);
size_t env_scratch_floats(int w, int h);
hipError_t tonemap(const float4* d_in, uint32_t* d_out, int64_t n, hipStream_t s);
hipError_t material_batch(const oxr_material* d_mat, const float* d_lut_bsdf, const float* d_lut_albedo, int mode,
                          const float* d_in, int64_t n, float* d_out, hipStream_t s);
// d_uvs / d_uvs_sorted may be null (no textures).
hipError_t gather_sorted_attributes(const float* d_normals, const uint32_t* d_mat_ids, const float* d_uvs,
                                    const uint32_t* d_order, int64_t n, float* d_normals_sorted,
                                    uint32_t* d_mat_sorted, float* d_uvs_sorted, hipStream_t s);

// One render request's share of a wavefront window.  A window holds
// `frames` frames of every pixel of each request, laid out request after
// request: sample id = path_begin + s * width * height + pixel.
// Exact unsigned 32-bit division by a divisor fixed ahead of time, as a
// multiply-high and two shifts (Granlund & Montgomery 1994, the round-up
// method): with L = ceil(log2 d) and m = floor(2^32 (2^L - d) / d) + 1,
// n / d = (t + ((n - t) >> min(L, 1))) >> max(L - 1, 0) where t = mulhi(m, n).
struct Divider {
    uint32_t m, s1, s2;
};
inline Divider make_divider(uint32_t d) {
    uint32_t L = 0;
    while (L < 32 && (1ull << L) < d) L++;
    const uint64_t m = ((((uint64_t)1 << L) - d) << 32) / d + 1;
    return Divider{(uint32_t)m, L < 1 ? L : 1u, L > 1 ? L - 1 : 0u};
}
__device__ __forceinline__ uint32_t udiv(uint32_t n, Divider d) {
    const uint32_t t = __umulhi(d.m, n);
    return (t + ((n - t) >> d.s1)) >> d.s2;
}

struct ReqDesc {
    int64_t path_begin;    // first sample id of this request in the window
    int64_t pix_begin;     // first pixel of this request among the window's pixels
    int width, height;
    int frame0, frames;    // frameIDs frame0 .. frame0 + frames - 1
    int filter;
    int path_frame_offset; // frame index of frame0 in path_radiance (debug output)
    float cam_pos[3], cam_dir[3], cam_h[3], cam_v[3];
    float4* color;
    float4* albedo;
    float4* normal;
    float* depth;          // first-hit depth d / (1 + d), 1 for a miss
    float* path_radiance;  // optional, frame-major 3 floats per sample
    // Camera-ray cover mask, one byte per kCoverTile x kCoverTile pixel tile
    // (row-major, tiles_w per row): 0 = no camera ray of the tile's pixels can
    // hit the scene, so its samples skip the first trace.  nullptr: none.
    const uint8_t* cover;
    int tiles_w;
    Divider by_pixels, by_width;  // division by width * height and by width
};

// Camera-ray culling (build_cover_mask): the boxes of the BVH's top levels,
// at most kCoverBoxes, projected per request onto tiles of kCoverTile pixels.
// kCoverMargin pixels cover the largest pixel-filter offset (the Gaussian's
// 0.5 + 0.375 sqrt(-2 ln FLT_MIN) < 5.5) plus 2 pixels for rounding.
constexpr int kCoverBoxes = 512;
constexpr int kCoverTile = 8;
constexpr double kCoverMargin = 8.0;
constexpr int cover_tiles(int n) { return (n + kCoverTile - 1) / kCoverTile; }

// The path pool is split into kShards independent shards of pool/kShards
// slots.  Wave w of any stage kernel serves shard w % kShards, so with the
// round-robin dispatch of workgroups over the 8 XCDs a shard's paths stay on
// one XCD, and every queue counter, fetch counter and sample cursor exists
// once per shard: same-address atomics serialise (about 11 ns each on
// MI355X), different addresses proceed in parallel.
constexpr int kShards = 32;
constexpr int kCtrStride = 64;  // ints between counters: one 256-byte line each
constexpr int kQueueRegions = 2;  // per shard: survivors, new paths

// Counter ids (int index = id * kCtrStride).
// Lists written by the stage at parity y and read by a later stage carry a
// parity index; see wave_trace for who writes, reads and clears each.
enum CtrId {
    // [parity][region]: region k survivors, kShards + k new paths
    kCtrQueue = 0,
    kCtrTerm = kCtrQueue + 2 * kQueueRegions * kShards,  // [parity][shard]: paths ended with shadow rays pending
    kCtrShEnv = kCtrTerm + 2 * kShards,  // [parity][shard]: env shadow rays listed
    kCtrShLight = kCtrShEnv + 2 * kShards,     // [parity][shard]: light shadow rays listed
    kCtrFetchTrace = kCtrShLight + 2 * kShards,
    kCtrFetchRegen = kCtrFetchTrace + kShards,
    kCtrSamplePos = kCtrFetchRegen + kShards,  // per shard: samples drawn so far
    kCtrWindowTotal = kCtrSamplePos + kShards, // samples in the current window
    kCtrShardSamples = kCtrWindowTotal + 1,    // per shard: samples listed for tracing by the sky pass
    kCtrTailMode = kCtrShardSamples + kShards, // 1 while the current iteration runs in tail mode (k_tail)
    kCtrClass = kCtrTailMode + 1,              // [parity][class][shard]: entries of the shade-class queues
    kCtrFetchClass = kCtrClass + 2 * 10 * kShards, // [class][shard]: chunks taken by the shade kernels
    kCtrIds = kCtrFetchClass + 10 * kShards
};
// Counter of queue x's region r (0 survivors, 1 new paths) of shard k.
constexpr int queue_ctr(int x, int r, int k) { return kCtrQueue + (x * kQueueRegions + r) * kShards + k; }
// The host's completion poll copies counters [kCtrQueue, kPollIds): the
// queue and pending-finish counts of both parities.
constexpr int kPollIds = kCtrTerm + 2 * kShards;
constexpr size_t ctr_ints() { return (size_t)kCtrIds * kCtrStride; }
// Shade classes: 0 = a path past bounce 0 whose ray left the scene, 1 + t =
// the ray hit a surface of material type t, 9 = a camera ray (bounce 0)
// that left the scene.
constexpr int kShadeClasses = 10;
constexpr int kClassCameraMiss = 9;

// Wavefront integrator state for one context.
struct WaveParams {
    // scene
    BvhView bvh;
    const float* normals;          // 9 floats per sorted triangle
    const uint32_t* mat_ids;       // per sorted triangle
    const oxr_material* mats;
    const oxr_light* lights;
    int n_lights;
    const oxr_medium* media;
    int n_media;
    const float4* env;
    const float4* env_cache;
    const float4* env_pdf;  // per texel (env r, g, b, cache pdf): the miss shade and env NEE read both
    int env_w, env_h;
    const float* lut_bsdf;
    const float* lut_albedo;
    // material textures: RGBA8 texels of every image, tex_info[id] = (first
    // texel, width, height); uvs: 6 floats per sorted triangle.  textured = 0:
    // no material uses a texture (uvs and texels unused)
    const uint32_t* texels;
    const int4* tex_info;
    const float* uvs;
    int textured;
    int cam_medium;
    int max_depth, rr_start;
    // path pool: `pool` slots, shard k owns slots [k*region, (k+1)*region)
    int64_t pool;
    int region;        // pool / kShards
    // one 64-byte record per slot (regeneration scatters slots over the
    // queues, so a slot's fields share a cache line):
    //   [0] origin, pre_pdf   [1] direction, rng state bits
    //   [2] throughput, bounce (int bits)   [3] radiance, sample id (uint bits; kNoSample: none)
    float4* path;
    uint32_t* queue[2];  // 2*kShards regions of `region` slot ids each
    uint32_t* fresh_g[2];  // sample ids of the new-path regions (kShards regions of `region`)
    float4* hits;      // per queue position (2*kShards regions)
    // the window's samples whose camera ray is traced, per shard in draw
    // order (shard k's at samples + k * shard_cap), listed by the sky pass
    uint32_t* samples;
    int64_t shard_cap;
    // shadow rays, one record of 3 float4 per slot and kind (env / light):
    //   [0] origin, tmax   [1] direction   [2] contribution, pending flag
    // The flag is 1 when the slot's last shade emitted the ray and the trace
    // found it unoccluded, 0 otherwise; the slot's next shade (or k_regen)
    // adds the contribution.
    float4* sh_env;
    float4* sh_light;
    uint32_t* sh_list[2];  // [parity]: slot ids with an env ray in kShards regions, then with a light ray
    uint32_t* term[2];     // [parity]: kShards regions of `region` slot ids
    // shade classes (kShadeClasses): cls_queue[class][shard] holds `region`
    // queue positions.
    uint32_t* cls_queue;
    uint32_t class_mask;  // classes the scene can produce (bit c): the shade launches
    // per-sample outputs of the window
    float4* s_rad;
    float4* s_alb;
    float4* s_nrm;
    float* s_dep;
    int* ctr;          // device counters, ctr_ints() ints
    const ReqDesc* reqs;  // device table of the window's requests
    int nreq;
    // [0,1] extension-ray node fetches / triangle tests (when counting),
    // [2] extension rays, [3] shadow rays, [4,5] shadow-ray node fetches /
    // triangle tests, [6] camera rays the sky pass resolved (counting: that
    // it would resolve), [7,8] node fetches / triangle tests of those
    // (counting only; included in [0,1]), [9,10] node fetches / triangle
    // tests of the tail kernel's rays (counting only; included in [0,1,4,5]),
    // [11,12] extension / shadow rays the tail kernel traced (in [2,3])
    unsigned long long* traversal;
    uint32_t* gstack;               // traversal stack spill, (kStackTotal - kStackLdsMin) * grid * kTraceBlock
    int stack_lds;                  // 8, 16 or 32 stack entries in LDS (selects the kernel instance)
    int* overflow;
    // 1: the sky pass shades the samples in uncovered tiles; 2: all traced,
    // and a counting trace tallies the ones it would shade; 0: all traced
    // (wave_resolve_mode)
    int resolve;
    // tail mode: an iteration whose queue holds at most tail_paths paths
    // (and at least one) shades them in k_tail, each carried to its end in
    // that launch, instead of one launch per bounce; 0 = never
    int64_t tail_paths;
    const float4* cover_boxes;  // the scene's cover set (build_cover_set)
    const int* cover_count;
};
constexpr uint32_t kNoSample = 0xffffffffu;
constexpr int kTravCounters = 16;  // entries of WaveParams::traversal

// Window set-up: rewrites frame0 / path_frame_offset / frames of request 0
// (single renders; the server writes its table from the host), zeroes every
// counter and stores the window's sample count.  Then build_cover_masks,
// wave_sky and wave_fill, which starts a listed sample in every slot (queue 0).
// grid: any wave_grid size.
hipError_t wave_window_init(const WaveParams& p, bool set_req0, int frame0, int path_frame_offset, int frames,
                            int64_t total_samples, hipStream_t s);
// Sky pass: every sample of the window whose pixel tile is clear in its
// request's cover mask is shaded by wave_sky_shade -- its camera ray leaves
// the scene, so it meets at most an analytic light or the environment --
// and wave_sky_list lists the others per shard for tracing
// (WaveParams::samples).  The shading is read only by the accumulation, so
// it may run on another stream beside the path loop.  grid: a wave_grid size.
hipError_t wave_sky_list(const WaveParams& p, int grid, hipStream_t s);
hipError_t wave_sky_shade(const WaveParams& p, int grid, hipStream_t s);
// The resolve mode for a render: the sky pass needs no medium around the
// camera, nor lights and media together (a light inside a medium would need
// the medium sampled); a counting render traces everything.
inline int wave_resolve_mode(const WaveParams& p, bool counting) {
    if (p.cam_medium != -1 || (p.n_lights > 0 && p.n_media > 0)) return 0;
    return counting ? 2 : 1;
}
hipError_t wave_fill(const WaveParams& p, int grid, hipStream_t s);
// Scene cover set: the boxes of the BVH's top levels (breadth first, at most
// kCoverBoxes; leaves met on the way included), every triangle inside one.
// boxes: 2 float4 (lo, hi) each; count: one int.
hipError_t build_cover_set(const BvhBuffers& b, int64_t n, float4* boxes, int* count, hipStream_t s);
// Cover masks of requests r0 .. r0 + nreq - 1 of the device table: a tile is
// 1 when the camera ray of some sample of its pixels may enter a cover box.
hipError_t build_cover_masks(const ReqDesc* reqs, int nreq, int max_tiles, const float4* boxes, const int* count,
                             hipStream_t s);
// One iteration at parity y is trace(y) -> regen(y) -> shade(y):
//   trace(y): extension rays of queue y and the shadow rays listed at parity
//     y^1 (by the previous shade) in one launch; hits go to the shade-class
//     lists of parity y, occluded shadow rays clear their flags.
//   regen(y): finishes the paths listed at parity y^1 as ended with shadow
//     rays pending and starts new samples in their slots (queue y^1).
//   shade(y): one launch per class in p.class_mask; adds the slot's pending
//     shadow contributions, then shades; survivors and new samples go to
//     queue y^1, shadow rays and pending finishes are listed at parity y.
// Counters of parity y are cleared by trace(y) for lists its iteration
// refills (queue y^1 regions, pending finishes and shadow rays of parity y,
// shade classes of parity y^1); fetch counters by the stage before their user.
hipError_t wave_trace(const WaveParams& p, int y, int grid, bool count, hipStream_t s);
// shade(y) as the shade-class launches and the tail kernel: the class
// kernels return at once in tail mode, the tail kernel at once otherwise.
// trace_grid: the tail kernel traverses, so it uses the trace grid's stacks.
hipError_t wave_shade(const WaveParams& p, int y, int grid, int trace_grid, bool count, hipStream_t s);
hipError_t wave_regen(const WaveParams& p, int y, int grid, hipStream_t s);
hipError_t wave_accumulate(const WaveParams& p, int64_t npix, hipStream_t s);
// ---- denoiser (denoise.hip) ----
// KPCNN 10 -> 32 -> 32 -> 32 -> 32 -> 25, 3x3 convolutions.  lanes: BF16
// weights in MFMA lane order, layer l at lane_offset[l] (in 8-element groups):
// for tap t, K step s, lane q, element e: W[co][ci = 16 s + 8 (q / 32) + e]
// [t / 3][t % 3] with co the output channel of MFMA row q % 32 (the row
// itself, except in the last layer: row 8 g + 4 h + i holds channel 4 g + i
// for h = 0 and 13 + 4 g + i for h = 1), zero where co or ci is padding.
// bias: 5 x 32 floats, by MFMA row.
constexpr int kDnLayers = 5;
constexpr int kDnChannels[kDnLayers + 1] = {10, 32, 32, 32, 32, 25};
struct DenoiserDev {
    const uint16_t* lanes;
    int64_t lane_offset[kDnLayers];
    const float* bias;
};
// The network in one kernel (activations in LDS, no scratch); cus: compute
// units of the device (one block each at a time).
hipError_t denoise_run(const DenoiserDev& d, const float4* color, const float4* albedo, const float4* normal,
                       const float* depth, int w, int h, int cus, float4* out, hipStream_t s);

// Largest resident workgroups per CU of the trace instance for stack_lds,
// and of shade / regen.
int wave_trace_occupancy(int stack_lds);
int wave_shade_occupancy();

}  // namespace oxr
