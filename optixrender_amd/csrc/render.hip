// Wavefront path tracer on gfx950 with path regeneration, plus the
// environment importance cache, the tone map and the test entry points.
// Replaces __raygen__renderFrame and its closest-hit/miss programs
// (core/cuda/devicePrograms.cu:68-532) and ComputeFinalPixelColorsKernel
// (core/cuda/PostProcessing.cu:26-48).
//
// A sample is (pixel, frame): RNG init(ix + W*iy, frameID), one sample per
// frame -- the reference with numPixelSamples = 1 in progressive mode.  A
// resident pool of path slots (SoA in HBM, 64 B/slot: origin+pre_pdf,
// direction+rng, throughput+bounce, radiance+sample id) runs the loop
//   extend -> shade -> connect(env) -> connect(lights) -> regen
// until the window's samples are done: regen writes each ended path's
// radiance to its sample's slot in s_rad and starts the next sample in the
// freed slot, so every launch works on a full pool instead of a shrinking
// batch.  The pool is split into kShards shards with their own queues and
// counters (common.h); stage kernels are persistent and a wave takes 64 queue
// entries of its shard per atomic.
#include <algorithm>
#include <mutex>
#include "common.h"
#include "shading.h"
#include "traverse.h"

namespace oxr {
namespace {

OXR_D int lane_id() { return (int)(threadIdx.x & 63); }
OXR_D uint64_t lanes_below() { int l = lane_id(); return l == 0 ? 0ull : (~0ull >> (64 - l)); }

// Wave-aggregated append: returns this lane's slot or -1.
OXR_D int wave_append(int* ctr, bool want) {
    uint64_t m = __ballot(want);
    if (m == 0) return -1;
    int base = 0;
    if (lane_id() == 0) base = atomicAdd(ctr, __popcll(m));
    base = __shfl(base, 0);
    return want ? base + (int)__popcll(m & lanes_below()) : -1;
}

OXR_D void wave_count_add(unsigned long long* dst, uint32_t v) {
    unsigned long long s = v;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane_id() == 0 && s) atomicAdd(dst, s);
}

OXR_D void wave_count_add64(unsigned long long* dst, unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (lane_id() == 0 && v) atomicAdd(dst, v);
}

OXR_D uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, off));
    return v;
}

// ------------------------------------------------------------ ray queries
template <bool ANY>
__global__ __launch_bounds__(kTraceBlock) void k_trace_rays(BvhView b, const float* __restrict__ rays, int64_t n,
                                                             float4* hits, uint8_t* occ, uint32_t* counts,
                                                             uint32_t* gstack, int* overflow) {
    __shared__ uint32_t lds[32 * kTraceBlock];
    int64_t gt = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t gstride = (int64_t)gridDim.x * blockDim.x;
    Stack<32> st(lds, gstack, gstride);
    for (int64_t i = gt; i < n; i += gstride) {
        const float* r = rays + 8 * i;
        uint32_t nn = 0, nt = 0;
        Hit h = traverse<ANY, true>(b, f3(r[0], r[1], r[2]), f3(r[4], r[5], r[6]), r[3], r[7], st, nn, nt, overflow);
        if (ANY) occ[i] = h.leaf != kNoHit;
        else hits[i] = make_float4(h.t, __uint_as_float(h.id), h.u, h.v);
        if (counts) { counts[2 * i] = nn; counts[2 * i + 1] = nt; }
    }
}

// ------------------------------------------------------------ environment cache
// HdrTexture::CalculateHdrCache (core/src/Scene.cpp:217-313) with every sum
// taken in the reference's order, so the cache is bit-identical to the
// host routine: the global luminance sum runs in row-major order on one
// lane, the per-column marginal and conditional CDFs run down each column on
// one lane per column, the marginal CDF runs across columns on one lane.
__global__ void k_env_lum(const float4* __restrict__ env, int64_t m, float* lum) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= m) return;
    float4 t = env[i];
    lum[i] = 0.299f * t.x + 0.587f * t.y + 0.114f * t.z;  // Utils.h:76-78
}
__global__ void k_env_columns(const float* __restrict__ lum, float total, int w, int h, float* pdf, float* marg,
                              float* cdfy) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= w) return;
    float m = 0.0f;
    for (int i = 0; i < h; i++) {
        float p = lum[(size_t)i * w + j] / total;
        pdf[(size_t)i * w + j] = p;
        m += p;
    }
    marg[j] = m;
    float acc = 0.0f;
    for (int i = 0; i < h; i++) {
        float c = pdf[(size_t)i * w + j] / m;
        acc = (i == 0) ? c : acc + c;
        cdfy[(size_t)j * h + i] = acc;
    }
}
__global__ void k_env_cdfx(const float* __restrict__ marg, int w, float* cdfx) {
    if (threadIdx.x != 0) return;
    float acc = marg[0];
    cdfx[0] = acc;
    for (int j = 1; j < w; j++) {
        acc = acc + marg[j];
        cdfx[j] = acc;
    }
}
// std::lower_bound as libstdc++ implements it (so NaN entries behave alike).
OXR_D int lower_bound_f(const float* a, int n, float v) {
    int first = 0, len = n;
    while (len > 0) {
        int half = len >> 1;
        if (a[first + half] < v) { first = first + half + 1; len = len - half - 1; }
        else len = half;
    }
    return first;
}
__global__ void k_env_search(const float* __restrict__ pdf, const float* __restrict__ cdfx,
                             const float* __restrict__ cdfy, int w, int h, float4* cache) {
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)w * h) return;
    int i = (int)(t / w), j = (int)(t % w);
    float xi1 = (float)i / (float)h, xi2 = (float)j / (float)w;
    int x = lower_bound_f(cdfx, w, xi1);
    int y = x < w ? lower_bound_f(cdfy + (size_t)x * h, h, xi2) : 0;
    cache[t] = make_float4((float)x / (float)w, (float)y / (float)h, pdf[t], 0.0f);
}

// ------------------------------------------------------------ tone map
// ComputeFinalPixelColorsKernel (PostProcessing.cu:26-48).
__global__ void k_tonemap(const float4* __restrict__ in, uint32_t* out, int64_t n) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 c = in[i];
    float lum = 0.299f * c.x + 0.587f * c.y + 0.114f * c.z;
    float f = 1.0f / (1.0f + lum / 1.5f);
    float r = clampf(powf(c.x * f, 1.0f / 2.2f), 0.0f, 1.0f);
    float g = clampf(powf(c.y * f, 1.0f / 2.2f), 0.0f, 1.0f);
    float b = clampf(powf(c.z * f, 1.0f / 2.2f), 0.0f, 1.0f);
    out[i] = (uint32_t)(r * 255.9f) | ((uint32_t)(g * 255.9f) << 8) | ((uint32_t)(b * 255.9f) << 16) | (255u << 24);
}

// ------------------------------------------------------------ material batch
__global__ void k_material_batch(const oxr_material* m, Lut lut, int mode, const float* __restrict__ in, int64_t n,
                                 float* out) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* a = in + 12 * i;
    Surf s{f3(a[0], a[1], a[2]), a[9] != 0.0f, m};
    f3 V(a[3], a[4], a[5]);
    float pdf = 0.0f;
    f3 f, L(a[6], a[7], a[8]);
    Rng rng{__float_as_uint(a[10])};
    if (mode == 0) f = eval_material(s, lut, V, L, pdf);
    else f = sample_material(s, lut, rng, V, L, pdf);
    float* o = out + 8 * i;
    o[0] = f.x; o[1] = f.y; o[2] = f.z; o[3] = pdf;
    o[4] = L.x; o[5] = L.y; o[6] = L.z; o[7] = __uint_as_float(rng.s);
}

__global__ void k_gather(const float* __restrict__ nrm, const uint32_t* __restrict__ mid,
                         const float* __restrict__ uv, const uint32_t* __restrict__ order, int64_t n, float* nrm_s,
                         uint32_t* mid_s, float* uv_s) {
    int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint32_t t = order[k];
#pragma unroll
    for (int c = 0; c < 9; c++) nrm_s[9 * k + c] = nrm[9 * (size_t)t + c];
    mid_s[k] = mid[t];
    if (uv) {
#pragma unroll
        for (int c = 0; c < 6; c++) uv_s[6 * k + c] = uv[6 * (size_t)t + c];
    }
}

// ------------------------------------------------------------ wavefront
OXR_D int* ctr_at(int* ctr, int id) { return ctr + (size_t)id * kCtrStride; }
// Shard served by the calling wave (wave-uniform).
OXR_D int wave_shard() { return (int)((blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) & (kShards - 1)); }
// A wave takes the next chunk index of its shard.
OXR_D int wave_next(int* ctr) {
    int j = 0;
    if (lane_id() == 0) j = atomicAdd(ctr, 1);
    return __shfl(j, 0);
}

// Queue `x` of shard k: region k (paths that survived the last shade) then
// region kShards + k (paths started by the last regen), in 64-slot chunks.
struct ShardQueue {
    int n_surv, n_new, c_surv, chunks;
    OXR_D ShardQueue(int* ctr, int x, int k) {
        n_surv = *ctr_at(ctr, queue_ctr(x, 0, k));
        n_new = *ctr_at(ctr, queue_ctr(x, 1, k));
        c_surv = (n_surv + 63) >> 6;
        chunks = c_surv + ((n_new + 63) >> 6);
    }
    // Queue position of this lane in chunk j, or -1 past the region's end.
    OXR_D int64_t position(int j, int k, int region) const {
        const bool surv = j < c_surv;
        const int slot = ((surv ? j : j - c_surv) << 6) + lane_id();
        if (slot >= (surv ? n_surv : n_new)) return -1;
        return (int64_t)(surv ? k : kShards + k) * region + slot;
    }
};

// Request owning a sample (or pixel) index: the last descriptor whose begin <= i.
template <bool PIX>
OXR_D int find_req(const ReqDesc* r, int n, int64_t i) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        int64_t b = PIX ? r[mid].pix_begin : r[mid].path_begin;
        if (b <= i) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Sample id of position `pos` of shard k's draw sequence: shard k owns the
// 64-sample blocks k, k + kShards, k + 2 kShards, ... so a wave that draws
// consecutive positions starts 64 neighbouring pixels of one frame.
OXR_D int64_t sample_of(int k, int pos) { return (((int64_t)(pos >> 6) * kShards + k) << 6) + (pos & 63); }

// Camera ray of sample g (devicePrograms.cu:214-236) and the RNG state after
// the two jitter draws.  A path at bounce 0 has no slot record: extend and
// shade recompute this from the sample id, and the record is written only
// once the path survives its first shade.
OXR_D void camera_ray(const WaveParams& p, int64_t g, f3& org, f3& dir, uint32_t& rng_state) {
    const ReqDesc& q = p.reqs[find_req<false>(p.reqs, p.nreq, g)];
    // a window holds fewer than 2^31 samples: 32-bit division is exact
    const uint32_t local = (uint32_t)(g - q.path_begin);
    const uint32_t npix = (uint32_t)(q.width * q.height), w = (uint32_t)q.width;
    const uint32_t s = udiv(local, q.by_pixels), pix = local - s * npix;
    const uint32_t iyu = udiv(pix, q.by_width);
    const int ix = (int)(pix - iyu * w), iy = (int)iyu;
    Rng rng{lcg_seed((uint32_t)(ix + q.width * iy), (uint32_t)q.frame0 + s)};
    float j0 = rng(), j1 = rng();
    f2 jit = filter_jitter(f2{j0, j1}, q.filter);                         // devicePrograms.cu:220-227
    float sx = ((float)ix + 0.5f + jit.x) / (float)q.width;
    float sy = ((float)iy + 0.5f + jit.y) / (float)q.height;
    dir = normalize(make3(q.cam_dir) + (sx - 0.5f) * make3(q.cam_h) + (sy - 0.5f) * make3(q.cam_v));
    org = make3(q.cam_pos);
    rng_state = rng.s;
}

// The new-path region of queue x: positions at or past kShards * region.
OXR_D bool is_fresh(const WaveParams& p, int64_t qp) { return qp >= (int64_t)kShards * p.region; }
OXR_D uint32_t fresh_sample(const WaveParams& p, int x, int64_t qp) {
    return p.fresh_g[x][qp - (int64_t)kShards * p.region];
}

// Whether the camera ray of sample g surely leaves the scene: its pixel's
// tile is clear in the request's cover mask.
OXR_D bool camera_ray_misses(const WaveParams& p, int64_t g) {
    const ReqDesc& q = p.reqs[find_req<false>(p.reqs, p.nreq, g)];
    if (!q.cover) return false;
    const uint32_t local = (uint32_t)(g - q.path_begin);  // a window holds fewer than 2^31 samples
    const uint32_t w = (uint32_t)q.width, npix = w * (uint32_t)q.height;
    const uint32_t pix = local - udiv(local, q.by_pixels) * npix;
    const uint32_t iy = udiv(pix, q.by_width), ix = pix - iy * w;
    return q.cover[(iy / kCoverTile) * (uint32_t)q.tiles_w + ix / kCoverTile] == 0;
}

// Draws the next listed sample of shard k for every lane that asks
// (consecutive positions of the shard's list, one atomic per wave), starts it
// in the lane's slot and appends the slot to the new-path region of queue y.
// All lanes of the wave must call it.  Returns whether this lane started a
// path.
OXR_D bool restart(const WaveParams& p, int k, int y, bool want, uint32_t pid) {
    const uint64_t m = __ballot(want);
    if (m == 0) return false;
    int base = 0;
    if (lane_id() == 0) base = atomicAdd(ctr_at(p.ctr, kCtrSamplePos + k), (int)__popcll(m));
    base = __shfl(base, 0);
    const int listed = *ctr_at(p.ctr, kCtrShardSamples + k);
    const int pos = base + (int)__popcll(m & lanes_below());
    const bool start = want && pos < listed;
    const int o = wave_append(ctr_at(p.ctr, queue_ctr(y, 1, k)), start);
    if (o >= 0) {
        p.queue[y][(size_t)(kShards + k) * p.region + o] = pid;
        p.fresh_g[y][(size_t)k * p.region + o] = p.samples[(size_t)k * p.shard_cap + pos];
    }
    return start;
}

__global__ void k_window_init(WaveParams p, ReqDesc* r0, int set_req0, int frame0, int offset, int frames,
                              int total) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t < (int64_t)ctr_ints()) {
        const int id = (int)(t / kCtrStride);
        int v = 0;
        if (t % kCtrStride == 0) {
            // k_fill draws positions 0 .. region-1 of every shard's list
            if (id >= kCtrSamplePos && id < kCtrSamplePos + kShards) v = p.region;
            if (id == kCtrWindowTotal) v = total;
        }
        p.ctr[t] = v;
    }
    if (t == 0 && set_req0) {
        r0->frame0 = frame0;
        r0->path_frame_offset = offset;
        r0->frames = frames;
    }
}

// lightTrace (defined with the shading below).
OXR_D bool light_trace(const WaveParams& p, f3 o, f3 d, int& mi, f3& Lr, float& lpdf, float& closest);

// Sky pass.  Shard k owns the window's 64-sample blocks k, k + kShards, ...
// (the order in which it draws them, so that a wave starts 64 neighbouring
// pixels of one frame).  Two kernels share the classification:
//   k_sky_list: lists the samples to trace.  A wave takes the shard's blocks
//     in a fixed stride; a first pass counts them and reserves their list
//     positions with one atomic, a second writes them.
//   k_sky_shade: shades the others.  Their camera ray leaves the scene, so
//     __miss__radiance / the light hit of devicePrograms.cu:342-373 at bounce
//     0 decide the sample, exactly as the camera-miss shade would.  Nothing
//     reads these samples before k_accumulate, so it runs beside the path
//     loop on a second stream.
// Needs no medium around the camera, nor one around a light the ray could
// meet (the caller's resolve mode).
OXR_D int sky_class(const WaveParams& p, int k, int j, int64_t total, int64_t& g) {  // 0 none, 1 trace, 2 here
    g = (((int64_t)j * kShards + k) << 6) + lane_id();
    if (g >= total) return 0;
    return (p.resolve == 1 && camera_ray_misses(p, g)) ? 2 : 1;
}
OXR_D int sky_blocks(int k, int64_t total) {
    const int64_t nb = (total + 63) >> 6;
    return nb > k ? (int)((nb - k + kShards - 1) / kShards) : 0;
}

// The first pass keeps each block's mask of samples to trace in LDS (up to
// kSkyMasks blocks per wave), so that the second pass, which writes the
// list, does not classify them again.
constexpr int kSkyMasks = 512;

__global__ __launch_bounds__(256) void k_sky_list(WaveParams p) {
    __shared__ uint64_t masks[256 / 64][kSkyMasks];
    uint64_t* mine = masks[threadIdx.x >> 6];
    const int k = wave_shard();
    const int wave = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    const int stride = (int)((gridDim.x * (blockDim.x >> 6)) / kShards);  // waves per shard
    const int64_t total = *ctr_at(p.ctr, kCtrWindowTotal);
    const int blocks_k = sky_blocks(k, total);
    const int first = wave / kShards;
    int n_trace = 0;
    for (int j = first, i = 0; j < blocks_k; j += stride, i++) {
        int64_t g;
        const uint64_t mt = __ballot(sky_class(p, k, j, total, g) == 1);
        if (i < kSkyMasks && lane_id() == 0) mine[i] = mt;
        n_trace += (int)__popcll(mt);
    }
    int base = 0;
    if (lane_id() == 0 && n_trace) base = atomicAdd(ctr_at(p.ctr, kCtrShardSamples + k), n_trace);
    base = __shfl(base, 0);
    uint32_t* list = p.samples + (size_t)k * p.shard_cap;
    for (int j = first, i = 0; j < blocks_k; j += stride, i++) {
        // the first pass's mask, or the classification again past the LDS array
        uint64_t mt;
        if (i < kSkyMasks) {
            mt = mine[i];
        } else {
            int64_t g;
            mt = __ballot(sky_class(p, k, j, total, g) == 1);
        }
        if ((mt >> lane_id()) & 1)
            list[base + (int)__popcll(mt & lanes_below())] = (uint32_t)((((int64_t)j * kShards + k) << 6) + lane_id());
        base += (int)__popcll(mt);
    }
}

__global__ __launch_bounds__(256) void k_sky_shade(WaveParams p) {
    const int k = wave_shard();
    const int wave = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    const int stride = (int)((gridDim.x * (blockDim.x >> 6)) / kShards);  // waves per shard
    const int64_t total = *ctr_at(p.ctr, kCtrWindowTotal);
    const int blocks_k = sky_blocks(k, total);
    uint32_t resolved = 0;
    for (int j = wave / kShards; j < blocks_k; j += stride) {
        int64_t g;
        if (sky_class(p, k, j, total, g) != 2) continue;
        resolved++;
        f3 org, dir;
        uint32_t rs;
        camera_ray(p, g, org, dir, rs);
        f3 radiance(0.0f), history(1.0f);
        float closest = kFltMax, light_pdf = 0.0f;
        f3 light_radiance(0.0f);
        int mi_light = 0;
        if (light_trace(p, org, dir, mi_light, light_radiance, light_pdf, closest)) {
            const float mw = 1.0f;                                          // devicePrograms.cu:342-351
            radiance = radiance + mw * history * light_radiance;
            p.s_nrm[g] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
            p.s_alb[g] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
            p.s_dep[g] = closest / (1.0f + closest);
            p.s_rad[g] = make_float4(radiance.x, radiance.y, radiance.z, 0.0f);
        } else {
            if (p.env_w > 0) {                                              // devicePrograms.cu:354-373
                f2 uv = sphere_to_plane(dir);
                float4 e = bilinear(p.env_pdf, p.env_w, p.env_h, uv.x, uv.y);
                const f3 lr(e.x, e.y, e.z);
                const float mw = 1.0f;
                radiance = radiance + mw * history * lr;
            }
            // default AOVs (albedo and normal 0, depth 1), flagged in w
            p.s_rad[g] = make_float4(radiance.x, radiance.y, radiance.z, 1.0f);
        }
    }
    wave_count_add(&p.traversal[6], resolved);
}

// Window start: slot s of shard k takes position s of the shard's list, if
// listed.  The started slots are a prefix of the shard, so queue position s
// of the new-path region of queue 0 is slot s and the region's count is set
// once.  Waves take the shard's 64-slot chunks in a fixed stride.
__global__ __launch_bounds__(256) void k_fill(WaveParams p) {
    const int k = wave_shard();
    const int wave = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    const int stride = (int)((gridDim.x * (blockDim.x >> 6)) / kShards);  // waves per shard
    const int listed = *ctr_at(p.ctr, kCtrShardSamples + k);
    const int chunks = p.region >> 6;  // region is a multiple of 64
    const size_t rbase = (size_t)k * p.region;
    uint32_t* qn = p.queue[0] + (size_t)(kShards + k) * p.region;
    uint32_t* gn = p.fresh_g[0] + rbase;
    const uint32_t* list = p.samples + (size_t)k * p.shard_cap;
    for (int j = wave / kShards; j < chunks; j += stride) {
        const int slot = (j << 6) + lane_id();
        if (slot < listed) {
            qn[slot] = (uint32_t)(rbase + slot);
            gn[slot] = list[slot];
        }
    }
    if (wave / kShards == 0 && lane_id() == 0) *ctr_at(p.ctr, queue_ctr(0, 1, k)) = min(listed, p.region);
}

// Adds slot pid's pending shadow contributions to r in emission order (the
// environment's, then the light's: devicePrograms.cu:376-471 add them to
// radiance in that order as soon as each shadow ray is found unoccluded).
OXR_D void add_pending(const WaveParams& p, uint32_t pid, float4& r) {
    if (p.env_w > 0) {
        const float4 c = p.sh_env[3 * (size_t)pid + 2];
        if (c.w != 0.0f) {
            r.x = r.x + c.x;
            r.y = r.y + c.y;
            r.z = r.z + c.z;
        }
    }
    if (p.n_lights != 0) {
        const float4 c = p.sh_light[3 * (size_t)pid + 2];
        if (c.w != 0.0f) {
            r.x = r.x + c.x;
            r.y = r.y + c.y;
            r.z = r.z + c.z;
        }
    }
}

// regen(y): finish the paths that the previous shade listed (parity y^1) as
// ended with shadow rays pending -- their rays were traced by trace(y) --
// then start new samples in their slots (new-path region of queue y^1).
__global__ __launch_bounds__(256) void k_regen(WaveParams p, int y) {
    if (blockIdx.x == 0) {  // fetch counters of the next trace and shade, idle now
        if (threadIdx.x < kShards) *ctr_at(p.ctr, kCtrFetchTrace + threadIdx.x) = 0;
        for (int i = threadIdx.x; i < kShadeClasses * kShards; i += blockDim.x) *ctr_at(p.ctr, kCtrFetchClass + i) = 0;
    }
    const int k = wave_shard();
    const int n = *ctr_at(p.ctr, kCtrTerm + (y ^ 1) * kShards + k);
    const int chunks = (n + 63) >> 6;
    if (chunks == 0) return;
    int* fetch = ctr_at(p.ctr, kCtrFetchRegen + k);
    const uint32_t* term = p.term[y ^ 1] + (size_t)k * p.region;
    while (true) {
        const int j = wave_next(fetch);
        if (j >= chunks) break;
        const int slot = (j << 6) + lane_id();
        const bool have = slot < n;
        uint32_t pid = 0;
        if (have) {
            pid = term[slot];
            float4 r = p.path[4 * (size_t)pid + 3];
            uint32_t g = __float_as_uint(r.w);
            if (g != kNoSample) {
                add_pending(p, pid, r);
                p.s_rad[g] = make_float4(r.x, r.y, r.z, 0.0f);
            }
        }
        restart(p, k, y ^ 1, have, pid);
    }
}

// trace(y): the shadow rays the previous shade listed at parity y^1, then
// the extension rays of queue y, in one launch.  Each shard's work is one
// sequence of 64-ray chunks, shadow chunks first: a shadow ray leaving a
// surface visits several times the nodes of an extension ray, so the long
// rays start first and the short ones fill the waves that finish early.
// Extension rays: hit record per queue position and the position listed
// under its shade class (parity y).  Shadow rays: an occluded ray clears its
// record's pending flag.
template <bool COUNT, int S>
__global__ __launch_bounds__(kTraceBlock, 8) void k_trace(WaveParams p, int y) {
    uint32_t* gstack = p.gstack;
    int* overflow = p.overflow;
    __shared__ uint32_t lds[S * kTraceBlock];
    if (blockIdx.x == 0) {
        // lists this iteration refills, whose readers have finished
        const int t = threadIdx.x;
        if (t < kQueueRegions * kShards) *ctr_at(p.ctr, queue_ctr(y ^ 1, 0, 0) + t) = 0;
        if (t < kShards) {
            *ctr_at(p.ctr, kCtrTerm + y * kShards + t) = 0;
            *ctr_at(p.ctr, kCtrShEnv + y * kShards + t) = 0;
            *ctr_at(p.ctr, kCtrShLight + y * kShards + t) = 0;
            *ctr_at(p.ctr, kCtrFetchRegen + t) = 0;
        }
        for (int i = t; i < kShadeClasses * kShards; i += blockDim.x)
            *ctr_at(p.ctr, kCtrClass + ((y ^ 1) * kShadeClasses) * kShards + i) = 0;
        if (t == 0) {
            unsigned long long rays = 0, shadow = 0;
            for (int r = 0; r < 2 * kShards; r++) rays += (unsigned)*ctr_at(p.ctr, queue_ctr(y, 0, 0) + r);
            for (int r = 0; r < kShards; r++)
                shadow += (unsigned)*ctr_at(p.ctr, kCtrShEnv + (y ^ 1) * kShards + r) +
                          (unsigned)*ctr_at(p.ctr, kCtrShLight + (y ^ 1) * kShards + r);
            if (rays) atomicAdd(&p.traversal[2], rays);
            if (shadow) atomicAdd(&p.traversal[3], shadow);
            *ctr_at(p.ctr, kCtrTailMode) = (rays > 0 && (int64_t)rays <= p.tail_paths) ? 1 : 0;
        }
    }
    const int k = wave_shard();
    const ShardQueue sq(p.ctr, y, k);
    const int n_env = *ctr_at(p.ctr, kCtrShEnv + (y ^ 1) * kShards + k);
    const int n_light = *ctr_at(p.ctr, kCtrShLight + (y ^ 1) * kShards + k);
    const int c_env = (n_env + 63) >> 6, c_shadow = c_env + ((n_light + 63) >> 6);
    const int chunks = c_shadow + sq.chunks;
    if (chunks == 0) return;
    int* fetch = ctr_at(p.ctr, kCtrFetchTrace + k);
    const uint32_t* q = p.queue[y];
    const uint32_t* shl = p.sh_list[y ^ 1] + (size_t)k * p.region;
    Stack<S> st(lds, gstack, (int64_t)gridDim.x * blockDim.x);
    uint32_t nn = 0, nt = 0, sn = 0, stt = 0;
    uint32_t cn = 0, cnn = 0, cnt = 0;  // counting: rays the cover mask would resolve, their fetches / tests
    // counting: SIMD use of the traversal loop -- node visits of active lanes,
    // and 64 x the most visits of any lane per 64-ray chunk (the iterations
    // the wave runs)
    unsigned long long lane_visits = 0, wave_slots = 0;
    while (true) {
        const int j = wave_next(fetch);
        if (j >= chunks) break;
        if (j < c_shadow) {
            const bool light = j >= c_env;
            const int slot = ((light ? j - c_env : j) << 6) + lane_id();
            uint32_t vis = 0;
            if (slot < (light ? n_light : n_env)) {
                const uint32_t pid = shl[(light ? (size_t)p.pool : 0) + slot];
                float4* rec = (light ? p.sh_light : p.sh_env) + 3 * (size_t)pid;
                const float4 a = rec[0], b = rec[1];
                const uint32_t sn0 = sn;
                Hit h = traverse<true, COUNT>(p.bvh, xyz(a), xyz(b), kEps, a.w, st, sn, stt, overflow);
                if (h.leaf != kNoHit) reinterpret_cast<float*>(rec + 2)[3] = 0.0f;
                vis = (sn - sn0) >> 1;
            }
            if (COUNT) {
                lane_visits += vis;
                wave_slots += 64ull * wave_max_u32(vis);
            }
            continue;
        }
        const int64_t qp = sq.position(j - c_shadow, k, p.region);
        int cls = -1;
        uint32_t vis = 0;
        if (qp >= 0) {
            f3 org, dir;
            const bool fresh = is_fresh(p, qp);
            bool tally = false;
            if (fresh) {
                uint32_t rs;
                const uint32_t g = fresh_sample(p, y, qp);
                camera_ray(p, g, org, dir, rs);
                if (COUNT) tally = p.resolve == 2 && camera_ray_misses(p, g);
            } else {
                const uint32_t pid = q[qp];
                org = xyz(p.path[4 * (size_t)pid + 0]);
                dir = xyz(p.path[4 * (size_t)pid + 1]);
            }
            const uint32_t nn0 = nn, nt0 = nt;
            Hit h = traverse<false, COUNT>(p.bvh, org, dir, kEps, kFltMax, st, nn, nt, overflow);
            if (COUNT && tally) {
                cn++;
                cnn += nn - nn0;
                cnt += nt - nt0;
            }
            if (COUNT) vis = (nn - nn0) >> 1;
            p.hits[qp] = make_float4(h.t, __uint_as_float(h.leaf), h.u, h.v);
            cls = h.leaf != kNoHit ? 1 + p.mats[p.mat_ids[h.leaf]].type : fresh ? kClassCameraMiss : 0;
        }
        if (COUNT) {
            lane_visits += vis;
            wave_slots += 64ull * wave_max_u32(vis);
        }
        // list the queue position under its shade class: one append per
        // class present in the wave
        uint64_t left = __ballot(cls >= 0);
        while (left) {
            const int c = __shfl(cls, (int)__builtin_ctzll(left));
            const bool mine = cls == c;
            const int o = wave_append(ctr_at(p.ctr, kCtrClass + (y * kShadeClasses + c) * kShards + k), mine);
            if (mine) p.cls_queue[((size_t)c * kShards + k) * p.region + o] = (uint32_t)qp;
            left &= ~__ballot(mine);
        }
    }
    if (COUNT) {
        wave_count_add(&p.traversal[0], nn);
        wave_count_add(&p.traversal[1], nt);
        wave_count_add(&p.traversal[4], sn);
        wave_count_add(&p.traversal[5], stt);
        wave_count_add(&p.traversal[6], cn);
        wave_count_add(&p.traversal[7], cnn);
        wave_count_add(&p.traversal[8], cnt);
        if (lane_id() == 0) atomicAdd(&p.traversal[14], wave_slots);  // wave-uniform
        wave_count_add64(&p.traversal[13], lane_visits);
    }
}

// __closesthit__radiance's texture overrides (devicePrograms.cu:86-115): the
// texture coordinate interpolated as the reference does, then every texture
// the material names replaces its parameter; a normal map bends Ns.
OXR_D void apply_textures(const WaveParams& p, uint32_t leaf, float w, float u, float v, oxr_material& m, f3& Ns) {
    const float* t = p.uvs + 6 * (size_t)leaf;
    const float tu = w * t[0] + u * t[2] + v * t[4];
    const float tv = w * t[1] + u * t[3] + v * t[5];
    auto tex = [&](int id) {
        const int4 ti = p.tex_info[id];
        return bilinear_rgba8(p.texels + ti.x, ti.y, ti.z, tu, tv);
    };
    if (m.tex[0] >= 0) {
        const float4 c = tex(m.tex[0]);
        m.albedo[0] = c.x;
        m.albedo[1] = c.y;
        m.albedo[2] = c.z;
    }
    if (m.tex[1] >= 0) m.roughness = tex(m.tex[1]).x;
    if (m.tex[2] >= 0) m.anisotropy = tex(m.tex[2]).x;
    if (m.tex[3] >= 0) {
        const float4 c = tex(m.tex[3]);
        m.specular[0] = c.x;
        m.specular[1] = c.y;
        m.specular[2] = c.z;
    }
    if (m.tex[4] >= 0) m.metallic = tex(m.tex[4]).x;
    if (m.tex[5] >= 0) m.coat_roughness_u = tex(m.tex[5]).x;
    if (m.tex[6] >= 0) m.coat_roughness_v = tex(m.tex[6]).x;
    if (m.tex[7] >= 0) {
        const float4 c = tex(m.tex[7]);
        Ns = normal_from_tangent(Ns, f3(c.x, c.y, c.z));
    }
}

struct ShadowOut {
    bool emit;
    f3 o, d, c;
    float tmax;
};

// lightTrace lambda (devicePrograms.cu:192-211): the nearest analytic light
// closer than `closest` wins; light_pdf is overwritten by every light tried.
OXR_D bool light_trace(const WaveParams& p, f3 o, f3 d, int& mi, f3& Lr, float& lpdf, float& closest) {
    bool hit = false;
    for (int i = 0; i < p.n_lights; i++) {
        const oxr_light* l = p.lights + i;
        float ld = 0.0f;
        f3 Li = eval_light(l, o, d, closest, lpdf, ld);
        if (lpdf == 0.0f) continue;
        hit = true;
        Lr = Li;
        closest = ld;
        mi = l->medium;
    }
    return hit;
}

// One bounce of __raygen__renderFrame's loop body (devicePrograms.cu:270-496)
// for one path, after its radiance ray was traced.  Returns whether the path
// continues; NEE shadow rays are returned for the next trace with their
// contribution precomputed (it is added only if the ray is unoccluded, by the
// slot's next shade or by k_regen).
// MT: the material type of the hit surface, fixed by the shade class; < 0 for
// rays that left the scene.  MEDIA: the scene has participating media (the
// medium code is compiled only then).
// fresh: the path is at bounce 0 with sample id g_fresh and no record yet.
// The slot record is written only if the path continues; rad_out returns
// (radiance, sample id) for the caller.
template <int MT, bool MEDIA>
OXR_D bool shade_path(const WaveParams& p, const Lut& lut, uint32_t pid, float4 hit, bool fresh, uint32_t g_fresh,
                      ShadowOut& sh_env, ShadowOut& sh_light, float4& rad_out, bool& aov_default) {
    sh_env.emit = false;
    sh_light.emit = false;
    f3 org, dir, history, radiance;
    float pre_pdf;
    Rng rng;
    int bounce;
    uint32_t g;
    if (fresh) {
        camera_ray(p, g_fresh, org, dir, rng.s);
        pre_pdf = 0.0f;
        history = f3(1.0f);
        radiance = f3(0.0f);
        bounce = 0;
        g = g_fresh;
    } else {
        const float4 ro = p.path[4 * (size_t)pid + 0], rd = p.path[4 * (size_t)pid + 1];
        const float4 th = p.path[4 * (size_t)pid + 2];
        float4 rad4 = p.path[4 * (size_t)pid + 3];
        add_pending(p, pid, rad4);  // the last bounce's unoccluded shadow rays
        org = xyz(ro);
        dir = xyz(rd);
        pre_pdf = ro.w;
        rng.s = __float_as_uint(rd.w);
        history = xyz(th);
        bounce = __float_as_int(th.w);
        radiance = xyz(rad4);
        g = __float_as_uint(rad4.w);
    }
    f3 V = -dir, L = dir;
    float mult_trans_pdf = 1.0f;

    // __closesthit__radiance / __miss__radiance (devicePrograms.cu:71-159)
    const uint32_t leaf = MT < 0 ? kNoHit : __float_as_uint(hit.y);
    // the hit's material, with its texture overrides (a copy per hit, as
    // __closesthit__radiance's `Material material = sbtData.material`)
    oxr_material mat;
    Surf s;
    s.m = &mat;
    s.front = true;
    s.n = f3(0.0f);
    bool surface = false;
    f3 pos = org;
    float distance;
    int mi_in, mi_out;
    if (leaf != kNoHit) {
        const float* nb = p.normals + 9 * (size_t)leaf;
        const float4* tb = p.bvh.tris + 3 * (size_t)leaf;
        float u = hit.z, v = hit.w, w = 1.0f - u - v;
        f3 Ns = normalize(w * f3(nb[0], nb[1], nb[2]) + u * f3(nb[3], nb[4], nb[5]) + v * f3(nb[6], nb[7], nb[8]));
        mat = p.mats[p.mat_ids[leaf]];
        surface = true;
        if (p.textured) apply_textures(p, leaf, w, u, v, mat, Ns);
        f3 P = w * xyz(tb[0]) + u * xyz(tb[1]) + v * xyz(tb[2]);
        s.front = dot(dir, Ns) < 0.0f;                                      // Utils.h:70-73
        s.n = s.front ? Ns : -Ns;
        distance = length(P - org);
        pos = P;
        mi_in = s.m->in_medium;
        mi_out = s.m->out_medium;
        if (bounce == 0) {
            p.s_nrm[g] = make_float4(Ns.x, Ns.y, Ns.z, 1.0f);
            p.s_alb[g] = make_float4(s.m->albedo[0], s.m->albedo[1], s.m->albedo[2], 1.0f);
        }
    } else {
        distance = kFltMax;
        mi_in = mi_out = p.cam_medium;
    }
    float closest = distance;
    f3 light_radiance(0.0f);
    float light_pdf = 0.0f;
    int mi_light = 0;
    bool hitLight = light_trace(p, org, dir, mi_light, light_radiance, light_pdf, closest);
    if (hitLight) mi_in = mi_out = mi_light;
    // A camera ray that met neither a surface nor a light has the default
    // AOVs (albedo and normal 0, depth 1); the caller writes them only if the
    // sample does not end now, else it flags them in the radiance record.
    aov_default = bounce == 0 && leaf == kNoHit && closest == kFltMax;
    if (bounce == 0 && leaf == kNoHit && !aov_default) {
        p.s_nrm[g] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
        p.s_alb[g] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    }
    // depth AOV (no reference counterpart; a denoiser input): first surface
    // or light along the camera ray as d / (1 + d), 1 for a miss
    if (bounce == 0 && !aov_default) p.s_dep[g] = closest < kFltMax ? closest / (1.0f + closest) : 1.0f;

    bool scattered = false;
    int medium_index = -1;
    bool alive = true;
    if (MEDIA && p.n_media != 0) {                                          // devicePrograms.cu:276-338
        medium_index = (hitLight || s.front) ? mi_out : mi_in;
        if (medium_index != -1) {
            const oxr_medium* md = p.media + medium_index;
            float dist = 0.0f, trans_pdf = 0.0f;
            f3 T;
            scattered = sample_medium_distance(md, closest, dist, trans_pdf, T, rng);
            history = history * (T / f3(trans_pdf));
            mult_trans_pdf *= trans_pdf;
            if (scattered) {
                pos = org + dist * L;
                if (p.n_lights != 0) {
                    int idx = min(max((int)((float)p.n_lights * rng()), 0), p.n_lights - 1);
                    const oxr_light* l = p.lights + idx;
                    float ld = 0.0f;
                    float u0 = rng(), u1 = rng();
                    f3 sd;
                    light_radiance = sample_light(l, pos, f2{u0, u1}, sd, ld, light_pdf);
                    light_pdf *= (1.0f / (float)p.n_lights);
                    f3 tv; float tp; int tm;
                    float cd = ld - kEps;
                    if (!light_trace(p, pos, sd, tm, tv, tp, cd)) {
                        float tpdf2 = 0.0f, phase_pdf = 0.0f;
                        f3 T2 = eval_medium_distance(md, false, cd, tpdf2);
                        f3 att = eval_phase(md, V, sd, phase_pdf);
                        float mw = power_heuristic(light_pdf, phase_pdf * tpdf2);
                        sh_light.emit = true;
                        sh_light.o = pos;
                        sh_light.d = sd;
                        sh_light.tmax = ld - kEps;
                        sh_light.c = mw * (T2 / f3(tpdf2)) * light_radiance * att * history / f3(light_pdf);
                    }
                }
                float phase_pdf = 0.0f;
                f3 att = sample_phase(md, rng, V, L, phase_pdf);
                history = history * (att / f3(phase_pdf));
                pre_pdf = phase_pdf;
            }
        }
    }

    if (!scattered) {
        if (!hitLight && !surface && closest == kFltMax && medium_index != -1) {
            // a ray inside a medium that left the scene unscattered: the
            // reference would shade stale surface data with zero throughput
            alive = false;
        } else if (hitLight) {                                              // devicePrograms.cu:342-351
            float mw = 1.0f;
            if (bounce != 0) {
                pre_pdf *= mult_trans_pdf;
                mw = power_heuristic(pre_pdf, light_pdf);
            }
            radiance = radiance + mw * history * light_radiance;
            alive = false;
        } else if (closest == kFltMax && medium_index == -1) {              // devicePrograms.cu:354-373
            if (p.env_w > 0) {
                f2 uv = sphere_to_plane(dir);
                float4 e = bilinear(p.env_pdf, p.env_w, p.env_h, uv.x, uv.y);  // radiance and cache pdf
                light_radiance = f3(e.x, e.y, e.z);
                float mw = 1.0f;
                if (bounce != 0) {
                    // the environment pdf only enters the MIS weight, which is
                    // 1 for camera rays
                    light_pdf = (float)(p.env_w * p.env_h) / (2.0f * sqr(kPi));
                    light_pdf *= e.w;
                    pre_pdf *= mult_trans_pdf;
                    mw = power_heuristic(pre_pdf, light_pdf);
                }
                radiance = radiance + mw * history * light_radiance;
            }
            alive = false;
        } else if constexpr (MT < 0) {
            // unreachable: a ray that left the scene has closest == FLT_MAX
            // unless it hit a light, and both cases are handled above
            alive = false;
        } else {
            if (p.env_w > 0) {                                              // devicePrograms.cu:376-424
                float x0 = rng(), x1 = rng();
                float4 uvt = bilinear(p.env_cache, p.env_w, p.env_h, x0, x1);
                f2 uv{uvt.x, uvt.y};
                f3 sd = plane_to_sphere(uv);
                float lp = (float)(p.env_w * p.env_h) / (2.0f * sqr(kPi));
                const float4 e = bilinear(p.env_pdf, p.env_w, p.env_h, uv.x, uv.y);  // radiance and cache pdf
                lp *= e.w;
                f3 lr(e.x, e.y, e.z);
                f3 tv; float tp; int tm;
                float cd = kFltMax;
                if (!light_trace(p, pos, sd, tm, tv, tp, cd)) {
                    float bpdf = 0.0f;
                    f3 f = eval_material_t<MT>(s, lut, V, sd, bpdf);
                    float ct = fabsf(dot(s.n, sd));
                    float mw = power_heuristic(lp, bpdf);
                    f3 thh(1.0f);
                    if (medium_index != -1) {
                        float tpdf2 = 0.0f;
                        f3 T2 = eval_medium_distance(p.media + medium_index, false, cd, tpdf2);
                        thh = T2 / f3(tpdf2);
                        mw = power_heuristic(lp, bpdf * tpdf2);
                    }
                    sh_env.emit = true;
                    sh_env.o = pos;
                    sh_env.d = sd;
                    sh_env.tmax = kFltMax;
                    sh_env.c = mw * thh * lr * f * ct * history / f3(lp);
                }
            }
            if (p.n_lights != 0) {                                          // devicePrograms.cu:426-471
                int idx = min(max((int)((float)p.n_lights * rng()), 0), p.n_lights - 1);
                const oxr_light* l = p.lights + idx;
                float ld = 0.0f;
                float u0 = rng(), u1 = rng();
                f3 sd;
                f3 lr = sample_light(l, pos, f2{u0, u1}, sd, ld, light_pdf);
                light_pdf *= (1.0f / (float)p.n_lights);
                f3 tv; float tp; int tm;
                float cd = ld - kEps;
                if (!light_trace(p, pos, sd, tm, tv, tp, cd)) {
                    float bpdf = 0.0f;
                    f3 f = eval_material_t<MT>(s, lut, V, sd, bpdf);
                    float ct = fabsf(dot(s.n, sd));
                    float mw = power_heuristic(light_pdf, bpdf);
                    f3 thh(1.0f);
                    if (medium_index != -1) {
                        float tpdf2 = 0.0f;
                        f3 T2 = eval_medium_distance(p.media + medium_index, false, cd, tpdf2);
                        thh = T2 / f3(tpdf2);
                        mw = power_heuristic(light_pdf, bpdf * tpdf2);
                    }
                    sh_light.emit = true;
                    sh_light.o = pos;
                    sh_light.d = sd;
                    sh_light.tmax = ld - kEps;
                    sh_light.c = mw * thh * lr * f * ct * history / f3(light_pdf);
                }
            }
            float bpdf = 0.0f;                                              // devicePrograms.cu:474-477
            f3 f = sample_material_t<MT>(s, lut, rng, V, L, bpdf);
            float ct = fabsf(dot(s.n, L));
            history = history * (f * ct / f3(bpdf));
            pre_pdf = bpdf;
        }
    }
    if (alive) {
        if (bounce >= p.rr_start) {                                         // devicePrograms.cu:481-485
            float prr = fminf((history.x + history.y + history.z) / 3.0f, 0.95f);
            if (rng() > prr) alive = false;
            else history = history / f3(prr);
        }
        if (alive && any_nan(history)) alive = false;                       // devicePrograms.cu:487-489
    }
    rad_out = make_float4(radiance.x, radiance.y, radiance.z, __uint_as_float(g));
    const bool cont = alive && bounce + 1 < p.max_depth;
    if (cont) {
        p.path[4 * (size_t)pid + 0] = make_float4(pos.x, pos.y, pos.z, pre_pdf);
        p.path[4 * (size_t)pid + 1] = make_float4(L.x, L.y, L.z, __uint_as_float(rng.s));
        p.path[4 * (size_t)pid + 2] = make_float4(history.x, history.y, history.z, __int_as_float(bounce + 1));
        p.path[4 * (size_t)pid + 3] = rad_out;
    }
    return cont;
}

// Writes slot pid's shadow record of one kind and lists the slot for the
// next trace.  `pending`: the path is not final, so its next shade or k_regen
// reads the record's flag, which must then say whether a ray was emitted.
// All lanes of the wave must call it.
OXR_D void emit_shadow(float4* recs, uint32_t* list, int* ctr, const ShadowOut& s, uint32_t pid, bool pending) {
    const int slot = wave_append(ctr, s.emit);
    if (slot >= 0) list[slot] = pid;
    float4* rec = recs + 3 * (size_t)pid;
    if (s.emit) {
        rec[0] = make_float4(s.o.x, s.o.y, s.o.z, s.tmax);
        rec[1] = make_float4(s.d.x, s.d.y, s.d.z, 0.0f);
        rec[2] = make_float4(s.c.x, s.c.y, s.c.z, 1.0f);
    } else if (pending) {
        reinterpret_cast<float*>(rec + 2)[3] = 0.0f;
    }
}

// One shade class: the paths of queue x whose extension ray left the scene
// (CLS 0, past bounce 0; CLS kClassCameraMiss, camera rays) or hit a surface
// of material type CLS - 1, listed by k_trace in cls_queue.  Each class
// compiles only its own BSDF, and the two miss classes know whether the path
// is at bounce 0, so the common cases run at high occupancy.
template <int CLS, bool MEDIA>
__global__ __launch_bounds__(256) void k_shade(WaveParams p, int x) {
    constexpr int MT = (CLS == 0 || CLS == kClassCameraMiss) ? -1 : CLS - 1;
    if (*ctr_at(p.ctr, kCtrTailMode)) return;  // k_tail shades this iteration
    const int k = wave_shard();
    const int n = *ctr_at(p.ctr, kCtrClass + (x * kShadeClasses + CLS) * kShards + k);
    const int chunks = (n + 63) >> 6;
    if (chunks == 0) return;
    int* fetch = ctr_at(p.ctr, kCtrFetchClass + CLS * kShards + k);
    const uint32_t* cq = p.cls_queue + ((size_t)CLS * kShards + k) * p.region;
    const uint32_t* q = p.queue[x];
    const size_t rbase = (size_t)k * p.region;
    uint32_t* qo = p.queue[x ^ 1] + rbase;  // survivors region of shard k
    int* qo_count = ctr_at(p.ctr, queue_ctr(x ^ 1, 0, k));
    uint32_t* term = p.term[x] + rbase;
    int* term_count = ctr_at(p.ctr, kCtrTerm + x * kShards + k);
    uint32_t* le = p.sh_list[x] + rbase;
    uint32_t* ll = p.sh_list[x] + p.pool + rbase;
    int* se_count = ctr_at(p.ctr, kCtrShEnv + x * kShards + k);
    int* sl_count = ctr_at(p.ctr, kCtrShLight + x * kShards + k);
    Lut lut{p.lut_bsdf, p.lut_albedo};
    while (true) {
        const int j = wave_next(fetch);
        if (j >= chunks) break;
        const int slot = (j << 6) + lane_id();
        const bool active = slot < n;
        uint32_t pid = 0;
        bool cont = false;
        ShadowOut sev, slv;
        sev.emit = slv.emit = false;
        float4 rad = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        bool aov_default = false;
        if (active) {
            const uint32_t qp = cq[slot];
            pid = q[qp];
            const bool fresh = CLS == kClassCameraMiss ? true : CLS == 0 ? false : is_fresh(p, qp);
            const float4 hit = MT < 0 ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : p.hits[qp];  // unused by misses
            cont = shade_path<MT, MEDIA>(p, lut, pid, hit, fresh, fresh ? fresh_sample(p, x, qp) : 0u, sev, slv,
                                         rad, aov_default);
        }
        // A path that ended with no shadow ray in flight is final now: its
        // radiance goes to its sample and the slot starts the next sample at
        // once.  One whose shadow rays are still to be traced keeps its
        // radiance in the slot record; k_regen adds them and finishes it.
        const bool ended = active && !cont;
        const bool final = ended && !sev.emit && !slv.emit;
        const uint32_t g = __float_as_uint(rad.w);
        // radiance record: w = 1 says the sample's AOVs are the defaults and
        // were not written
        if (final) p.s_rad[g] = make_float4(rad.x, rad.y, rad.z, aov_default ? 1.0f : 0.0f);
        if (aov_default && !final) {
            p.s_nrm[g] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
            p.s_alb[g] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
            p.s_dep[g] = 1.0f;
        }
        if (ended && !final) p.path[4 * (size_t)pid + 3] = rad;
        int o = wave_append(qo_count, cont);
        if (o >= 0) qo[o] = pid;
        o = wave_append(term_count, ended && !final);
        if (o >= 0) term[o] = pid;
        const bool pending = active && !final;
        if (p.env_w > 0) emit_shadow(p.sh_env, le, se_count, sev, pid, pending);
        if (p.n_lights != 0) emit_shadow(p.sh_light, ll, sl_count, slv, pid, pending);
        restart(p, k, x ^ 1, final, pid);
    }
}

// A shadow ray in the tail kernel: traced at once and recorded as the slot's
// next shade (or the finish) reads it in add_pending.
template <bool COUNT, class StackT>
OXR_D void tail_shadow(const WaveParams& p, float4* recs, const ShadowOut& s, uint32_t pid, bool pending,
                       StackT& st, uint32_t& sn, uint32_t& stt, uint32_t& rays) {
    float4* rec = recs + 3 * (size_t)pid;
    if (s.emit) {
        rays++;
        const Hit h = traverse<true, COUNT>(p.bvh, s.o, s.d, kEps, s.tmax, st, sn, stt, p.overflow);
        rec[2] = make_float4(s.c.x, s.c.y, s.c.z, h.leaf == kNoHit ? 1.0f : 0.0f);
    } else if (pending) {
        reinterpret_cast<float*>(rec + 2)[3] = 0.0f;
    }
}

// Tail mode (k_trace sets the flag when queue x holds few paths): the paths
// listed under every shade class of parity x are carried to their end here,
// each lane one path at a time: shade (run-time material dispatch), trace its
// shadow rays at once, finish it or trace its next ray, and so on; a lane
// whose path ended takes the next listed path (one atomic per wave), so the
// launch lasts about as long as its longest path rather than the sum of the
// longest path of each 64-path chunk a wave is given.  A finished path's slot
// starts a new sample if any is left (queue x^1).  The operations and their
// order per path are those of the wavefront, so the result is the same.
// Launched on the trace grid: the traversal stacks are that grid's.
template <bool COUNT, bool MEDIA>
__global__ __launch_bounds__(kTraceBlock) void k_tail(WaveParams p, int x) {
    constexpr int S = 16;
    if (!*ctr_at(p.ctr, kCtrTailMode)) return;
    __shared__ uint32_t lds[S * kTraceBlock];
    const int k = wave_shard();
    int first[kShadeClasses + 1];  // the classes' lists back to back, in paths
    first[0] = 0;
    for (int c = 0; c < kShadeClasses; c++)
        first[c + 1] = first[c] + *ctr_at(p.ctr, kCtrClass + (x * kShadeClasses + c) * kShards + k);
    const int total = first[kShadeClasses];
    if (total == 0) return;
    int* fetch = ctr_at(p.ctr, kCtrFetchClass + k);  // class 0's counter (in paths): the class kernels are idle
    const uint32_t* q = p.queue[x];
    Stack<S> st(lds, p.gstack, (int64_t)gridDim.x * blockDim.x);
    Lut lut{p.lut_bsdf, p.lut_albedo};
    uint32_t nn = 0, nt = 0, sn = 0, stt = 0, ext_rays = 0, sh_rays = 0;
    bool have = false, fresh = false, listed_left = true;
    uint32_t pid = 0, g_fresh = 0;
    float4 hit = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    while (true) {
        // lanes without a path take the next listed ones
        const uint64_t need = __ballot(!have);
        if (need != 0 && listed_left) {
            int base = 0;
            if (lane_id() == 0) base = atomicAdd(fetch, (int)__popcll(need));
            base = __shfl(base, 0);
            listed_left = base + (int)__popcll(need) < total;
            const int e = base + (int)__popcll(need & lanes_below());
            if (!have && e < total) {
                int c = 0;
                while (e >= first[c + 1]) c++;
                const uint32_t qp = p.cls_queue[((size_t)c * kShards + k) * p.region + (e - first[c])];
                pid = q[qp];
                fresh = is_fresh(p, qp);
                g_fresh = fresh ? fresh_sample(p, x, qp) : 0u;
                hit = (c == 0 || c == kClassCameraMiss) ? make_float4(kFltMax, __uint_as_float(kNoHit), 0.0f, 0.0f)
                                                        : p.hits[qp];
                have = true;
            }
        }
        if (__ballot(have) == 0) break;
        bool ended = false;
        if (have) {
            ShadowOut sev, slv;
            sev.emit = slv.emit = false;
            float4 rad;
            bool aov_default = false;
            const bool cont = shade_path<kMatAny, MEDIA>(p, lut, pid, hit, fresh, g_fresh, sev, slv, rad, aov_default);
            fresh = false;
            const bool final = !cont && !sev.emit && !slv.emit;
            if (p.env_w > 0) tail_shadow<COUNT>(p, p.sh_env, sev, pid, !final, st, sn, stt, sh_rays);
            if (p.n_lights != 0) tail_shadow<COUNT>(p, p.sh_light, slv, pid, !final, st, sn, stt, sh_rays);
            if (!cont) {
                const uint32_t g = __float_as_uint(rad.w);
                if (final) {
                    p.s_rad[g] = make_float4(rad.x, rad.y, rad.z, aov_default ? 1.0f : 0.0f);
                } else {
                    float4 r = rad;
                    add_pending(p, pid, r);  // as k_regen finishes a path
                    p.s_rad[g] = make_float4(r.x, r.y, r.z, 0.0f);
                }
                if (aov_default && !final) {
                    p.s_nrm[g] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
                    p.s_alb[g] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
                    p.s_dep[g] = 1.0f;
                }
                have = false;
                ended = true;
            } else {
                ext_rays++;
                const f3 org = xyz(p.path[4 * (size_t)pid + 0]), dir = xyz(p.path[4 * (size_t)pid + 1]);
                const Hit h = traverse<false, COUNT>(p.bvh, org, dir, kEps, kFltMax, st, nn, nt, p.overflow);
                hit = make_float4(h.t, __uint_as_float(h.leaf), h.u, h.v);
            }
        }
        restart(p, k, x ^ 1, ended, pid);
    }
    wave_count_add(&p.traversal[2], ext_rays);
    wave_count_add(&p.traversal[3], sh_rays);
    wave_count_add(&p.traversal[11], ext_rays);
    wave_count_add(&p.traversal[12], sh_rays);
    if (COUNT) {
        wave_count_add(&p.traversal[0], nn);
        wave_count_add(&p.traversal[1], nt);
        wave_count_add(&p.traversal[4], sn);
        wave_count_add(&p.traversal[5], stt);
        wave_count_add(&p.traversal[9], nn + sn);
        wave_count_add(&p.traversal[10], nt + stt);
    }
}

// devicePrograms.cu:499-531 with numPixelSamples = 1: a NaN radiance drops
// the frame's sample (all three outputs become 0 for that frame).
__global__ void k_accumulate(WaveParams p, int64_t npix_total) {
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= npix_total) return;
    const ReqDesc& q = p.reqs[find_req<true>(p.reqs, p.nreq, t)];
    const int64_t pix = t - q.pix_begin;
    const int64_t npix = (int64_t)q.width * q.height;
    float4 c = q.color[pix], a = q.albedo[pix], nr = q.normal[pix];
    float dz = q.depth[pix];
    for (int s = 0; s < q.frames; s++) {
        size_t pid = (size_t)(q.path_begin + s * npix + pix);
        int frame = q.frame0 + s;
        const float4 r = p.s_rad[pid];
        float4 al = make_float4(0.0f, 0.0f, 0.0f, 1.0f), no = al;
        float z = 1.0f;
        if (r.w == 0.0f) {  // AOVs written by the shade, else the defaults
            al = p.s_alb[pid];
            no = p.s_nrm[pid];
            z = p.s_dep[pid];
        }
        float4 x = make_float4(r.x, r.y, r.z, 1.0f);
        if (isnan(r.x) || isnan(r.y) || isnan(r.z)) {
            x = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
            al = x;
            no = x;
            z = 0.0f;
        }
        if (q.path_radiance) {
            float* pr = q.path_radiance + 3 * ((size_t)(q.path_frame_offset + s) * npix + pix);
            pr[0] = r.x; pr[1] = r.y; pr[2] = r.z;
        }
        if (frame > 0) {
            float f = (float)frame, g = (float)frame + 1.0f;
            x.x = (x.x + f * c.x) / g; x.y = (x.y + f * c.y) / g; x.z = (x.z + f * c.z) / g;
            // Where sample and mean are both 0 or both 1 the quotient is
            // known exactly -- (0 + f 0) / g = +0, and (1 + f 1) / g = 1, the
            // numerator and g being the same sum rounded alike -- so those
            // divisions are skipped: the colour's w (always 1), and the AOVs
            // of a default-AOV sample (albedo and normal (0, 0, 0, 1), depth
            // 1) on default means, as every sky pixel has them.
            if (!(x.w == 1.0f && c.w == 1.0f)) x.w = (x.w + f * c.w) / g;
            const bool unit = a.x == 0.0f && a.y == 0.0f && a.z == 0.0f && a.w == 1.0f && nr.x == 0.0f &&
                              nr.y == 0.0f && nr.z == 0.0f && nr.w == 1.0f && dz == 1.0f;
            // a sample with default AOVs that was not dropped (a dropped
            // one has depth 0): al, no and z are the constants
            const bool dflt = r.w != 0.0f && z == 1.0f;
            if (!(unit && dflt)) {
                al.x = (al.x + f * a.x) / g; al.y = (al.y + f * a.y) / g; al.z = (al.z + f * a.z) / g; al.w = (al.w + f * a.w) / g;
                no.x = (no.x + f * nr.x) / g; no.y = (no.y + f * nr.y) / g; no.z = (no.z + f * nr.z) / g; no.w = (no.w + f * nr.w) / g;
                z = (z + f * dz) / g;
            }
        }
        c = x;
        a = al;
        nr = no;
        dz = z;
    }
    q.color[pix] = c;
    q.albedo[pix] = a;
    q.normal[pix] = nr;
    q.depth[pix] = dz;
}

// The scene's cover set (one wave): starting from the root's two children,
// replace every internal node by its children, level by level, while the
// list stays within kCoverBoxes.  Entries are (lo.xyz, ref), (hi.xyz, 0) with
// the exact boxes of the Karras layout; an expanded entry's children take its
// place in the list (left, then right), so entry i of a level lands at i
// plus the number of internal entries before it.  The levels are built in
// LDS and the last one copied out.
__global__ __launch_bounds__(64) void k_cover_set(int64_t nt, const uint32_t* left, const uint32_t* right,
                                                  const float* lo, const float* hi, float4* boxes, int* count) {
    __shared__ float4 lv[2][2 * kCoverBoxes];
    const int lane = threadIdx.x;
    if (nt <= 0) {
        if (lane == 0) *count = 0;
        return;
    }
    const int64_t ni = nt - 1;
    auto put = [&](uint32_t ref, float4* out) {
        const int64_t idx = (ref & kLeafBit) ? ni + (ref & ~kLeafBit) : ref;
        out[0] = make_float4(lo[3 * idx], lo[3 * idx + 1], lo[3 * idx + 2], __uint_as_float(ref));
        out[1] = make_float4(hi[3 * idx], hi[3 * idx + 1], hi[3 * idx + 2], 0.0f);
    };
    if (nt == 1) {
        if (lane == 0) {
            put(kLeafBit, boxes);
            *count = 1;
        }
        return;
    }
    if (lane == 0) put(left[0], lv[0]);
    if (lane == 1) put(right[0], lv[0] + 2);
    int n = 2, cur = 0;
    while (true) {
        __syncthreads();
        int internal = 0;
        for (int i = lane; i < n; i += 64) internal += (__float_as_uint(lv[cur][2 * i].w) & kLeafBit) ? 0 : 1;
        for (int o = 32; o > 0; o >>= 1) internal += __shfl_xor(internal, o);
        if (internal == 0 || n + internal > kCoverBoxes) break;
        int carry = 0;  // internal entries before this chunk
        for (int i0 = 0; i0 < n; i0 += 64) {
            const int i = i0 + lane;
            const bool in = i < n;
            const float4 blo = in ? lv[cur][2 * i] : make_float4(0, 0, 0, 0), bhi = in ? lv[cur][2 * i + 1] : blo;
            const uint32_t ref = __float_as_uint(blo.w);
            const bool inner = in && !(ref & kLeafBit);
            const uint64_t b = __ballot(inner);
            const int at = i + carry + __popcll(b & ((1ull << lane) - 1));
            if (inner) {
                put(left[ref], lv[cur ^ 1] + 2 * at);
                put(right[ref], lv[cur ^ 1] + 2 * at + 2);
            } else if (in) {
                lv[cur ^ 1][2 * at] = blo;
                lv[cur ^ 1][2 * at + 1] = bhi;
            }
            carry += __popcll(b);
        }
        n += internal;
        cur ^= 1;
    }
    for (int i = lane; i < 2 * n; i += 64) boxes[i] = lv[cur][i];
    if (lane == 0) *count = n;
}

// Cover mask of one request (one block).  A camera ray of pixel (ix, iy)
// passes through the image-plane point X = ix + 0.5 + jx, Y = iy + 0.5 + jy
// (pixel units, |j| below the largest filter offset): its direction is
// cam_dir + (X / W - 0.5) cam_h + (Y / H - 0.5) cam_v.  Solving P - cam_pos =
// a cam_dir + b cam_h + c cam_v for a box corner P (in double) gives its
// image point X = (b / a + 0.5) W, Y = (c / a + 0.5) H when a > 0; every
// point a camera ray reaches has a > 0, and the image of a box in front of
// the camera lies in the rectangle of its corners' images.  A box wholly
// behind the camera is skipped; one that straddles the camera plane marks
// every tile.
__global__ __launch_bounds__(256) void k_cover_mask(const ReqDesc* reqs, const float4* boxes, const int* count) {
    const ReqDesc& q = reqs[blockIdx.x];
    if (!q.cover) return;
    uint8_t* cover = const_cast<uint8_t*>(q.cover);
    const int tw = q.tiles_w, th = cover_tiles(q.height);
    const int ntiles = tw * th;
    __shared__ int all;
    if (threadIdx.x == 0) all = 0;
    for (int i = threadIdx.x; i < ntiles; i += blockDim.x) cover[i] = 0;
    __syncthreads();
    // inverse of the camera basis [cam_dir cam_h cam_v] (columns)
    double m[3][3];
    for (int r = 0; r < 3; r++) {
        m[r][0] = q.cam_dir[r];
        m[r][1] = q.cam_h[r];
        m[r][2] = q.cam_v[r];
    }
    const double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                       m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                       m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    const double scale = fabs(m[0][0]) + fabs(m[1][0]) + fabs(m[2][0]) + fabs(m[0][1]) + fabs(m[1][1]) +
                         fabs(m[2][1]) + fabs(m[0][2]) + fabs(m[1][2]) + fabs(m[2][2]);
    const bool singular = !(fabs(det) > 1e-12 * scale * scale * scale);
    double inv[3][3];
    if (!singular) {
        inv[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) / det;
        inv[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) / det;
        inv[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) / det;
        inv[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) / det;
        inv[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) / det;
        inv[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) / det;
        inv[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) / det;
        inv[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) / det;
        inv[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) / det;
    }
    const int nb = count ? *count : 0;
    if (!boxes || !count) all = 1;  // no cover set: every tile traced
    for (int bi = threadIdx.x; bi < nb; bi += blockDim.x) {
        if (singular) {
            all = 1;
            break;
        }
        const float4 lo = boxes[2 * bi], hi = boxes[2 * bi + 1];
        double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
        int front = 0, behind = 0;
        for (int c = 0; c < 8; c++) {
            const double d[3] = {(double)((c & 1) ? hi.x : lo.x) - q.cam_pos[0],
                                 (double)((c & 2) ? hi.y : lo.y) - q.cam_pos[1],
                                 (double)((c & 4) ? hi.z : lo.z) - q.cam_pos[2]};
            const double a = inv[0][0] * d[0] + inv[0][1] * d[1] + inv[0][2] * d[2];
            const double bb = inv[1][0] * d[0] + inv[1][1] * d[1] + inv[1][2] * d[2];
            const double cc = inv[2][0] * d[0] + inv[2][1] * d[1] + inv[2][2] * d[2];
            if (a > 1e-9) {
                front++;
                const double X = (bb / a + 0.5) * q.width, Y = (cc / a + 0.5) * q.height;
                x0 = fmin(x0, X);
                x1 = fmax(x1, X);
                y0 = fmin(y0, Y);
                y1 = fmax(y1, Y);
            } else if (a <= 0.0) {
                behind++;
            }
        }
        if (front == 0 && behind == 8) continue;  // no camera ray reaches it
        if (front < 8) {
            all = 1;
            continue;
        }
        // pixels ix whose samples may fall in [x0, x1]: ix + 0.5 + j in it, |j| < margin
        const double fx0 = floor((x0 - kCoverMargin - 0.5) / kCoverTile), fx1 = floor((x1 + kCoverMargin - 0.5) / kCoverTile);
        const double fy0 = floor((y0 - kCoverMargin - 0.5) / kCoverTile), fy1 = floor((y1 + kCoverMargin - 0.5) / kCoverTile);
        if (fx1 < 0.0 || fy1 < 0.0 || fx0 > tw - 1 || fy0 > th - 1) continue;
        const int tx0 = (int)fmax(fx0, 0.0), tx1 = (int)fmin(fx1, (double)(tw - 1));
        const int ty0 = (int)fmax(fy0, 0.0), ty1 = (int)fmin(fy1, (double)(th - 1));
        for (int ty = ty0; ty <= ty1; ty++)
            for (int tx = tx0; tx <= tx1; tx++) cover[ty * tw + tx] = 1;
    }
    __syncthreads();
    if (all)
        for (int i = threadIdx.x; i < ntiles; i += blockDim.x) cover[i] = 1;
}

inline int blocks(int64_t n, int b) { return (int)((n + b - 1) / b); }

}  // namespace

hipError_t build_cover_set(const BvhBuffers& b, int64_t n, float4* boxes, int* count, hipStream_t s) {
    k_cover_set<<<1, 64, 0, s>>>(n, b.left, b.right, b.lo, b.hi, boxes, count);
    return hipGetLastError();
}

hipError_t build_cover_masks(const ReqDesc* reqs, int nreq, int max_tiles, const float4* boxes, const int* count,
                             hipStream_t s) {
    (void)max_tiles;
    if (nreq <= 0) return hipSuccess;
    k_cover_mask<<<nreq, 256, 0, s>>>(reqs, boxes, count);
    return hipGetLastError();
}

hipError_t trace_rays(const BvhView& b, const float* rays, int64_t n, int any, float4* hits, uint8_t* occ,
                      uint32_t* counts, uint32_t* gstack, int grid, int* overflow, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (any) k_trace_rays<true><<<grid, kTraceBlock, 0, s>>>(b, rays, n, hits, occ, counts, gstack, overflow);
    else k_trace_rays<false><<<grid, kTraceBlock, 0, s>>>(b, rays, n, hits, occ, counts, gstack, overflow);
    return hipGetLastError();
}

// ------------------------------------------------------------ ray queries that allocate nothing
// Scratch of oxr_trace_rays_noalloc, resident for the life of the process on
// every device the library runs on: the traversal stack spill of
// kNoAllocTraceGrid blocks of k_trace_rays (32 entries per lane in LDS, the
// other kStackTotal - 32 here: entry i of grid lane g at g + i * lanes) and,
// after it, the overflow word.  Host code never passes `scratch` itself: it
// uses the device address noalloc_scratch() obtained with
// hipGetSymbolAddress, one per device.
static __device__ __align__(256) uint32_t scratch[kNoAllocScratchWords];

namespace {
std::mutex g_noalloc_mu;
uint32_t* g_noalloc_ptr[kMaxDevices] = {};

__global__ void k_clear_word(uint32_t* w) {
    if (threadIdx.x == 0) *w = 0u;
}
}  // namespace

hipError_t noalloc_scratch(int device, uint32_t** out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[device]) {
        hipError_t e = hipSetDevice(device);   // the array's copy on this device
        if (e != hipSuccess) return e;
        uint32_t* p = nullptr;
        e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_noalloc_ptr[device] = p;
    }
    *out = g_noalloc_ptr[device];
    return hipSuccess;
}

hipError_t trace_rays_noalloc(const BvhView& b, const float* rays, int64_t n, int any, float4* hits, uint8_t* occ,
                              uint32_t* scratch_ptr, hipStream_t s) {
    uint32_t* overflow = scratch_ptr + kNoAllocSpillWords;
    k_clear_word<<<1, 64, 0, s>>>(overflow);   // a fill kernel, not a memset: callers may capture this
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || n <= 0) return e;
    const int64_t want = (n + kTraceBlock - 1) / kTraceBlock;
    const int grid = (int)std::min<int64_t>(want, kNoAllocTraceGrid);   // the spill area holds this many blocks
    if (any) k_trace_rays<true><<<grid, kTraceBlock, 0, s>>>(b, rays, n, hits, occ, nullptr, scratch_ptr, (int*)overflow);
    else k_trace_rays<false><<<grid, kTraceBlock, 0, s>>>(b, rays, n, hits, occ, nullptr, scratch_ptr, (int*)overflow);
    return hipGetLastError();
}

size_t env_scratch_floats(int w, int h) { return (size_t)3 * w * h + 2 * (size_t)w + 16; }

// The environment radiance and the cache's pdf channel side by side: the
// miss shade and next-event estimation read both at one coordinate, so one
// bilinear footprint of 16-B texels serves both (the per-channel arithmetic
// is the same as two separate lookups).
__global__ void k_env_pack(const float4* __restrict__ env, const float4* __restrict__ cache, int64_t m,
                           float4* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) out[i] = make_float4(env[i].x, env[i].y, env[i].z, cache[i].z);
}

hipError_t env_cache_build(const float4* env, int w, int h, float lum_sum, float4* cache, float4* env_pdf,
                           float* scratch, hipStream_t s) {
    int64_t m = (int64_t)w * h;
    float* lum = scratch;
    float* pdf = lum + m;
    float* cdfy = pdf + m;
    float* marg = cdfy + m;
    float* cdfx = marg + w;
    k_env_lum<<<blocks(m, 256), 256, 0, s>>>(env, m, lum);
    k_env_columns<<<blocks(w, 64), 64, 0, s>>>(lum, lum_sum, w, h, pdf, marg, cdfy);
    k_env_cdfx<<<1, 64, 0, s>>>(marg, w, cdfx);
    k_env_search<<<blocks(m, 256), 256, 0, s>>>(pdf, cdfx, cdfy, w, h, cache);
    k_env_pack<<<blocks(m, 256), 256, 0, s>>>(env, cache, m, env_pdf);
    return hipGetLastError();
}

hipError_t tonemap(const float4* in, uint32_t* out, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_tonemap<<<blocks(n, 256), 256, 0, s>>>(in, out, n);
    return hipGetLastError();
}

hipError_t material_batch(const oxr_material* m, const float* lb, const float* la, int mode, const float* in,
                          int64_t n, float* out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_material_batch<<<blocks(n, 128), 128, 0, s>>>(m, Lut{lb, la}, mode, in, n, out);
    return hipGetLastError();
}

hipError_t gather_sorted_attributes(const float* nrm, const uint32_t* mid, const float* uv, const uint32_t* order,
                                    int64_t n, float* nrm_s, uint32_t* mid_s, float* uv_s, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_gather<<<blocks(n, 256), 256, 0, s>>>(nrm, mid, uv_s ? uv : nullptr, order, n, nrm_s, mid_s, uv_s);
    return hipGetLastError();
}

hipError_t wave_window_init(const WaveParams& p, bool set_req0, int frame0, int path_frame_offset, int frames,
                            int64_t total_samples, hipStream_t s) {
    const int64_t n = (int64_t)ctr_ints();
    k_window_init<<<blocks(n, 256), 256, 0, s>>>(p, const_cast<ReqDesc*>(p.reqs), set_req0 ? 1 : 0, frame0,
                                                 path_frame_offset, frames, (int)total_samples);
    return hipGetLastError();
}

hipError_t wave_regen(const WaveParams& p, int y, int grid, hipStream_t s) {
    k_regen<<<grid, 256, 0, s>>>(p, y);
    return hipGetLastError();
}

hipError_t wave_fill(const WaveParams& p, int grid, hipStream_t s) {
    k_fill<<<grid, 256, 0, s>>>(p);
    return hipGetLastError();
}

hipError_t wave_sky_list(const WaveParams& p, int grid, hipStream_t s) {
    k_sky_list<<<grid, 256, 0, s>>>(p);
    return hipGetLastError();
}

hipError_t wave_sky_shade(const WaveParams& p, int grid, hipStream_t s) {
    k_sky_shade<<<grid, 256, 0, s>>>(p);
    return hipGetLastError();
}

template <bool C>
void launch_trace(const WaveParams& p, int y, int grid, hipStream_t s) {
    if (p.stack_lds <= 8) k_trace<C, 8><<<grid, kTraceBlock, 0, s>>>(p, y);
    else if (p.stack_lds <= 16) k_trace<C, 16><<<grid, kTraceBlock, 0, s>>>(p, y);
    else k_trace<C, 32><<<grid, kTraceBlock, 0, s>>>(p, y);
}

hipError_t wave_trace(const WaveParams& p, int y, int grid, bool count, hipStream_t s) {
    if (count) launch_trace<true>(p, y, grid, s);
    else launch_trace<false>(p, y, grid, s);
    return hipGetLastError();
}

template <bool MEDIA>
void launch_shade(const WaveParams& p, int x, int grid, hipStream_t s) {
    const uint32_t m = p.class_mask;
    if (m & 1u) k_shade<0, MEDIA><<<grid, 256, 0, s>>>(p, x);
    if (m & 2u) k_shade<1, MEDIA><<<grid, 256, 0, s>>>(p, x);
    if (m & 4u) k_shade<2, MEDIA><<<grid, 256, 0, s>>>(p, x);
    if (m & 8u) k_shade<3, MEDIA><<<grid, 256, 0, s>>>(p, x);
    if (m & 16u) k_shade<4, MEDIA><<<grid, 256, 0, s>>>(p, x);
    if (m & 32u) k_shade<5, MEDIA><<<grid, 256, 0, s>>>(p, x);
    if (m & 64u) k_shade<6, MEDIA><<<grid, 256, 0, s>>>(p, x);
    if (m & 128u) k_shade<7, MEDIA><<<grid, 256, 0, s>>>(p, x);
    if (m & 256u) k_shade<8, MEDIA><<<grid, 256, 0, s>>>(p, x);
    if (m & 512u) k_shade<kClassCameraMiss, MEDIA><<<grid, 256, 0, s>>>(p, x);
}

hipError_t wave_shade(const WaveParams& p, int x, int grid, int trace_grid, bool count, hipStream_t s) {
    if (p.n_media != 0) launch_shade<true>(p, x, grid, s);
    else launch_shade<false>(p, x, grid, s);
    if (p.tail_paths > 0) {
        if (p.n_media != 0) {
            if (count) k_tail<true, true><<<trace_grid, kTraceBlock, 0, s>>>(p, x);
            else k_tail<false, true><<<trace_grid, kTraceBlock, 0, s>>>(p, x);
        } else {
            if (count) k_tail<true, false><<<trace_grid, kTraceBlock, 0, s>>>(p, x);
            else k_tail<false, false><<<trace_grid, kTraceBlock, 0, s>>>(p, x);
        }
    }
    return hipGetLastError();
}

template <class K>
int occupancy(K kernel, int block) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, block, 0) != hipSuccess) return 1;
    return std::max(n, 1);
}

int wave_trace_occupancy(int stack_lds) {
    if (stack_lds <= 8) return occupancy(k_trace<false, 8>, kTraceBlock);
    if (stack_lds <= 16) return occupancy(k_trace<false, 16>, kTraceBlock);
    return occupancy(k_trace<false, 32>, kTraceBlock);
}

int wave_shade_occupancy() {
    // the grid serves every class kernel; each is capped by what fits anyway
    return std::max({occupancy(k_shade<0, false>, 256), occupancy(k_shade<0, true>, 256), occupancy(k_regen, 256)});
}

hipError_t wave_accumulate(const WaveParams& p, int64_t npix, hipStream_t s) {
    k_accumulate<<<blocks(npix, 256), 256, 0, s>>>(p, npix);
    return hipGetLastError();
}

}  // namespace oxr
