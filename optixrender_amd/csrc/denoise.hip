// KPCNN denoiser inference on gfx950: BF16 operands on MFMA, FP32 accumulate.
//
// Replaces optixDenoiserInvoke (core/src/Renderer.cpp:738-780): reads the
// accumulated colour, albedo, normal and depth buffers of a render and writes
// the denoised colour that the tone map then reads (Renderer.cpp:786).  The
// reference's network is closed, so the network is our own:
//
//   features (10 per pixel): log(1 + colour) (3), albedo (3), normal (3),
//       depth d / (1 + d) (1), stored as BF16 packed three pixels wide (the
//       left neighbour's, the pixel's, the right neighbour's: 30 channels and
//       2 of zero), so that layer 1 runs as a 3x1 convolution over 32
//       channels (12 MFMAs per row of 64 pixels instead of 18 over 16);
//   five 3x3 convolutions 10 -> 32 -> 32 -> 32 -> 32 -> 25 (zero padding),
//       ReLU after the first four, activations BF16 NHWC with 32 channels;
//   softmax over the 25 outputs = a 5x5 kernel per pixel, applied to the
//       linear colour with clamped borders.
//
// Each convolution is an implicit GEMM  D[co][px] = sum_k W[co][k] X[k][px],
// k = tap * Cin + ci, on v_mfma_f32_32x32x16_bf16: one wave computes 32
// output channels x 32 consecutive pixels of a row, K in steps of 16.
// Operand layout of the instruction (lane l, element e):
//   A 32x16: row l % 32, k = 8 (l / 32) + e
//   B 16x32: column l % 32, k = 8 (l / 32) + e
//   D 32x32: register r holds row 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32
// so a lane's B operand is 8 consecutive channels of one pixel at one tap
// (one 16-byte load) and its A operand 8 consecutive input channels of one
// output channel, pre-arranged in lane order on the host.
#include "common.h"

namespace oxr {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

typedef short i16x2 __attribute__((ext_vector_type(2)));

// Two floats to packed BF16 (lo in the low half), rounded to nearest even:
// one v_cvt_pk_bf16_f32 on gfx950.
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){lo, hi}, bf16x2));
}

// ReLU then BF16 as BF16 then ReLU (a negative value rounds to a negative
// BF16 or -0, so the order does not change the result): the ReLU is a packed
// signed 16-bit max with 0 (v_pk_max_i16), which takes the MFMA results as
// they are rather than canonicalizing each before a float max.
__device__ __forceinline__ uint32_t pack_bf16_relu(float lo, float hi) {
    const i16x2 v = __builtin_bit_cast(i16x2, __builtin_convertvector((f32x2){lo, hi}, bf16x2));
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(v, (i16x2){0, 0}));
}

// log(1 + x) for x >= 0 from the hardware log2 (v_log_f32) of u = 1 + x
// and the first-order correction (x - (u - 1)) / u for the rounding of u:
// a few fp32 ulp, far inside the BF16 rounding of the features, at about a
// tenth of the device library's log1pf.
__device__ __forceinline__ float log1p_nonneg(float x) {
    const float u = 1.0f + x;
    const float c = (x - (u - 1.0f)) * __builtin_amdgcn_rcpf(u);
    const float l = __builtin_amdgcn_logf(u) * 0.693147180559945309f;
    return u == __builtin_inff() ? u : l + c;
}

// The whole network in one kernel.  A block owns a strip of kStripW output
// columns and seg_h output rows and walks down it one row step at a time;
// every layer keeps only the last four rows of its input in LDS (ring
// buffers), so activations never go to HBM.  Seven waves, each with one job:
//
//   wave 0       stores the feature row of the step (colour, albedo, normal,
//                depth -> BF16 features, zero outside the image) and the
//                linear colour of the last rows, each into an LDS ring;
//   layers 1..4  one wave each: 64 pixels (two MFMA tiles of 32) x 32
//                channels of the layer's output row y_l(t) from rows
//                y_l - 1 .. y_l + 1 of the previous layer's ring, with the
//                layer's weights held in registers;
//   layer 5      two waves, one tile each: logits, their softmax as a 5x5
//                kernel, applied to the colour ring.
//
// Layer l runs three steps behind layer l - 1 and a step has one barrier:
// in step t layer l reads rows y_l - 1 .. y_l + 1 of ring l - 1, all stored
// in earlier steps, while layer l - 1 stores row y_l + 2 into the slot of
// y_l - 2, which was last read in step t - 1.  So a ring of four rows lets
// every wave read, compute and store in the same phase.
// Zero padding: every stored row or pixel outside the image is zero, as the
// layer-by-layer network sees it.  Ring pixel p is image column x0 - 5 + p;
// layer l computes pixels p = l .. l + 63 (its needed range l .. 60 + 5 - l
// plus a few that nothing reads), the last layer as two waves of 32.
constexpr int kStripW = 56;              // output columns per block
constexpr int kRingW = 70;               // ring row: columns x0 - 5 .. x0 + 64
constexpr int kRingC = 32;               // channels of an activation ring (features: 16)
// A ring pixel's channels take its channel count plus 8 BF16 of padding (80
// or 48 bytes, odd multiples of 16): the 16-byte operand reads of a wave's
// consecutive pixels then spread over the LDS banks.
template <int C> constexpr int ring_stride() { return C + 8; }
// Wave roles.  Waves go to SIMD wave % 4, and the two waves of a SIMD share
// its matrix core and vector ALU, so the table spreads MFMAs and VALU work.
enum DnRole { kLoader = 0, kLayer1, kLayer2, kLayer3, kLayer4, kLastTile0, kLastTile1, kIdle };
constexpr int kRoles[] = {kLoader, kLastTile0, kLayer2, kLayer3, kLayer4, kLastTile1, kLayer1};
constexpr int kFusedWaves = sizeof(kRoles) / sizeof(kRoles[0]);
constexpr int kRingRows = 4;
// colour ring rows: in step t the last layer filters row ys + t - 15 with the
// colour of rows ys + t - 17 .. ys + t - 13 while the loader stores row ys + t
constexpr int kColourRows = 32;

__device__ __forceinline__ int ring_slot(int y) { return y & (kRingRows - 1); }  // y mod 4, also for y < 0

// The block's step barrier: this wave's LDS writes complete, then every wave
// meets.  Unlike __syncthreads() it does not wait for the wave's global loads
// in flight (the loader's prefetch of the next row), whose registers the
// compiler waits for where they are used.
__device__ __forceinline__ void step_barrier() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);  // no compiler motion of LDS accesses across
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_s_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// One layer's output row at ring pixels p0 .. p0 + 32 NT - 1: NT 32-pixel
// tiles, NT independent MFMA accumulation chains (taps in the order tap row,
// tap column, channel step), the B operands read from the input ring.
template <int CIN, int NT, int TX = 3>
__device__ __forceinline__ void dn_layer_row(const uint16_t* __restrict__ ring, const bf16x8* a, int y, int p0,
                                             int lane, const f32x16& bias, f32x16 (&acc)[NT]) {
    constexpr int STEPS = CIN / 16;
    constexpr int PS = ring_stride<CIN>();
    constexpr int G = TX * STEPS * NT;           // operands (and MFMAs) of one tap row
    const int px = lane & 31, half = lane >> 5;
#pragma unroll
    for (int tile = 0; tile < NT; tile++) acc[tile] = bias;  // the sum starts from the bias
    // Software pipeline over the three tap rows: the operands of row ty + 1
    // are read while row ty's MFMAs run, so that the LDS latency hides
    // behind the matrix core rather than stalling every MFMA.
    bf16x8 b[2][G];
    auto load = [&](int ty, bf16x8* dst) {
        const uint16_t* row = ring + ring_slot(y - 1 + ty) * kRingW * PS + (p0 + px - (TX == 3 ? 1 : 0)) * PS + 8 * half;
#pragma unroll
        for (int tx = 0; tx < TX; tx++)
#pragma unroll
            for (int s = 0; s < STEPS; s++)
#pragma unroll
                for (int tile = 0; tile < NT; tile++)
                    dst[(tx * STEPS + s) * NT + tile] =
                        *reinterpret_cast<const bf16x8*>(row + (tx + 32 * tile) * PS + 16 * s);
    };
    load(0, b[0]);
#pragma unroll
    for (int ty = 0; ty < 3; ty++) {
        if (ty < 2) load(ty + 1, b[(ty + 1) & 1]);
#pragma unroll
        for (int j = 0; j < TX * STEPS; j++) {
            const int k = ty * TX * STEPS + j;
#pragma unroll
            for (int tile = 0; tile < NT; tile++)
                acc[tile] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[k], b[ty & 1][NT * j + tile], acc[tile], 0, 0, 0);
        }
    }
    // The order to keep: reads of row 0, then each MFMA of rows 0 and 1
    // followed by one read of the next row (as many reads as MFMAs in a
    // row), then row 2's MFMAs: about a row of reads stays in flight.
    __builtin_amdgcn_sched_group_barrier(0x100, G, 0);
#pragma unroll
    for (int i = 0; i < 2 * G; i++) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
    }
    __builtin_amdgcn_sched_group_barrier(0x008, G, 0);
}

static_assert(5 + 63 < kRingW, "every layer's 64 pixels lie inside the ring row");

// What a layer does with its output row once the MFMAs are done: layers 1..4
// store their two tiles into their ring (bias already in, ReLU, BF16, zero
// outside the image).  The last layer has one tile per wave, its 25 outputs
// placed so that accumulator register r (r < 13) of a lane in the lower half
// of the wave holds tap r and that of the upper half tap 13 + r (r < 12):
// each lane takes its pixel's softmax over its taps, applies them to the
// colour ring, and the two lanes of the pixel add their halves through a
// v_permlane32_swap.  No branch depends on the lane before the final store,
// so the scheduler can interleave the work with the next row's MFMAs.
// The colour taps of the last layer's row y for this lane (see finish_row),
// read one step ahead so that the filter finds them in registers.
struct ColourTaps {
    float r[13], g[13], b[13];
};
__device__ __forceinline__ void load_colour_taps(ColourTaps& ct, const float4* crow, int p0, int lane, int y, int ys) {
    const int px = lane & 31, half = lane >> 5;
    const int pc = min(p0 + px, kRingW - 3);     // keeps the taps of lanes without a pixel inside the row
    const int s0 = y - ys - 2;                   // colour ring row of tap row 0
#pragma unroll
    for (int r = 0; r < 13; r++) {
        const int dy = half ? (13 + r) / 5 : r / 5, dx = half ? (13 + r) % 5 : r % 5;
        const float4 c = crow[((s0 + dy) & (kColourRows - 1)) * kRingW + pc + dx - 2];
        ct.r[r] = c.x;
        ct.g[r] = c.y;
        ct.b[r] = c.z;
    }
}

template <bool LAST, int NT>
__device__ __forceinline__ void finish_row(int p0, int lane, int y, const f32x16* acc, uint16_t* ring,
                                           const ColourTaps& ct, int x0, int y0, int w, int h, int seg_h,
                                           float4* __restrict__ out) {
    const int px = lane & 31, half = lane >> 5;
    if (!LAST) {
#pragma unroll
        for (int tile = 0; tile < NT; tile++) {
            const int p = p0 + 32 * tile + px;
            const int xo = x0 - 5 + p;
            const bool in = y >= 0 && y < h && xo >= 0 && xo < w;
            uint16_t* dst = ring + (ring_slot(y) * kRingW + p) * ring_stride<kRingC>();
#pragma unroll
            for (int g = 0; g < 4; g++) {
                const uint32_t lo = pack_bf16_relu(acc[tile][4 * g + 0], acc[tile][4 * g + 1]);
                const uint32_t hi = pack_bf16_relu(acc[tile][4 * g + 2], acc[tile][4 * g + 3]);
                *reinterpret_cast<uint2*>(dst + 8 * g + 4 * half) = make_uint2(in ? lo : 0u, in ? hi : 0u);
            }
        }
    } else {
        const int p = p0 + px;
        const int xo = x0 - 5 + p;
        const bool valid = y >= 0 && y < h && xo >= 0 && xo < w && p - 5 < kStripW && y >= y0 && y < y0 + seg_h;
        const f32x16 lg = acc[0];
        // lane's taps: r (lower half) or 13 + r (upper half); r = 12 only below
        float m = lg[0];
#pragma unroll
        for (int r = 1; r < 12; r++) m = fmaxf(m, lg[r]);
        m = half ? m : fmaxf(m, lg[12]);
        const auto ms = __builtin_amdgcn_permlane32_swap(__float_as_uint(m), __float_as_uint(m), false, false);
        m = fmaxf(__uint_as_float(ms[0]), __uint_as_float(ms[1]));
        constexpr float kLog2e = 1.44269504088896341f;
        const float mb = m * kLog2e;
        float sum = 0.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
#pragma unroll
        for (int r = 0; r < 13; r++) {
            float e = __builtin_amdgcn_exp2f(fmaf(lg[r], kLog2e, -mb));
            if (r == 12) e = half ? 0.0f : e;
            sum += e;
            cr = fmaf(e, ct.r[r], cr);
            cg = fmaf(e, ct.g[r], cg);
            cb = fmaf(e, ct.b[r], cb);
        }
        // both lanes of the pixel end with the same totals (lower + upper)
        const auto s1 = __builtin_amdgcn_permlane32_swap(__float_as_uint(sum), __float_as_uint(sum), false, false);
        const auto s2 = __builtin_amdgcn_permlane32_swap(__float_as_uint(cr), __float_as_uint(cr), false, false);
        const auto s3 = __builtin_amdgcn_permlane32_swap(__float_as_uint(cg), __float_as_uint(cg), false, false);
        const auto s4 = __builtin_amdgcn_permlane32_swap(__float_as_uint(cb), __float_as_uint(cb), false, false);
        const float inv = __builtin_amdgcn_rcpf(__uint_as_float(s1[0]) + __uint_as_float(s1[1]));  // >= 1
        float4 v = make_float4((__uint_as_float(s2[0]) + __uint_as_float(s2[1])) * inv,
                               (__uint_as_float(s3[0]) + __uint_as_float(s3[1])) * inv,
                               (__uint_as_float(s4[0]) + __uint_as_float(s4[1])) * inv, 1.0f);
        // computed by every lane before the store's branch, so that the work
        // stays in the MFMAs' basic block rather than sinking into the branch
        asm volatile("" : "+v"(v.x), "+v"(v.y), "+v"(v.z));
        if (valid && half == 0) out[(int64_t)y * w + xo] = v;
    }
}

// Layer l's walk down the strip, weights in registers, over NT 32-pixel
// tiles from tile `tile` on.  Its output row at
// step t is y_l(t) = ys + t + 1 - 3 l: three steps behind layer l - 1, so
// that the row computed in step t is finished (stored) in step t + 1 while
// the MFMAs of the next row run, and its readers take it in step t + 2.
template <int CIN, bool LAST, int NT, int TX = 3>
__device__ __forceinline__ void layer_walk(int l, int tile, int lane, const uint16_t* feat,
                                           uint16_t (*act)[kRingRows * kRingW * ring_stride<kRingC>()],
                                           const float4* crow, const bf16x8* __restrict__ wl, const DenoiserDev& d,
                                           int x0, int y0, int ys, int w, int h, int seg_h, int steps,
                                           float4* __restrict__ out) {
    const int half = lane >> 5;
    const int p0 = l + 32 * tile;                // its first ring pixel
    constexpr int NA = 3 * TX * CIN / 16;
    bf16x8 a[NA];
    const bf16x8* src = wl + d.lane_offset[l - 1];
#pragma unroll
    for (int i = 0; i < NA; i++) a[i] = src[i * 64 + lane];
    // this lane's biases in the accumulator's layout (register 4 g + i:
    // MFMA row 8 g + 4 half + i), the accumulation's starting value
    f32x16 bias;
#pragma unroll
    for (int g = 0; g < 4; g++)
#pragma unroll
        for (int i = 0; i < 4; i++) bias[4 * g + i] = d.bias[(l - 1) * 32 + 8 * g + 4 * half + i];
    const uint16_t* in_ring = l == 1 ? feat : act[l - 2];
    uint16_t* out_ring = LAST ? nullptr : act[l - 1];
    f32x16 prev[NT];
    int prev_y = 0;
    bool prev_needed = false;
    ColourTaps ct;                               // the last layer's: colour taps of prev_y
    for (int t = 0; t < steps; t++) {
        const int y = ys + t + 1 - 3 * l;        // this layer's output row
        const bool row_needed = y >= y0 - (5 - l) && y < y0 + seg_h + (5 - l);
        const bool compute = row_needed && y >= 0 && y < h;
        f32x16 acc[NT];
#pragma unroll
        for (int i = 0; i < NT; i++) acc[i] = bias;
        if (compute && prev_needed) {            // the steady state: one block for the scheduler
            dn_layer_row<CIN, NT, TX>(in_ring, a, y, p0, lane, bias, acc);
            finish_row<LAST, NT>(p0, lane, prev_y, prev, out_ring, ct, x0, y0, w, h, seg_h, out);
        } else {
            if (prev_needed) finish_row<LAST, NT>(p0, lane, prev_y, prev, out_ring, ct, x0, y0, w, h, seg_h, out);
            if (compute) dn_layer_row<CIN, NT, TX>(in_ring, a, y, p0, lane, bias, acc);
        }
        // row y's colour taps for the next step: its rows y - 2 .. y + 2 were
        // stored by the loader at least 12 steps ago, and the ring keeps 32
        if (LAST) load_colour_taps(ct, crow, p0, lane, y, ys);
#pragma unroll
        for (int i = 0; i < NT; i++) prev[i] = acc[i];
        prev_y = y;
        prev_needed = row_needed;
        step_barrier();
    }
}

__global__ __launch_bounds__(64 * kFusedWaves, 2) void k_dn_fused(
    const float4* __restrict__ color, const float4* __restrict__ albedo, const float4* __restrict__ normal,
    const float* __restrict__ depth, const bf16x8* __restrict__ wl, DenoiserDev d, int w, int h, int seg_h,
    float4* __restrict__ out) {
    // features packed three pixels wide for layer 1 (see the loader)
    __shared__ __align__(16) uint16_t feat[kRingRows * kRingW * ring_stride<32>()];
    __shared__ __align__(16) uint16_t act[4][kRingRows * kRingW * ring_stride<kRingC>()];
    // linear colour of the last 16 feature rows at clamped coordinates (the
    // 5x5 kernel's clamped borders), for the last layer
    __shared__ __align__(16) float4 crow[kColourRows * kRingW];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int strips = (w + kStripW - 1) / kStripW;
    const int x0 = (blockIdx.x % strips) * kStripW;
    const int y0 = (blockIdx.x / strips) * seg_h;
    const int ys = y0 - 5;                       // first feature row the segment needs
    const int steps = seg_h + 20;                // the last layer finishes its last row at step seg_h + 19
    // Two roles, one loop each: every wave passes the same barrier once per
    // step (waves are wave-uniform in their role), and the roles' registers
    // never overlap.
    const int role = kRoles[wave];
    if (role == kLoader) {
        // Loader: the inputs of a row (colour at clamped coordinates, albedo,
        // normal, depth; zero albedo/normal/depth outside the image) are
        // fetched three steps before the row is stored, into one of two
        // register sets, and turned into features one step before.
        struct Raw {
            float4 c[2], a[2], n[2];
            float d[2];
        };
        auto fetch_row = [&](int y, Raw& r) {
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const int p = lane + 64 * k;
                const int x = x0 - 5 + p;
                // clamped coordinates: always a pixel of the image; what lies
                // outside is zeroed when the row is converted
                const int64_t ci = (int64_t)min(max(y, 0), h - 1) * w + min(max(x, 0), w - 1);
                r.c[k] = color[ci];
                r.a[k] = albedo[ci];
                r.n[k] = normal[ci];
                r.d[k] = depth[ci];
            }
        };
        uint32_t fw[2][8];                       // features of the next row to store
        float4 cn[2];                            // and its colour
        auto convert = [&](int y, const Raw& r) {
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const int p = lane + 64 * k;
                const int x = x0 - 5 + p;
                const bool inside = p < kRingW && y >= 0 && y < h && x >= 0 && x < w;
                const float4 c = r.c[k];
                const float f[10] = {log1p_nonneg(fmaxf(c.x, 0.0f)), log1p_nonneg(fmaxf(c.y, 0.0f)),
                                     log1p_nonneg(fmaxf(c.z, 0.0f)), r.a[k].x, r.a[k].y, r.a[k].z,
                                     r.n[k].x, r.n[k].y, r.n[k].z, r.d[k]};
#pragma unroll
                for (int q = 0; q < 8; q++)
                    fw[k][q] = inside ? pack_bf16(2 * q < 10 ? f[2 * q] : 0.0f, 2 * q + 1 < 10 ? f[2 * q + 1] : 0.0f)
                                      : 0u;
                cn[k] = c;
            }
        };
        // step t: store row ys + t (converted last step), convert row
        // ys + t + 1 from `next`, then refill `next` with row ys + t + 3
        auto step = [&](int t, Raw& next) {
            // Layer 1 reads ring pixels 1 .. 64, each holding the 10 features
            // of the pixel to its left, its own and the pixel to its right
            // (32 channels, the last two zero): this lane's pixel `lane` and,
            // in lane 0, pixel 64.  The neighbours' come from the lanes that
            // converted them.
            // Whole-wave DPP shifts (v_mov_b32 wave_shr:1 / wave_shl:1) and
            // lane reads, no LDS traffic.
            uint32_t lf[5], rt[5], lf64[5], rt64[5];
#pragma unroll
            for (int i = 0; i < 5; i++) {
                const uint32_t up = (uint32_t)__builtin_amdgcn_mov_dpp((int)fw[0][i], 0x138, 0xf, 0xf, false);
                const uint32_t down = (uint32_t)__builtin_amdgcn_mov_dpp((int)fw[0][i], 0x130, 0xf, 0xf, false);
                const uint32_t p64 = (uint32_t)__builtin_amdgcn_readlane((int)fw[1][i], 0);
                lf[i] = up;                                  // pixel lane - 1
                rt[i] = lane == 63 ? p64 : down;             // pixel lane + 1
                lf64[i] = (uint32_t)__builtin_amdgcn_readlane((int)fw[0][i], 63);
                rt64[i] = (uint32_t)__builtin_amdgcn_readlane((int)fw[1][i], 1);
            }
            auto put = [&](int p, const uint32_t* l, const uint32_t* c, const uint32_t* r) {
                uint4* dst = reinterpret_cast<uint4*>(feat + (ring_slot(ys + t) * kRingW + p) * ring_stride<32>());
                dst[0] = make_uint4(l[0], l[1], l[2], l[3]);
                dst[1] = make_uint4(l[4], c[0], c[1], c[2]);
                dst[2] = make_uint4(c[3], c[4], r[0], r[1]);
                dst[3] = make_uint4(r[2], r[3], r[4], 0u);
            };
            put(lane, lf, fw[0], rt);
            if (lane == 0) put(64, lf64, fw[1], rt64);
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const int p = lane + 64 * k;
                if (p < kRingW) crow[(t & (kColourRows - 1)) * kRingW + p] = cn[k];
            }
            if (t + 1 < steps) {
                convert(ys + t + 1, next);
                if (t + 3 < steps) fetch_row(ys + t + 3, next);
            }
            step_barrier();
        };
        Raw r0, r1;
        fetch_row(ys, r0);
        fetch_row(ys + 1, r1);
        convert(ys, r0);
        if (2 < steps) fetch_row(ys + 2, r0);
        int t = 0;
        for (; t + 1 < steps; t += 2) {
            step(t, r1);                         // row t + 1 is in r1, row t + 2 in r0
            step(t + 1, r0);
        }
        if (t < steps) step(t, r1);
    } else if (role == kIdle) {
        for (int t = 0; t < steps; t++) step_barrier();
    } else if (role == kLayer1) {
        layer_walk<32, false, 2, 1>(1, 0, lane, feat, act, crow, wl, d, x0, y0, ys, w, h, seg_h, steps, out);
    } else if (role == kLastTile0 || role == kLastTile1) {
        layer_walk<32, true, 1>(5, role == kLastTile1, lane, feat, act, crow, wl, d, x0, y0, ys, w, h, seg_h, steps,
                                out);
    } else {
        layer_walk<32, false, 2>(role, 0, lane, feat, act, crow, wl, d, x0, y0, ys, w, h, seg_h, steps, out);
    }
}

}  // namespace

hipError_t denoise_run(const DenoiserDev& d, const float4* color, const float4* albedo, const float4* normal,
                       const float* depth, int w, int h, int cus, float4* out, hipStream_t s) {
    if ((int64_t)w * h == 0) return hipSuccess;
    // One block per CU at a time (its LDS rings take 136 KB of the CU's
    // 160 KB).  Rows per block: the fewest rounds of blocks times the steps
    // of a walk (its rows plus 20 steps of filling and draining the pipeline).
    const int64_t strips = (w + kStripW - 1) / kStripW;
    const int64_t slots = cus > 0 ? cus : 256;
    int seg_h = 16;
    int64_t best = -1;
    for (int sh = 16; sh < h + 8; sh += 8) {
        const int64_t blocks = strips * ((h + sh - 1) / sh);
        const int64_t cost = ((blocks + slots - 1) / slots) * (sh + 20);
        if (best < 0 || cost < best) { best = cost; seg_h = sh; }
    }
    const int64_t blocks = strips * ((h + seg_h - 1) / seg_h);
    k_dn_fused<<<(unsigned)blocks, 64 * kFusedWaves, 0, s>>>(color, albedo, normal, depth,
                                                            reinterpret_cast<const bf16x8*>(d.lanes), d, w, h, seg_h,
                                                            out);
    return hipGetLastError();
}

}  // namespace oxr
