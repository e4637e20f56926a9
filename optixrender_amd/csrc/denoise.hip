// KPCNN denoiser inference on gfx950: BF16 operands on MFMA, FP32 accumulate.
//
// Replaces optixDenoiserInvoke (core/src/Renderer.cpp:738-780): reads the
// accumulated colour, albedo, normal and depth buffers of a render and writes
// the denoised colour that the tone map then reads (Renderer.cpp:786).  The
// reference's network is closed, so the network is our own:
//
//   features (10 per pixel): log(1 + colour) (3), albedo (3), normal (3),
//       depth d / (1 + d) (1), stored as BF16 NHWC padded to 16 channels;
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

__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even, NaN kept
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__global__ void k_dn_features(const float4* __restrict__ color, const float4* __restrict__ albedo,
                              const float4* __restrict__ normal, const float* __restrict__ depth, int64_t n,
                              uint16_t* __restrict__ feat) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 c = color[i], a = albedo[i], nr = normal[i];
    const float f[10] = {log1pf(fmaxf(c.x, 0.0f)), log1pf(fmaxf(c.y, 0.0f)), log1pf(fmaxf(c.z, 0.0f)),
                         a.x, a.y, a.z, nr.x, nr.y, nr.z, depth[i]};
    uint32_t w[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const float lo = 2 * k < 10 ? f[2 * k] : 0.0f, hi = 2 * k + 1 < 10 ? f[2 * k + 1] : 0.0f;
        w[k] = (uint32_t)f2bf(lo) | ((uint32_t)f2bf(hi) << 16);
    }
    uint4* dst = reinterpret_cast<uint4*>(feat + 16 * i);
    dst[0] = make_uint4(w[0], w[1], w[2], w[3]);
    dst[1] = make_uint4(w[4], w[5], w[6], w[7]);
}

// One 3x3 convolution.  CIN in {16, 32} (padded input channels), 32 padded
// outputs.  A wave computes kDnRows rows of 32 pixels: its weights (all taps)
// stay in registers, and each input row segment it loads feeds the up to three
// output rows that read it, so operand traffic per MFMA falls by about 3x.
// Every accumulator receives its taps in the same order as one row at a time
// would give it (tap row, tap column, channel step).  A block of 4 waves
// covers 4 such row groups.  LAST: the 25 logits become the 5x5 kernel
// applied to `color`, written to `out`.
constexpr int kDnRows = 4;
template <int CIN, bool LAST>
__global__ __launch_bounds__(256) void k_dn_conv(const uint16_t* __restrict__ in, const bf16x8* __restrict__ wl,
                                                 const float* __restrict__ bias, int w, int h, int tiles_x,
                                                 uint16_t* __restrict__ out_act, const float4* __restrict__ color,
                                                 float4* __restrict__ out) {
    constexpr int STEPS = CIN / 16;
    constexpr int R = kDnRows;
    constexpr int SROW = 40;  // staging row of one pixel's 32 outputs, padded (bf16)
    __shared__ float logits[LAST ? 4 : 1][32][33];
    __shared__ __align__(16) uint16_t stage[LAST ? 1 : 4][32 * SROW];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tile = blockIdx.x * 4 + wave;
    const int row_groups = (h + R - 1) / R;
    // a wave past the last tile runs masked, so every wave reaches the barriers
    const bool live = tile < tiles_x * row_groups;
    const int y0 = live ? (tile / tiles_x) * R : 0, x0 = live ? (tile % tiles_x) * 32 : w;
    const int px = lane & 31, half = lane >> 5;
    const int x = x0 + px;
    bf16x8 a[9 * STEPS];
#pragma unroll
    for (int t = 0; t < 9 * STEPS; t++) a[t] = wl[t * 64 + lane];
    f32x16 acc[R];
#pragma unroll
    for (int r = 0; r < R; r++) acc[r] = f32x16{};
#pragma unroll
    for (int iy = 0; iy < R + 2; iy++) {  // input row y0 - 1 + iy: tap row iy - r of output row r
        const int sy = y0 - 1 + iy;
#pragma unroll
        for (int dx = 0; dx < 3; dx++) {
            const int sx = x + dx - 1;
            const bool inside = live && sy >= 0 && sy < h && sx >= 0 && sx < w;
            const uint16_t* src = in + ((int64_t)(inside ? sy : 0) * w + (inside ? sx : 0)) * CIN + 8 * half;
#pragma unroll
            for (int s = 0; s < STEPS; s++) {
                bf16x8 b = {};
                if (inside) b = *reinterpret_cast<const bf16x8*>(src + 16 * s);
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int ty = iy - r;
                    if (ty >= 0 && ty <= 2)
                        acc[r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[(ty * 3 + dx) * STEPS + s], b, acc[r], 0, 0,
                                                                         0);
                }
            }
        }
    }
    if (!LAST) {
        // Each row's outputs go through the wave's LDS staging area so that the
        // global writes are whole 16-byte pieces of consecutive pixels (one
        // contiguous 1 KB per store instruction).  Only this wave touches its
        // area, and a wave's LDS operations complete in order.
        uint16_t* sw = stage[wave];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int y = y0 + r;
#pragma unroll
            for (int g = 0; g < 4; g++) {
                const int co = 8 * g + 4 * half;
                const uint32_t lo = (uint32_t)f2bf(fmaxf(acc[r][4 * g + 0] + bias[co + 0], 0.0f)) |
                                    ((uint32_t)f2bf(fmaxf(acc[r][4 * g + 1] + bias[co + 1], 0.0f)) << 16);
                const uint32_t hi = (uint32_t)f2bf(fmaxf(acc[r][4 * g + 2] + bias[co + 2], 0.0f)) |
                                    ((uint32_t)f2bf(fmaxf(acc[r][4 * g + 3] + bias[co + 3], 0.0f)) << 16);
                *reinterpret_cast<uint2*>(sw + px * SROW + co) = make_uint2(lo, hi);
            }
            __atomic_signal_fence(__ATOMIC_SEQ_CST);  // no compiler motion of the LDS accesses across
        __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const int c = lane + 64 * k, p = c >> 2, part = c & 3;  // pixel p of the row, 8 channels
                const uint4 v = *reinterpret_cast<const uint4*>(sw + p * SROW + 8 * part);
                if (live && y < h && x0 + p < w)
                    *reinterpret_cast<uint4*>(out_act + ((int64_t)y * w + x0 + p) * 32 + 8 * part) = v;
            }
            __atomic_signal_fence(__ATOMIC_SEQ_CST);  // no compiler motion of the LDS accesses across
        __builtin_amdgcn_wave_barrier();  // the next row reuses the area
        }
        return;
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int y = y0 + r;
#pragma unroll
        for (int g = 0; g < 4; g++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int co = 8 * g + 4 * half + i;
                logits[wave][px][co] = acc[r][4 * g + i] + bias[co];
            }
        // The wave's own logits only: a wave's LDS operations complete in order.
        __atomic_signal_fence(__ATOMIC_SEQ_CST);  // no compiler motion of the LDS accesses across
        __builtin_amdgcn_wave_barrier();
        // Both halves of the wave share the pixel's 25 taps (rows 0-2 and 3-4
        // of the 5x5 kernel); the partial sums meet across the halves.
        float sum = 0.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
        if (live && x < w && y < h) {
            float m = logits[wave][px][0];
            for (int t = 1; t < 25; t++) m = fmaxf(m, logits[wave][px][t]);
            const int t0 = half ? 15 : 0, t1 = half ? 25 : 15;
            for (int t = t0; t < t1; t++) {
                const float e = expf(logits[wave][px][t] - m);
                const int sy = min(max(y + t / 5 - 2, 0), h - 1), sx = min(max(x + t % 5 - 2, 0), w - 1);
                const float4 c = color[(int64_t)sy * w + sx];
                sum += e;
                cr += e * c.x;
                cg += e * c.y;
                cb += e * c.z;
            }
        }
        sum += __shfl_xor(sum, 32);
        cr += __shfl_xor(cr, 32);
        cg += __shfl_xor(cg, 32);
        cb += __shfl_xor(cb, 32);
        if (live && half == 0 && x < w && y < h) out[(int64_t)y * w + x] = make_float4(cr / sum, cg / sum, cb / sum, 1.0f);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);  // no compiler motion of the LDS accesses across
        __builtin_amdgcn_wave_barrier();  // the next row's logits reuse the array
    }
}

inline int nblocks(int64_t n, int b) { return (int)((n + b - 1) / b); }

}  // namespace

size_t denoise_scratch_bytes(int w, int h) { return (size_t)w * h * (16 + 32 + 32) * sizeof(uint16_t); }

hipError_t denoise_run(const DenoiserDev& d, const float4* color, const float4* albedo, const float4* normal,
                       const float* depth, int w, int h, void* scratch, float4* out, hipStream_t s) {
    const int64_t n = (int64_t)w * h;
    if (n == 0) return hipSuccess;
    uint16_t* f0 = static_cast<uint16_t*>(scratch);  // 16 channels
    uint16_t* a = f0 + 16 * n;                        // 32 channels
    uint16_t* b = a + 32 * n;                         // 32 channels
    const bf16x8* wl = reinterpret_cast<const bf16x8*>(d.lanes);
    k_dn_features<<<nblocks(n, 256), 256, 0, s>>>(color, albedo, normal, depth, n, f0);
    const int tiles_x = (w + 31) / 32;
    const int nb = nblocks((int64_t)tiles_x * ((h + kDnRows - 1) / kDnRows), 4);
    k_dn_conv<16, false><<<nb, 256, 0, s>>>(f0, wl + d.lane_offset[0], d.bias + 0 * 32, w, h, tiles_x, a, nullptr,
                                            nullptr);
    k_dn_conv<32, false><<<nb, 256, 0, s>>>(a, wl + d.lane_offset[1], d.bias + 1 * 32, w, h, tiles_x, b, nullptr,
                                            nullptr);
    k_dn_conv<32, false><<<nb, 256, 0, s>>>(b, wl + d.lane_offset[2], d.bias + 2 * 32, w, h, tiles_x, a, nullptr,
                                            nullptr);
    k_dn_conv<32, false><<<nb, 256, 0, s>>>(a, wl + d.lane_offset[3], d.bias + 3 * 32, w, h, tiles_x, b, nullptr,
                                            nullptr);
    k_dn_conv<32, true><<<nb, 256, 0, s>>>(b, wl + d.lane_offset[4], d.bias + 4 * 32, w, h, tiles_x, nullptr, color,
                                           out);
    return hipGetLastError();
}

}  // namespace oxr
