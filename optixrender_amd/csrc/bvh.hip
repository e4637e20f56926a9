// LBVH build on gfx950 (replaces optixAccelBuild/optixAccelCompact,
// core/src/Renderer.cpp:146-285).
//   1. centroid bounds: block min/max reduction, order-preserving integer
//      atomics (exact).
//   2. 30-bit Morton codes of triangle centroids ((v0+v1+v2)*(1/3)),
//      quantised as min(max(n*1024, 0), 1023) per axis.
//   3. LSD radix sort of (code, triangle id) in 4 passes of 8 bits: per-tile
//      LDS histograms, one scan, stable scatter ranked with wave64 ballots.
//   4. Karras 2012 hierarchy, one thread per internal node; equal codes are
//      split on the sorted position (augmented key).
//   5. bottom-up passes, one kernel per tree depth (depths by pointer
//      jumping): the refit, or rounds of treelet restructuring, then the
//      relabelling into Karras's layout of the new leaf order.
//   6. the traversal's 32-byte compressed node records (k_quantize).
#include <algorithm>
#include "common.h"

namespace oxr {
namespace {

constexpr int kSortBlock = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortBlock * kSortItems;  // keys per block
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;

__device__ __forceinline__ uint32_t float_to_ordered(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ordered_to_float(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__device__ __forceinline__ void centroid(const float* __restrict__ p, float c[3]) {
#pragma unroll
    for (int a = 0; a < 3; a++) c[a] = ((p[a] + p[3 + a]) + p[6 + a]) * (1.0f / 3.0f);
}

__global__ void k_init_bounds(uint32_t* b) {
    if (threadIdx.x < 3) b[threadIdx.x] = 0xffffffffu;      // min (ordered)
    else if (threadIdx.x < 8) b[threadIdx.x] = 0u;          // max (ordered), error word, pad
}

__global__ __launch_bounds__(256) void k_centroid_bounds(const float* __restrict__ pos, int64_t n, uint32_t* bounds) {
    __shared__ uint32_t smin[3][256], smax[3][256];
    uint32_t lo[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, hi[3] = {0u, 0u, 0u};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        float c[3];
        centroid(pos + 9 * i, c);
#pragma unroll
        for (int a = 0; a < 3; a++) {
            uint32_t o = float_to_ordered(c[a]);
            lo[a] = min(lo[a], o);
            hi[a] = max(hi[a], o);
        }
    }
#pragma unroll
    for (int a = 0; a < 3; a++) { smin[a][threadIdx.x] = lo[a]; smax[a][threadIdx.x] = hi[a]; }
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
#pragma unroll
            for (int a = 0; a < 3; a++) {
                smin[a][threadIdx.x] = min(smin[a][threadIdx.x], smin[a][threadIdx.x + s]);
                smax[a][threadIdx.x] = max(smax[a][threadIdx.x], smax[a][threadIdx.x + s]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 3) atomicMin(&bounds[threadIdx.x], smin[threadIdx.x][0]);
    else if (threadIdx.x < 6) atomicMax(&bounds[threadIdx.x], smax[threadIdx.x - 3][0]);
}

__device__ __forceinline__ uint32_t expand10(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
__device__ __forceinline__ uint32_t quantize(float c, float lo, float ext) {
    float nrm = ext > 0.0f ? (c - lo) / ext : 0.0f;
    float q = fminf(fmaxf(nrm * 1024.0f, 0.0f), 1023.0f);
    return (uint32_t)q;
}

__global__ __launch_bounds__(256) void k_morton(const float* __restrict__ pos, int64_t n, const uint32_t* bounds,
                                                 uint32_t* keys, uint32_t* vals) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    float lo[3], ext[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        lo[a] = ordered_to_float(bounds[a]);
        ext[a] = ordered_to_float(bounds[3 + a]) - lo[a];
    }
    float c[3];
    centroid(pos + 9 * i, c);
    keys[i] = (expand10(quantize(c[0], lo[0], ext[0])) << 2) | (expand10(quantize(c[1], lo[1], ext[1])) << 1) |
              expand10(quantize(c[2], lo[2], ext[2]));
    vals[i] = (uint32_t)i;
}

// ---- radix sort ----
// hist layout: [digit][block] so that one exclusive scan gives every
// (digit, block) its global output offset.
__global__ __launch_bounds__(kSortBlock) void k_radix_hist(const uint32_t* __restrict__ keys, int64_t n, int shift,
                                                            uint32_t* hist, int nblocks) {
    __shared__ uint32_t h[kRadix];
    h[threadIdx.x] = 0;
    __syncthreads();
    int64_t base = (int64_t)blockIdx.x * kSortTile;
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        int64_t i = base + j * kSortBlock + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & (kRadix - 1)], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of m uint32 in place by one block of 1024 threads:
// each thread owns a contiguous chunk.
__global__ __launch_bounds__(1024) void k_scan_single(uint32_t* data, int64_t m) {
    __shared__ uint32_t part[1024];
    int64_t chunk = (m + 1023) / 1024;
    int64_t b = threadIdx.x * chunk, e = min(b + chunk, m);
    uint32_t s = 0;
    for (int64_t i = b; i < e; i++) s += data[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        uint32_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - s;
    for (int64_t i = b; i < e; i++) {
        uint32_t v = data[i];
        data[i] = run;
        run += v;
    }
}

// A longer exclusive scan in three steps: tile sums, a scan of those (one
// block), then each tile scanned in LDS from its offset.  A tile is 2048
// values, eight consecutive per thread.
constexpr int kScanTile = 2048;

__global__ __launch_bounds__(256) void k_scan_reduce(const uint32_t* __restrict__ data, int64_t m, uint32_t* tsum) {
    __shared__ uint32_t red[256 / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanTile / 256; j++) {
        const int64_t i = base + j * 256 + threadIdx.x;
        if (i < m) s += data[i];
    }
    for (int o = 32; o > 0; o >>= 1) s += (uint32_t)__shfl_xor((int)s, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) tsum[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(256) void k_scan_apply(uint32_t* data, int64_t m, const uint32_t* __restrict__ toff) {
    __shared__ uint32_t wsum[256 / 64];
    constexpr int kPer = kScanTile / 256;
    const int64_t b = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kPer;
    uint32_t v[kPer], t = 0;
#pragma unroll
    for (int j = 0; j < kPer; j++) {
        v[j] = b + j < m ? data[b + j] : 0u;
        t += v[j];
    }
    // inclusive scan of the thread totals over the wave, then over the waves
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t inc = t;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = (uint32_t)__shfl_up((int)inc, o);
        if (lane >= o) inc += u;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t run = toff[blockIdx.x] + inc - t;
    for (int k = 0; k < wave; k++) run += wsum[k];
#pragma unroll
    for (int j = 0; j < kPer; j++) {
        if (b + j < m) data[b + j] = run;
        run += v[j];
    }
}

__global__ __launch_bounds__(kSortBlock) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                               const uint32_t* __restrict__ vin, int64_t n, int shift,
                                                               const uint32_t* __restrict__ offs, int nblocks,
                                                               uint32_t* kout, uint32_t* vout) {
    constexpr int kWaves = kSortBlock / 64;
    __shared__ uint32_t base[kRadix];
    __shared__ uint32_t run[kRadix];
    __shared__ uint32_t wcount[kWaves][kRadix];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    base[threadIdx.x] = offs[(size_t)threadIdx.x * nblocks + blockIdx.x];
    run[threadIdx.x] = 0;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int64_t tile = (int64_t)blockIdx.x * kSortTile;
    for (int j = 0; j < kSortItems; j++) {
#pragma unroll
        for (int w = 0; w < kWaves; w++) wcount[w][threadIdx.x] = 0;
        __syncthreads();
        int64_t i = tile + j * kSortBlock + threadIdx.x;
        bool valid = i < n;
        uint32_t key = valid ? kin[i] : 0u, val = valid ? vin[i] : 0u;
        uint32_t d = (key >> shift) & (kRadix - 1);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kRadixBits; b++) {
            uint64_t m = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? m : ~m;
        }
        uint32_t lrank = __popcll(peers & lt);
        if (valid && lrank == 0) wcount[wave][d] = __popcll(peers);
        __syncthreads();
        uint32_t rank = 0;
        if (valid) {
            uint32_t before = run[d];
            for (int w = 0; w < wave; w++) before += wcount[w][d];
            rank = before + lrank;
        }
        __syncthreads();
        {
            uint32_t tot = 0;
#pragma unroll
            for (int w = 0; w < kWaves; w++) tot += wcount[w][threadIdx.x];
            run[threadIdx.x] += tot;
        }
        if (valid) {
            uint32_t dst = base[d] + rank;
            kout[dst] = key;
            vout[dst] = val;
        }
        __syncthreads();
    }
}

// ---- Karras hierarchy ----
__device__ __forceinline__ int delta(const uint32_t* __restrict__ c, int64_t n, int64_t i, int64_t j) {
    if (j < 0 || j >= n) return -1;
    uint32_t ci = c[i], cj = c[j];
    if (ci != cj) return __clz((int)(ci ^ cj));
    return 32 + __clz((int)((uint32_t)i ^ (uint32_t)j));
}

__global__ __launch_bounds__(256) void k_karras(const uint32_t* __restrict__ codes, int64_t n, uint32_t* left,
                                                 uint32_t* right, uint32_t* parent) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t ni = n - 1;
    if (i >= ni) return;
    int d = (delta(codes, n, i, i + 1) - delta(codes, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta(codes, n, i, i - d);
    int64_t lmax = 2;
    while (delta(codes, n, i, i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
        if (delta(codes, n, i, i + (l + t) * d) > dmin) l += t;
    int64_t j = i + l * d;
    int dnode = delta(codes, n, i, j);
    int64_t s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(codes, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int64_t gamma = i + s * d + min(d, 0);
    uint32_t L = (min(i, j) == gamma) ? (kLeafBit | (uint32_t)gamma) : (uint32_t)gamma;
    uint32_t R = (max(i, j) == gamma + 1) ? (kLeafBit | (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
    left[i] = L;
    right[i] = R;
    parent[(L & kLeafBit) ? ni + (L & ~kLeafBit) : L] = (uint32_t)i;
    parent[(R & kLeafBit) ? ni + (R & ~kLeafBit) : R] = (uint32_t)i;
}

// One thread per leaf: its box and its triangle record in sorted order.
// The internal boxes follow level by level (k_tree_level).
__global__ __launch_bounds__(256) void k_leaf_boxes(const float* __restrict__ pos, int64_t n,
                                                     const uint32_t* __restrict__ order, float* lo, float* hi,
                                                     float4* tris) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t tid = order[k];
    const float* p = pos + 9 * (size_t)tid;
    const int64_t node = (n - 1) + k;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        lo[3 * node + a] = fminf(fminf(p[a], p[3 + a]), p[6 + a]);
        hi[3 * node + a] = fmaxf(fmaxf(p[a], p[3 + a]), p[6 + a]);
    }
    tris[3 * k + 0] = make_float4(p[0], p[1], p[2], __uint_as_float(tid));
    tris[3 * k + 1] = make_float4(p[3], p[4], p[5], 0.0f);
    tris[3 * k + 2] = make_float4(p[6], p[7], p[8], 0.0f);
}

// exact 2^e (|e| < 1023)
__device__ __forceinline__ double pow2(int e) { return __longlong_as_double((long long)(e + 1023) << 52); }

// Smallest scale exponent whose 253 steps reach the extent (kQuantEmin for
// an empty or NaN extent).
__device__ int quant_exponent(double extent) {
    if (!(extent > 0.0)) return kQuantEmin;
    const int k = (int)((__double_as_longlong(extent) >> 52) & 0x7ff) - 1022;  // extent < 2^k when normal
    int e = max(kQuantEmin, k - 9);
    while (e < kQuantEmax && extent > 253.0 * pow2(e)) e++;
    return e;
}

// One thread per internal node: its compressed record (layout in common.h).
// Bounds are computed in double, so they round outwards whatever the
// magnitudes: a child's lower bound lies at least half a step below its box
// (or at the origin), its upper bound at least half a step above.
__global__ __launch_bounds__(256) void k_quantize(int64_t ni, const uint32_t* __restrict__ left,
                                                   const uint32_t* __restrict__ right, const float* __restrict__ lo,
                                                   const float* __restrict__ hi, uint4* qnodes) {
    const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (v >= ni) return;
    const uint32_t cl = left[v], cr = right[v];
    const int64_t ch[2] = {(cl & kLeafBit) ? ni + (cl & ~kLeafBit) : cl, (cr & kLeafBit) ? ni + (cr & ~kLeafBit) : cr};
    int e[3];
    float org[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        org[a] = lo[3 * v + a];
        e[a] = quant_exponent((double)hi[3 * v + a] - (double)org[a]);
    }
    uint32_t q[3] = {0u, 0u, 0u};
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const double step_inv = pow2(-e[a]);
            const double xl = ((double)lo[3 * ch[c] + a] - (double)org[a]) * step_inv;
            const double xh = ((double)hi[3 * ch[c] + a] - (double)org[a]) * step_inv;
            const uint32_t ql = (uint32_t)(xl - 0.5 >= 0.0 ? floor(xl - 0.5) : 0.0);
            const uint32_t qh = (uint32_t)(xh + 0.5 < 254.0 ? floor(xh + 0.5) + 1.0 : 255.0);
            const int bl = 6 * c + a, bh = 6 * c + 3 + a;  // byte positions
            q[bl >> 2] |= ql << (8 * (bl & 3));
            q[bh >> 2] |= qh << (8 * (bh & 3));
        }
    const uint32_t w3 = ((cl & ~kLeafBit) & kQGamma) | ((cl & kLeafBit) ? kQLeft : 0u) | ((cr & kLeafBit) ? kQRight : 0u);
    const uint32_t w4 = (uint32_t)(uint8_t)(int8_t)e[0] | (uint32_t)(uint8_t)(int8_t)e[1] << 8 |
                        (uint32_t)(uint8_t)(int8_t)e[2] << 16;
    qnodes[2 * v] = make_uint4(__float_as_uint(org[0]), __float_as_uint(org[1]), __float_as_uint(org[2]), w3);
    qnodes[2 * v + 1] = make_uint4(w4, q[0], q[1], q[2]);
}

// ---- bottom-up passes: refit and treelet restructuring (Karras & Aila
// 2013), then Karras relabelling ----
// A bottom-up pass runs one kernel per tree depth, deepest first: a node's
// children are finished by the kernels before its own, and the nodes of one
// depth have disjoint subtrees.  (A climb with per-node arrival counters
// would need a device-scope fence per node, and on this chip a device-scope
// fence writes back and invalidates the XCD's L2.)  Depths come from
// pointer jumping over the parent links.  At each node the pass recomputes
// box, leaf count and SAH cost from the children and, in a restructuring
// round, if at least kTreeletLeaves << round leaves lie below it, one wave
// optimises its treelet: the treelet leaf (internal) of largest area is
// replaced by its two children until there are seven, a dynamic program over
// the subsets of those (in LDS, a lane per subset, subsets by size) finds the
// binary tree of least SAH cost, and if it is cheaper the treelet's internal
// nodes are rewired to it (the root keeps its id; the others are reused in
// the order they were expanded, depth first, left part first).  Ties go to
// the first candidate in a fixed order, so the tree is the one a serial
// post-order pass builds.  The tree is then relabelled as Karras's layout of
// its own leaf order: every node's leaves are consecutive, internal ids
// follow the endpoint rule (a left child takes the last position of its
// range, a right child the first), so a node's children are again gamma and
// gamma + 1.
constexpr int kTreeletLeaves = 7;
constexpr int kTreeletSubsets = 1 << kTreeletLeaves;
constexpr int kDepthJumps = 13;   // pointer-jumping steps: depths below 2^13
constexpr float kCostNode = 1.2f, kCostTri = 1.0f;

__device__ __forceinline__ float box_area(const float lo[3], const float hi[3]) {
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return 2.0f * ((dx * dy + dy * dz) + dz * dx);
}

// The tree as the passes see it.  Every node a kernel reads was written by
// an earlier kernel of the pass (its children lie deeper), so plain loads
// and stores suffice.
struct TreeDev {
    int64_t ni;
    uint32_t *left, *right, *parent, *nl;
    float *lo, *hi, *cost;
    __device__ int64_t index(uint32_t ref) const { return (ref & kLeafBit) ? ni + (ref & ~kLeafBit) : ref; }
    // internal node v from its two children: box, leaf count and SAH cost
    __device__ void join(int64_t v) const {
        const int64_t a = index(left[v]), c = index(right[v]);
        float l[3], h[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            l[k] = fminf(lo[3 * a + k], lo[3 * c + k]);
            h[k] = fmaxf(hi[3 * a + k], hi[3 * c + k]);
            lo[3 * v + k] = l[k];
            hi[3 * v + k] = h[k];
        }
        nl[v] = nl[a] + nl[c];
        cost[v] = kCostNode * box_area(l, h) + cost[a] + cost[c];
    }
};

// One wave's treelet (LDS).  Slots hold the treelet leaves; subsets of them
// are bit masks over the slots.
struct TreeletLds {
    float cost[kTreeletSubsets];         // the dynamic program's best SAH cost per subset
    float sarea[kTreeletSubsets];        // each subset's box area (slot boxes folded in ascending order)
    float jbox[kTreeletLeaves][6];       // the rewired nodes' boxes, by join position
    float jcost[kTreeletLeaves];         // ... their SAH costs
    uint32_t jnl[kTreeletLeaves];        // ... their leaf counts
    uint8_t jat[kTreeletSubsets];        // a rewired node's join position, by its leaf subset
    float lb[kTreeletLeaves][6];         // slot boxes (lo.xyz, hi.xyz)
    float pbox[8 + 16][6];               // boxes of the subsets of slots 0..2, then of slots 3..6
    float area[kTreeletLeaves];          // slot box areas
    float scost[kTreeletLeaves];         // slot SAH costs
    uint32_t snl[kTreeletLeaves];        // slot leaf counts
    uint32_t ref[kTreeletLeaves];        // slot node refs
    uint32_t cl[kTreeletLeaves], cr[kTreeletLeaves];  // an internal slot's children
    uint32_t inner[kTreeletLeaves];      // the treelet's internal nodes below its root, in expansion order
    int64_t st_node[kTreeletLeaves], order_node[kTreeletLeaves];
    int st_set[kTreeletLeaves], order_set[kTreeletLeaves];
    uint8_t part[kTreeletSubsets];       // the subset's best left part
    int m;
    float root_cost;
};

// The wave's LDS writes before this point are seen by all its lanes after it.
__device__ __forceinline__ void wave_lds_sync() {
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_wave_barrier();
}

// The subsets of the seven slots ordered by size, then value: those of
// size k are mask[start[k]] .. mask[start[k + 1] - 1].
struct SubsetTable {
    uint8_t mask[kTreeletSubsets];
    uint8_t start[kTreeletLeaves + 2];
};
constexpr SubsetTable make_subset_table() {
    SubsetTable t{};
    int o = 0;
    for (int k = 0; k <= kTreeletLeaves; k++) {
        t.start[k] = (uint8_t)o;
        for (int s = 0; s < kTreeletSubsets; s++)
            if (__builtin_popcount(s) == k) t.mask[o++] = (uint8_t)s;
    }
    t.start[kTreeletLeaves + 1] = (uint8_t)o;
    return t;
}
__constant__ SubsetTable kSubsets = make_subset_table();

// The deposit table: at (r >> 1) * 64 + u, the bits of u placed, in order,
// on the set bits of r, for every r without bit 0 and every u < 64 (the
// subset of r numbered u)
struct alignas(16) DepositTable {
    uint8_t d[64 * 64];
};
constexpr DepositTable make_deposit_table() {
    DepositTable t{};
    for (int k = 0; k < 64 * 64; k++) {
        int u = k & 63, r = (k >> 6) << 1, d = 0;
        for (int b = 0; r; b++) {
            const int low = r & -r;
            if ((u >> b) & 1) d |= low;
            r ^= low;
        }
        t.d[k] = (uint8_t)d;
    }
    return t;
}
__constant__ DepositTable kDeposit = make_deposit_table();
struct SubsetStart {
    int v[kTreeletLeaves + 2];
};
constexpr SubsetStart make_subset_start() {
    SubsetStart r{};
    const SubsetTable t = make_subset_table();
    for (int k = 0; k < kTreeletLeaves + 2; k++) r.v[k] = t.start[k];
    return r;
}
constexpr SubsetStart kSubsetStart = make_subset_start();


// slot i of the treelet := node ref r (its box, area, leaf count, cost, and
// for an internal node its children, read with the rest so that expanding
// the slot needs no further dependent load)
__device__ __forceinline__ void treelet_slot(const TreeDev& T, TreeletLds& L, int i, uint32_t r) {
    const int64_t x = T.index(r);
    if (!(r & kLeafBit)) {
        L.cl[i] = T.left[x];
        L.cr[i] = T.right[x];
    }
    float l[3], h[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        l[k] = T.lo[3 * x + k];
        h[k] = T.hi[3 * x + k];
        L.lb[i][k] = l[k];
        L.lb[i][3 + k] = h[k];
    }
    L.ref[i] = r;
    L.area[i] = box_area(l, h);
    L.snl[i] = T.nl[x];
    L.scost[i] = T.cost[x];
}

// Treelet of internal node v, by the whole wave (every lane calls it);
// smask is the subset table and dtab the deposit table in LDS.
__device__ void treelet_wave(const TreeDev& T, int64_t v, TreeletLds& L, const uint8_t* smask, const uint8_t* dtab,
                             int lane) {
    if (lane == 0) {
        // formation: expand the internal slot of largest area (the first of
        // equal ones) until there are seven slots
        L.root_cost = T.cost[v];
        treelet_slot(T, L, 0, T.left[v]);
        treelet_slot(T, L, 1, T.right[v]);
        int m = 2, ninner = 0;
        while (m < kTreeletLeaves) {
            int best = -1;
            float best_area = -1.0f;
            for (int i = 0; i < m; i++) {
                if (L.ref[i] & kLeafBit) continue;
                if (L.area[i] > best_area) { best_area = L.area[i]; best = i; }
            }
            if (best < 0) break;
            const uint32_t c = L.ref[best];
            L.inner[ninner++] = c;
            const uint32_t cl = L.cl[best], cr = L.cr[best];
            treelet_slot(T, L, best, cl);
            treelet_slot(T, L, m++, cr);
        }
        L.m = m;
    }
    wave_lds_sync();
    const int m = L.m;
    if (m < 3) return;
    const int full = (1 << m) - 1;
    if (lane < m) L.cost[1 << lane] = L.scost[lane];
    // subset boxes in two steps: lanes 0..7 fold the boxes of the subsets
    // of slots 0..2, lanes 8..23 those of slots 3..6; then a subset's box is
    // the union of its two parts.  Minima and maxima are exact, so the
    // grouping changes no box.
    if ((lane >= 1 && lane < 8) || (lane >= 9 && lane < 24)) {   // the non-empty parts
        const int base = lane < 8 ? 0 : 3, bits = lane < 8 ? lane : lane - 8;
        const int b0 = base + __builtin_ctz(bits);
        float l[3], h[3];
#pragma unroll
        for (int k = 0; k < 3; k++) { l[k] = L.lb[b0][k]; h[k] = L.lb[b0][3 + k]; }
        for (int t = bits & (bits - 1); t; t &= t - 1) {
            const int b = base + __builtin_ctz(t);
#pragma unroll
            for (int k = 0; k < 3; k++) {
                l[k] = fminf(l[k], L.lb[b][k]);
                h[k] = fmaxf(h[k], L.lb[b][3 + k]);
            }
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
            L.pbox[lane][k] = l[k];
            L.pbox[lane][3 + k] = h[k];
        }
    }
    wave_lds_sync();
    // this lane's subsets s = lane + 1 and lane + 65: box and area
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int s = lane + 1 + 64 * j;
        if (s > full || (s & (s - 1)) == 0) continue;
        const int sa = s & 7, sb = s >> 3;   // one of them may be empty
        const float* a = L.pbox[sa ? sa : 8 + sb];
        const float* b = L.pbox[sb ? 8 + sb : sa];
        float l[3], h[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            l[k] = fminf(a[k], b[k]);
            h[k] = fmaxf(a[3 + k], b[3 + k]);
        }
        L.sarea[s] = box_area(l, h);
    }
    wave_lds_sync();
    // subsets by size: each one's best split into a left part holding its
    // lowest slot and a right part.  A subset of size k has 2^(k-1) - 1
    // splits; candidate t is the t-th of them in descending order of the
    // left part (the left part is the lowest slot plus the subset of the
    // other slots numbered 2^(k-1) - 2 - t).  A segment of 2^(k-3) lanes (at
    // least one) takes one subset, a lane four consecutive candidates, and
    // the best is the least finite cost, the lowest t among equal ones: the
    // first found by a scan in that order keeping a candidate only when
    // strictly cheaper than the best so far.
#pragma unroll
    for (int size = 2; size <= kTreeletLeaves; size++) {
        if (size > m) break;
        const int W = 1 << (size - 1);
        const int SL = W >= 4 ? W / 4 : 1, per = 64 / SL;  // lanes per subset, subsets per pass
        const int first = kSubsetStart.v[size], cnt = kSubsetStart.v[size + 1] - first;
        const int g = lane / SL, u = lane & (SL - 1);
        for (int b0 = 0; b0 < cnt; b0 += per) {
            const int idx = b0 + g;
            const int s = idx < cnt ? smask[first + idx] : 0;
            const bool live = idx < cnt && s <= full;
            const int low = s & -s, r = s ^ low;
            float bc = __builtin_inff();
            int bt = W;  // none
            if (live) {
                // candidates 4u .. 4u + 3 in order: a later one only when
                // strictly cheaper.  The other slots of candidate t + 1 are
                // the next smaller subset of r: (sub - 1) & r.
                int sub = dtab[(r >> 1) * 64 + (W - 2 - 4 * u)];
#pragma unroll
                for (int v = 0; v < 4; v++, sub = (sub - 1) & r) {
                    const int t = 4 * u + v;
                    if (t < W - 1) {
                        const int q = sub | low;
                        const float c = L.cost[q] + L.cost[s ^ q];
                        if (c < __builtin_inff() && (bt == W || c < bc)) { bc = c; bt = t; }
                    }
                }
            }
            for (int o = 1; o < SL; o <<= 1) {
                const float oc = __shfl_xor(bc, o);
                const int ot = __shfl_xor(bt, o);
                if (ot < W && (bt == W || oc < bc || (oc == bc && ot < bt))) { bc = oc; bt = ot; }
            }
            if (live && u == 0) {
                L.cost[s] = kCostNode * L.sarea[s] + bc;
                L.part[s] = (uint8_t)(bt < W ? (dtab[(r >> 1) * 64 + (W - 2 - bt)] | low) : 0);
            }
        }
        wave_lds_sync();
    }
    if (lane != 0 || !(L.cost[full] < L.root_cost)) return;
    // rewire: pre-order over the chosen splits, left part first
    int next = 0, sp = 0, nn = 0;
    L.st_node[sp] = v;
    L.st_set[sp++] = full;
    while (sp > 0) {
        --sp;
        const int64_t node = L.st_node[sp];
        const int set = L.st_set[sp];
        const int pl = L.part[set], pr = set ^ pl;
        const uint32_t rl = (pl & (pl - 1)) == 0 ? L.ref[__builtin_ctz(pl)] : L.inner[next++];
        const uint32_t rr = (pr & (pr - 1)) == 0 ? L.ref[__builtin_ctz(pr)] : L.inner[next++];
        T.left[node] = rl;
        T.right[node] = rr;
        T.parent[T.index(rl)] = (uint32_t)node;
        T.parent[T.index(rr)] = (uint32_t)node;
        L.order_node[nn] = node;
        L.order_set[nn++] = set;
        if ((pr & (pr - 1)) != 0) { L.st_node[sp] = rr; L.st_set[sp++] = pr; }
        if ((pl & (pl - 1)) != 0) { L.st_node[sp] = rl; L.st_set[sp++] = pl; }
    }
    // the rewired nodes, children first, joined from the slots' and each
    // other's values as the join of each node would compute them
    for (int i = nn - 1; i >= 0; i--) {
        const int set = L.order_set[i];
        const int pl = L.part[set], pr = set ^ pl;
        const bool sa = (pl & (pl - 1)) == 0, sc = (pr & (pr - 1)) == 0;
        const int ja = sa ? 0 : L.jat[pl], jc = sc ? 0 : L.jat[pr];
        const float* ba = sa ? L.lb[__builtin_ctz(pl)] : L.jbox[ja];
        const float* bc = sc ? L.lb[__builtin_ctz(pr)] : L.jbox[jc];
        const float ca = sa ? L.scost[__builtin_ctz(pl)] : L.jcost[ja];
        const float cc = sc ? L.scost[__builtin_ctz(pr)] : L.jcost[jc];
        const uint32_t na = sa ? L.snl[__builtin_ctz(pl)] : L.jnl[ja];
        const uint32_t nc = sc ? L.snl[__builtin_ctz(pr)] : L.jnl[jc];
        const int64_t node = L.order_node[i];
        float l[3], h[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            l[k] = fminf(ba[k], bc[k]);
            h[k] = fmaxf(ba[3 + k], bc[3 + k]);
            L.jbox[i][k] = l[k];
            L.jbox[i][3 + k] = h[k];
            T.lo[3 * node + k] = l[k];
            T.hi[3 * node + k] = h[k];
        }
        const float cst = kCostNode * box_area(l, h) + ca + cc;
        L.jcost[i] = cst;
        L.jnl[i] = na + nc;
        L.jat[set] = (uint8_t)i;
        T.nl[node] = na + nc;
        T.cost[node] = cst;
    }
}

// leaves: leaf count 1 and SAH cost kCostTri * area
__global__ void k_tree_leaves(int64_t n, TreeDev T) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t x = T.ni + k;
    float l[3], h[3];
#pragma unroll
    for (int a = 0; a < 3; a++) { l[a] = T.lo[3 * x + a]; h[a] = T.hi[3 * x + a]; }
    T.nl[x] = 1;
    T.cost[x] = kCostTri * box_area(l, h);
}

// Depths of the internal nodes by pointer jumping: anc = an ancestor (the
// root 0 its own), dep = the number of links to it.  A parent link out of
// range, or a node whose ancestor is not the root after kDepthJumps steps
// (a cycle or a depth of 2^13 or more), marks the hierarchy malformed.
__global__ void k_depth_init(int64_t ni, const uint32_t* __restrict__ parent, uint32_t* anc, uint32_t* dep,
                             uint32_t* error) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= ni) return;
    uint32_t p = 0;
    if (i > 0) {
        p = parent[i];
        if (p >= (uint32_t)ni) {
            atomicOr(error, 1u);
            p = 0;
        }
    }
    anc[i] = p;
    dep[i] = i > 0 ? 1u : 0u;
}
__global__ void k_depth_jump(int64_t ni, const uint32_t* __restrict__ anc, const uint32_t* __restrict__ dep,
                             uint32_t* anc2, uint32_t* dep2) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= ni) return;
    const uint32_t a = anc[i];
    anc2[i] = anc[a];
    dep2[i] = dep[i] + dep[a];
}
// The deepest depth as 64 partial maxima, one atomic per block.  A node
// whose ancestor is not yet the root adds bit 31 to its depth, so a maximum
// with that bit set asks for more jumps; after the last jumps (last != 0)
// it marks the hierarchy malformed.
__global__ __launch_bounds__(256) void k_depth_max(int64_t ni, const uint32_t* __restrict__ anc,
                                                   const uint32_t* __restrict__ dep, uint32_t* maxd, uint32_t* error,
                                                   int last) {
    __shared__ uint32_t wmax[256 / 64];
    uint32_t d = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ni; i += (int64_t)gridDim.x * blockDim.x)
        d = max(d, dep[i] | (anc[i] != 0 ? 0x80000000u : 0u));
    for (int o = 32; o > 0; o >>= 1) d = max(d, (uint32_t)__shfl_xor((int)d, o));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = d;
    __syncthreads();
    if (threadIdx.x == 0) {
        d = max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3]));
        atomicMax(&maxd[blockIdx.x % 64], d);
        if (last && (d >> 31)) atomicOr(error, 1u);
    }
}

// One depth of a bottom-up pass: the nodes at depth d join their children;
// with gamma > 0 those with gamma leaves or more below them are listed for
// treelet optimisation.  The list is kSubLists sub-lists of cap entries, one
// per block residue (block b appends to sub-list b % kSubLists with one
// atomic per block), so that no counter takes every block's atomic.
constexpr int kSubLists = 64;

__global__ __launch_bounds__(256) void k_tree_level(TreeDev T, const uint32_t* __restrict__ dep, uint32_t d,
                                                    uint32_t gamma, uint32_t* list, int64_t cap, uint32_t* count) {
    __shared__ uint32_t wcount[256 / 64], wbase[256 / 64];
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    bool need = false;
    if (i < T.ni && dep[i] == d) {
        T.join(i);
        need = gamma > 0 && T.nl[i] >= gamma;
    }
    if (gamma == 0) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t b = __ballot(need);
    if (lane == 0) wcount[wave] = (uint32_t)__popcll(b);
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t total = (wcount[0] + wcount[1]) + (wcount[2] + wcount[3]);
        uint32_t base = total ? atomicAdd(&count[blockIdx.x % kSubLists], total) : 0u;
        for (int k = 0; k < 256 / 64; k++) {
            wbase[k] = base;
            base += wcount[k];
        }
    }
    __syncthreads();
    if (need)
        list[(int64_t)(blockIdx.x % kSubLists) * cap + wbase[wave] + __popcll(b & ((1ull << lane) - 1))] =
            (uint32_t)i;
}

// The listed treelets of one depth, a wave each.
__global__ __launch_bounds__(256, 8) void k_treelets(TreeDev T, const uint32_t* __restrict__ list, int64_t cap,
                                                     const uint32_t* __restrict__ count) {
    __shared__ TreeletLds lds[256 / 64];
    __shared__ uint8_t smask[kTreeletSubsets];
    __shared__ __align__(16) uint8_t dtab[64 * 64];   // the deposit table
    if (threadIdx.x < kTreeletSubsets) smask[threadIdx.x] = kSubsets.mask[threadIdx.x];
    static_assert(sizeof(DepositTable) == 16 * 256, "a 16-byte piece per thread");
    reinterpret_cast<uint4*>(dtab)[threadIdx.x] = reinterpret_cast<const uint4*>(kDeposit.d)[threadIdx.x];
    __syncthreads();
    TreeletLds& L = lds[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63;
    static_assert(kSubLists == 64, "a lane per sub-list");
    const uint32_t c = count[lane];
    uint32_t inc = c;  // inclusive scan of the sub-list lengths
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = (uint32_t)__shfl_up((int)inc, o);
        if (lane >= o) inc += u;
    }
    const uint32_t total = (uint32_t)__shfl((int)inc, 63);
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    for (uint32_t j = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); j < total; j += waves) {
        const int sub = __builtin_ctzll(__ballot(inc > j));         // the sub-list holding entry j
        const uint32_t start = (uint32_t)__shfl((int)(inc - c), sub);
        treelet_wave(T, list[(int64_t)sub * cap + (j - start)], L, smask, dtab, lane);
    }
}

// Relabelling: the leaf position where every node's range starts (the
// leaves of the left siblings on its path from the root), then ids.
__global__ void k_first(int64_t n, const uint32_t* __restrict__ left, const uint32_t* __restrict__ right,
                        const uint32_t* __restrict__ parent, const uint32_t* __restrict__ nl, uint32_t* first,
                        uint32_t* error) {
    const int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t ni = n - 1;
    if (x >= 2 * n - 1) return;
    uint32_t f = 0;
    int64_t c = x;
    for (int depth = 0; c != 0; depth++) {
        const uint32_t p = parent[c];
        if (depth > 4096 || p >= (uint32_t)ni) {  // a malformed hierarchy
            atomicOr(error, 1u);
            break;
        }
        const uint32_t lc = left[p];
        const int64_t li = (lc & kLeafBit) ? ni + (lc & ~kLeafBit) : lc;
        if (li != c) f += nl[li];  // c is the right child
        c = p;
    }
    first[x] = f;
}

__global__ void k_relabel(int64_t n, const uint32_t* __restrict__ left, const uint32_t* __restrict__ right,
                          const uint32_t* __restrict__ parent, const uint32_t* __restrict__ nl,
                          const uint32_t* __restrict__ first, uint32_t* newid) {
    const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t ni = n - 1;
    if (v >= ni) return;
    if (v == 0) { newid[0] = 0; return; }
    const uint32_t p = parent[v];
    newid[v] = left[p] == (uint32_t)v ? first[v] + nl[v] - 1 : first[v];
}

__global__ void k_scatter_nodes(int64_t n, const uint32_t* __restrict__ left, const uint32_t* __restrict__ right,
                                const float* __restrict__ lo, const float* __restrict__ hi,
                                const uint32_t* __restrict__ first, const uint32_t* __restrict__ newid, uint32_t* L2,
                                uint32_t* R2, uint32_t* P2, float* lo2, float* hi2) {
    const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t ni = n - 1;
    if (v >= ni) return;
    auto remap = [&](uint32_t ref) -> uint32_t {
        return (ref & kLeafBit) ? (kLeafBit | first[ni + (ref & ~kLeafBit)]) : newid[ref];
    };
    const uint32_t w = newid[v];
    const uint32_t a = remap(left[v]), c = remap(right[v]);
    L2[w] = a;
    R2[w] = c;
    P2[(a & kLeafBit) ? ni + (a & ~kLeafBit) : a] = w;
    P2[(c & kLeafBit) ? ni + (c & ~kLeafBit) : c] = w;
#pragma unroll
    for (int k = 0; k < 3; k++) { lo2[3 * w + k] = lo[3 * v + k]; hi2[3 * w + k] = hi[3 * v + k]; }
}

__global__ void k_scatter_leaves(int64_t n, const float* __restrict__ lo, const float* __restrict__ hi,
                                 const uint32_t* __restrict__ order, const float4* __restrict__ tris,
                                 const uint32_t* __restrict__ first, float* lo2, float* hi2, uint32_t* order2,
                                 float4* tris2) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t ni = n - 1;
    if (k >= n) return;
    const int64_t q = first[ni + k];
    order2[q] = order[k];
#pragma unroll
    for (int i = 0; i < 3; i++) tris2[3 * q + i] = tris[3 * k + i];
#pragma unroll
    for (int a = 0; a < 3; a++) { lo2[3 * (ni + q) + a] = lo[3 * (ni + k) + a]; hi2[3 * (ni + q) + a] = hi[3 * (ni + k) + a]; }
}

__global__ void k_zero_words(uint32_t* w, int64_t m) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < m) w[i] = 0;
}

// entries per treelet sub-list: the nodes of the blocks of one residue
inline int64_t sublist_cap(int64_t n) { return ((n / 256 + 1) / kSubLists + 1) * 256; }

inline int blocks_for(int64_t n, int b) { return (int)((n + b - 1) / b); }

// Word offsets of the restructuring's arrays in the build scratch, after the
// sort's (each aligned to 64 words).
struct TreeScratch {
    size_t tsum, cost, nl, first, newid, L2, R2, P2, lo2, hi2, order2, tris2, anc, dep, anc2, dep2, tlist, tcount, maxd, end;
    explicit TreeScratch(int64_t n) {
        const int64_t nb = (n + kSortTile - 1) / kSortTile;
        auto al = [](size_t x) { return (x + 63) & ~(size_t)63; };
        const int64_t tiles = ((int64_t)kRadix * nb + kScanTile - 1) / kScanTile;
        size_t o = al((size_t)(4 * n + (int64_t)kRadix * nb + 8));
        tsum = o; o = al(o + tiles);
        cost = o; o = al(o + 2 * n);
        nl = o; o = al(o + 2 * n);
        first = o; o = al(o + 2 * n);
        newid = o; o = al(o + n);
        L2 = o; o = al(o + n);
        R2 = o; o = al(o + n);
        P2 = o; o = al(o + 2 * n);
        lo2 = o; o = al(o + 6 * n);
        hi2 = o; o = al(o + 6 * n);
        order2 = o; o = al(o + n);
        tris2 = o; o = al(o + 12 * n);
        anc = o; o = al(o + n);
        dep = o; o = al(o + n);
        anc2 = o; o = al(o + n);
        dep2 = o; o = al(o + n);
        tlist = o; o = al(o + (size_t)kSubLists * sublist_cap(n));
        tcount = o; o = al(o + ((size_t)kSubLists << kDepthJumps));
        maxd = o; o = al(o + 64);
        end = o;
    }
};

}  // namespace

uint32_t* bvh_error_word(void* scratch, int64_t n) {
    int64_t nb = (n + kSortTile - 1) / kSortTile;
    return static_cast<uint32_t*>(scratch) + 4 * n + (size_t)kRadix * nb + 6;
}

size_t bvh_scratch_bytes(int64_t n) {
    // keys/vals ping-pong (4n words) + histogram + bounds, then the tree
    // passes' costs, leaf counts, relabelled copies and depth arrays
    return TreeScratch(n).end * sizeof(uint32_t);
}

hipError_t bvh_build(const float* pos, int64_t n, BvhBuffers& o, void* scratch, hipStream_t s, int rounds) {
    if (n <= 0) return hipSuccess;
    uint32_t* w = static_cast<uint32_t*>(scratch);
    uint32_t* k0 = w;
    uint32_t* v0 = k0 + n;
    uint32_t* k1 = v0 + n;
    uint32_t* v1 = k1 + n;
    int nb = (int)((n + kSortTile - 1) / kSortTile);
    uint32_t* hist = v1 + n;
    uint32_t* bounds = hist + (size_t)kRadix * nb;

    k_init_bounds<<<1, 64, 0, s>>>(bounds);
    int gb = std::min(blocks_for(n, 256), 4096);
    k_centroid_bounds<<<gb, 256, 0, s>>>(pos, n, bounds);
    k_morton<<<blocks_for(n, 256), 256, 0, s>>>(pos, n, bounds, k0, v0);
    uint32_t *ki = k0, *vi = v0, *ko = k1, *vo = v1;
    const int64_t hm = (int64_t)kRadix * nb, tiles = (hm + kScanTile - 1) / kScanTile;
    uint32_t* tsum = w + TreeScratch(n).tsum;
    for (int shift = 0; shift < 30; shift += kRadixBits) {
        k_radix_hist<<<nb, kSortBlock, 0, s>>>(ki, n, shift, hist, nb);
        k_scan_reduce<<<(int)tiles, 256, 0, s>>>(hist, hm, tsum);
        k_scan_single<<<1, 1024, 0, s>>>(tsum, tiles);
        k_scan_apply<<<(int)tiles, 256, 0, s>>>(hist, hm, tsum);
        k_radix_scatter<<<nb, kSortBlock, 0, s>>>(ki, vi, n, shift, hist, nb, ko, vo);
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    hipError_t e = hipMemcpyAsync(o.codes, ki, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(o.order, vi, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return e;
    if (n > 1) k_karras<<<blocks_for(n - 1, 256), 256, 0, s>>>(o.codes, n, o.left, o.right, o.parent);
    k_leaf_boxes<<<blocks_for(n, 256), 256, 0, s>>>(pos, n, o.order, o.lo, o.hi, o.tris);
    if (n > 1) {
        const TreeScratch ts(n);
        const int64_t ni = n - 1;
        float* cost = reinterpret_cast<float*>(w + ts.cost);
        uint32_t* nl = w + ts.nl;
        uint32_t* error = bounds + 6;
        uint32_t* maxd = w + ts.maxd;
        TreeDev T{ni, o.left, o.right, o.parent, nl, o.lo, o.hi, cost};
        k_tree_leaves<<<blocks_for(n, 256), 256, 0, s>>>(n, T);
        // a refit pass (rounds 0), else one pass per restructuring round
        const int passes = (n > 2 && rounds > 0) ? rounds : 1;
        for (int r = 0; r < passes; r++) {
            uint32_t *anc = w + ts.anc, *dep = w + ts.dep, *anc2 = w + ts.anc2, *dep2 = w + ts.dep2;
            k_depth_init<<<blocks_for(ni, 256), 256, 0, s>>>(ni, o.parent, anc, dep, error);
            // six jumps reach the root from depths below 64; the other seven
            // run only when some node has not reached it
            constexpr int kFirstJumps = 6;
            int64_t depth = 0;
            for (int j0 = 0; j0 < kDepthJumps;) {
                const int j1 = j0 == 0 ? kFirstJumps : kDepthJumps;
                for (; j0 < j1; j0++) {
                    k_depth_jump<<<blocks_for(ni, 256), 256, 0, s>>>(ni, anc, dep, anc2, dep2);
                    std::swap(anc, anc2);
                    std::swap(dep, dep2);
                }
                k_zero_words<<<1, 64, 0, s>>>(maxd, 64);
                k_depth_max<<<std::min(blocks_for(ni, 256), 2048), 256, 0, s>>>(ni, anc, dep, maxd, error,
                                                                                j0 == kDepthJumps);
                uint32_t host[65] = {};   // error, then the 64 partial maxima
                if ((e = hipMemcpyAsync(host, error, 4, hipMemcpyDeviceToHost, s)) != hipSuccess ||
                    (e = hipMemcpyAsync(host + 1, maxd, 4 * 64, hipMemcpyDeviceToHost, s)) != hipSuccess ||
                    (e = hipStreamSynchronize(s)) != hipSuccess)
                    return e;
                if (host[0] != 0) return hipSuccess;   // malformed: the caller reads the error word
                depth = *std::max_element(host + 1, host + 65);
                if (!(depth >> 31)) break;
            }
            const uint32_t gamma = (n > 2 && rounds > 0) ? (uint32_t)kTreeletLeaves << r : 0u;
            const int64_t cap = sublist_cap(n);
            uint32_t *tlist = w + ts.tlist, *tcount = w + ts.tcount;
            if (gamma > 0) k_zero_words<<<blocks_for((depth + 1) * kSubLists, 256), 256, 0, s>>>(tcount, (depth + 1) * kSubLists);
            for (int64_t d = depth; d >= 0; d--) {
                uint32_t* cnt = tcount + d * kSubLists;
                k_tree_level<<<blocks_for(ni, 256), 256, 0, s>>>(T, dep, (uint32_t)d, gamma, tlist, cap, cnt);
                if (gamma > 0)
                    k_treelets<<<std::min(blocks_for(ni, 256), 2048), 256, 0, s>>>(T, tlist, cap, cnt);
            }
        }
    }
    if (n > 2 && rounds > 0) {
        const TreeScratch ts(n);
        uint32_t* nl = w + ts.nl;
        uint32_t *first = w + ts.first, *newid = w + ts.newid, *L2 = w + ts.L2, *R2 = w + ts.R2, *P2 = w + ts.P2,
                 *order2 = w + ts.order2;
        float *lo2 = reinterpret_cast<float*>(w + ts.lo2), *hi2 = reinterpret_cast<float*>(w + ts.hi2);
        float4* tris2 = reinterpret_cast<float4*>(w + ts.tris2);
        k_first<<<blocks_for(2 * n - 1, 256), 256, 0, s>>>(n, o.left, o.right, o.parent, nl, first, bounds + 6);
        k_relabel<<<blocks_for(n - 1, 256), 256, 0, s>>>(n, o.left, o.right, o.parent, nl, first, newid);
        k_scatter_nodes<<<blocks_for(n - 1, 256), 256, 0, s>>>(n, o.left, o.right, o.lo, o.hi, first, newid, L2, R2,
                                                                P2, lo2, hi2);
        k_scatter_leaves<<<blocks_for(n, 256), 256, 0, s>>>(n, o.lo, o.hi, o.order, o.tris, first, lo2, hi2, order2,
                                                             tris2);
        const size_t ni = (size_t)(n - 1);
        if ((e = hipMemcpyAsync(o.left, L2, 4 * ni, hipMemcpyDeviceToDevice, s)) != hipSuccess ||
            (e = hipMemcpyAsync(o.right, R2, 4 * ni, hipMemcpyDeviceToDevice, s)) != hipSuccess ||
            (e = hipMemcpyAsync(o.parent + 1, P2 + 1, 4 * (2 * n - 2), hipMemcpyDeviceToDevice, s)) != hipSuccess ||
            (e = hipMemcpyAsync(o.lo, lo2, 12 * (2 * n - 1), hipMemcpyDeviceToDevice, s)) != hipSuccess ||
            (e = hipMemcpyAsync(o.hi, hi2, 12 * (2 * n - 1), hipMemcpyDeviceToDevice, s)) != hipSuccess ||
            (e = hipMemcpyAsync(o.order, order2, 4 * n, hipMemcpyDeviceToDevice, s)) != hipSuccess ||
            (e = hipMemcpyAsync(o.tris, tris2, 48 * n, hipMemcpyDeviceToDevice, s)) != hipSuccess)
            return e;
    }
    if (n > 1) k_quantize<<<blocks_for(n - 1, 256), 256, 0, s>>>(n - 1, o.left, o.right, o.lo, o.hi, o.qnodes);
    return hipGetLastError();
}

}  // namespace oxr
