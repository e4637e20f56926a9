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
//   5. bottom-up AABB refit with per-node arrival counters; each finished box
//      is also written into its parent's 64-byte child-pair record used by
//      traversal.
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

// One thread per leaf: leaf box, sorted triangle record, then climb while
// this thread is the second to reach a node.
__global__ __launch_bounds__(256) void k_refit(const float* __restrict__ pos, int64_t n,
                                                const uint32_t* __restrict__ order, const uint32_t* __restrict__ left,
                                                const uint32_t* __restrict__ right, const uint32_t* __restrict__ parent,
                                                uint32_t* flags, float* lo, float* hi, float4* tris, uint32_t* error) {
    int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t ni = n - 1;
    uint32_t tid = order[k];
    const float* p = pos + 9 * (size_t)tid;
    float bl[3], bh[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        bl[a] = fminf(fminf(p[a], p[3 + a]), p[6 + a]);
        bh[a] = fmaxf(fmaxf(p[a], p[3 + a]), p[6 + a]);
    }
    tris[3 * k + 0] = make_float4(p[0], p[1], p[2], __uint_as_float(tid));
    tris[3 * k + 1] = make_float4(p[3], p[4], p[5], 0.0f);
    tris[3 * k + 2] = make_float4(p[6], p[7], p[8], 0.0f);
    int64_t node = ni + k;
#pragma unroll
    for (int a = 0; a < 3; a++) { lo[3 * node + a] = bl[a]; hi[3 * node + a] = bh[a]; }
    if (ni == 0) return;
    uint32_t par = parent[node];
    for (int depth = 0;; depth++) {
        if (depth > 256 || par >= (uint32_t)ni) {  // LBVH depth <= 30 + log2(n): a malformed tree
            atomicOr(error, 1u);
            return;
        }
        __threadfence();
        if (atomicAdd(&flags[par], 1u) == 0u) return;  // sibling not finished yet
        __threadfence();
        uint32_t cl = left[par], cr = right[par];
        int64_t il = (cl & kLeafBit) ? ni + (cl & ~kLeafBit) : cl;
        int64_t ir = (cr & kLeafBit) ? ni + (cr & ~kLeafBit) : cr;
        float nl[3], nh[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            float l0 = __hip_atomic_load(&lo[3 * il + a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            float l1 = __hip_atomic_load(&lo[3 * ir + a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            float h0 = __hip_atomic_load(&hi[3 * il + a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            float h1 = __hip_atomic_load(&hi[3 * ir + a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            nl[a] = fminf(l0, l1);
            nh[a] = fmaxf(h0, h1);
        }
#pragma unroll
        for (int a = 0; a < 3; a++) {
            __hip_atomic_store(&lo[3 * (size_t)par + a], nl[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&hi[3 * (size_t)par + a], nh[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (par == 0) return;
        par = parent[par];
    }
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

// ---- treelet restructuring (Karras & Aila 2013), then Karras relabelling ----
// Each round climbs the tree bottom up (the second thread to reach a node
// goes on, as in the refit); at a node it recomputes box, leaf count and SAH
// cost from the children and, if at least kTreeletLeaves << round leaves lie
// below it, optimises its treelet: the treelet leaf (internal) of largest
// area is replaced by its two children until there are seven, a dynamic
// program over the subsets of those finds the binary tree of least SAH cost,
// and if it is cheaper the treelet's internal nodes are rewired to it (the
// root keeps its id; the others are reused in the order they were expanded,
// depth first, left part first).  The tree is then relabelled as Karras's
// layout of its own leaf order: every node's leaves are consecutive, internal
// ids follow the endpoint rule (a left child takes the last position of its
// range, a right child the first), so a node's children are again gamma and
// gamma + 1.  Reads and writes of the tree during the climb are relaxed
// device-scope atomics, so a thread sees what the threads below it wrote.
constexpr int kTreeletLeaves = 7;
constexpr float kCostNode = 1.2f, kCostTri = 1.0f;

template <typename T>
__device__ __forceinline__ T ald(const T* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
template <typename T>
__device__ __forceinline__ void ast(T* p, T v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

__device__ __forceinline__ float box_area(const float lo[3], const float hi[3]) {
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return 2.0f * ((dx * dy + dy * dz) + dz * dx);
}

struct TreeDev {
    int64_t ni;
    uint32_t *left, *right, *parent, *nl;
    float *lo, *hi, *cost;
    __device__ int64_t index(uint32_t ref) const { return (ref & kLeafBit) ? ni + (ref & ~kLeafBit) : ref; }
    __device__ void box(int64_t x, float l[3], float h[3]) const {
#pragma unroll
        for (int k = 0; k < 3; k++) { l[k] = ald(&lo[3 * x + k]); h[k] = ald(&hi[3 * x + k]); }
    }
    // internal node v from its two children: box, leaf count and SAH cost
    __device__ void join(int64_t v) const {
        const int64_t a = index(ald(&left[v])), c = index(ald(&right[v]));
        float la[3], ha[3], lc[3], hc[3], l[3], h[3];
        box(a, la, ha);
        box(c, lc, hc);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            l[k] = fminf(la[k], lc[k]);
            h[k] = fmaxf(ha[k], hc[k]);
            ast(&lo[3 * v + k], l[k]);
            ast(&hi[3 * v + k], h[k]);
        }
        ast(&nl[v], ald(&nl[a]) + ald(&nl[c]));
        ast(&cost[v], kCostNode * box_area(l, h) + ald(&cost[a]) + ald(&cost[c]));
    }
};

__device__ void treelet(const TreeDev& T, int64_t v) {
    uint32_t ref[kTreeletLeaves], inner[kTreeletLeaves];
    int m = 2, ninner = 0;
    ref[0] = ald(&T.left[v]);
    ref[1] = ald(&T.right[v]);
    while (m < kTreeletLeaves) {
        int best = -1;
        float best_area = -1.0f;
        for (int i = 0; i < m; i++) {
            if (ref[i] & kLeafBit) continue;
            float l[3], h[3];
            T.box(ref[i], l, h);
            const float a = box_area(l, h);
            if (a > best_area) { best_area = a; best = i; }
        }
        if (best < 0) break;
        const uint32_t c = ref[best];
        inner[ninner++] = c;
        ref[best] = ald(&T.left[c]);
        ref[m++] = ald(&T.right[c]);
    }
    if (m < 3) return;
    const int full = (1 << m) - 1;
    float slo[1 << kTreeletLeaves][3], shi[1 << kTreeletLeaves][3], scost[1 << kTreeletLeaves];
    uint8_t part[1 << kTreeletLeaves];
    for (int s = 1; s <= full; s++) {
        if ((s & (s - 1)) == 0) {
            const int64_t x = T.index(ref[__builtin_ctz(s)]);
            T.box(x, slo[s], shi[s]);
            scost[s] = ald(&T.cost[x]);
            continue;
        }
        const int top = 31 - __builtin_clz((unsigned)s), rest = s & ~(1 << top);
        float l[3], h[3];
        T.box(T.index(ref[top]), l, h);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            slo[s][k] = fminf(slo[rest][k], l[k]);
            shi[s][k] = fmaxf(shi[rest][k], h[k]);
        }
        const int low = s & -s;
        float best = __builtin_inff();
        int bp = 0;
        for (int q = (s - 1) & s; q > 0; q = (q - 1) & s) {
            if (!(q & low)) continue;
            const float c = scost[q] + scost[s ^ q];
            if (c < best) { best = c; bp = q; }
        }
        scost[s] = kCostNode * box_area(slo[s], shi[s]) + best;
        part[s] = (uint8_t)bp;
    }
    if (!(scost[full] < ald(&T.cost[v]))) return;
    int next = 0;
    int64_t st_node[kTreeletLeaves], order_nodes[kTreeletLeaves];
    int st_set[kTreeletLeaves];
    int sp = 0, nn = 0;
    st_node[sp] = v;
    st_set[sp++] = full;
    while (sp > 0) {
        --sp;
        const int64_t node = st_node[sp];
        const int set = st_set[sp];
        const int pl = part[set], pr = set ^ pl;
        const uint32_t rl = (pl & (pl - 1)) == 0 ? ref[__builtin_ctz(pl)] : inner[next++];
        const uint32_t rr = (pr & (pr - 1)) == 0 ? ref[__builtin_ctz(pr)] : inner[next++];
        ast(&T.left[node], rl);
        ast(&T.right[node], rr);
        ast(&T.parent[T.index(rl)], (uint32_t)node);
        ast(&T.parent[T.index(rr)], (uint32_t)node);
        order_nodes[nn++] = node;
        if ((pr & (pr - 1)) != 0) { st_node[sp] = rr; st_set[sp++] = pr; }
        if ((pl & (pl - 1)) != 0) { st_node[sp] = rl; st_set[sp++] = pl; }
    }
    for (int i = nn - 1; i >= 0; i--) T.join(order_nodes[i]);
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

// one round: thread per leaf climbs; the second arrival joins the node and,
// with gamma leaves or more below it (gamma 0: join only), optimises its treelet
__global__ __launch_bounds__(256) void k_treelet_round(int64_t n, TreeDev T, uint32_t gamma, uint32_t* flags,
                                                       uint32_t* error) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint32_t node = ald(&T.parent[T.ni + k]);
    for (int depth = 0;; depth++) {
        if (depth > 4096 || node >= (uint32_t)T.ni) {
            atomicOr(error, 1u);
            return;
        }
        __threadfence();
        if (atomicAdd(&flags[node], 1u) == 0u) return;  // the other child's thread is not done yet
        __threadfence();
        T.join(node);
        if (gamma > 0 && ald(&T.nl[node]) >= gamma) treelet(T, node);
        __threadfence();
        if (node == 0) return;
        node = ald(&T.parent[node]);
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

inline int blocks_for(int64_t n, int b) { return (int)((n + b - 1) / b); }

// Word offsets of the restructuring's arrays in the build scratch, after the
// sort's (each aligned to 64 words).
struct TreeScratch {
    size_t cost, nl, first, newid, L2, R2, P2, lo2, hi2, order2, tris2, end;
    explicit TreeScratch(int64_t n) {
        const int64_t nb = (n + kSortTile - 1) / kSortTile;
        auto al = [](size_t x) { return (x + 63) & ~(size_t)63; };
        size_t o = al((size_t)(4 * n + (int64_t)kRadix * nb + 8 + n + 64));
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
        end = o;
    }
};

}  // namespace

uint32_t* bvh_error_word(void* scratch, int64_t n) {
    int64_t nb = (n + kSortTile - 1) / kSortTile;
    return static_cast<uint32_t*>(scratch) + 4 * n + (size_t)kRadix * nb + 6;
}

size_t bvh_scratch_bytes(int64_t n) {
    // keys/vals ping-pong (4n words) + histogram + bounds + flags, then the
    // restructuring's costs, leaf counts and relabelled copies
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
    uint32_t* flags = bounds + 8;

    k_init_bounds<<<1, 64, 0, s>>>(bounds);
    int gb = std::min(blocks_for(n, 256), 4096);
    k_centroid_bounds<<<gb, 256, 0, s>>>(pos, n, bounds);
    k_morton<<<blocks_for(n, 256), 256, 0, s>>>(pos, n, bounds, k0, v0);
    uint32_t *ki = k0, *vi = v0, *ko = k1, *vo = v1;
    for (int shift = 0; shift < 30; shift += kRadixBits) {
        k_radix_hist<<<nb, kSortBlock, 0, s>>>(ki, n, shift, hist, nb);
        k_scan_single<<<1, 1024, 0, s>>>(hist, (int64_t)kRadix * nb);
        k_radix_scatter<<<nb, kSortBlock, 0, s>>>(ki, vi, n, shift, hist, nb, ko, vo);
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    hipError_t e = hipMemcpyAsync(o.codes, ki, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(o.order, vi, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return e;
    if (n > 1) {
        k_karras<<<blocks_for(n - 1, 256), 256, 0, s>>>(o.codes, n, o.left, o.right, o.parent);
        e = hipMemsetAsync(flags, 0, (size_t)(n - 1) * sizeof(uint32_t), s);
        if (e != hipSuccess) return e;
    }
    k_refit<<<blocks_for(n, 256), 256, 0, s>>>(pos, n, o.order, o.left, o.right, o.parent, flags, o.lo, o.hi,
                                               o.tris, bounds + 6);
    if (n > 2 && rounds > 0) {
        const TreeScratch ts(n);
        float* cost = reinterpret_cast<float*>(w + ts.cost);
        uint32_t* nl = w + ts.nl;
        TreeDev T{n - 1, o.left, o.right, o.parent, nl, o.lo, o.hi, cost};
        k_tree_leaves<<<blocks_for(n, 256), 256, 0, s>>>(n, T);
        for (int r = 0; r < rounds; r++) {
            if ((e = hipMemsetAsync(flags, 0, (size_t)(n - 1) * sizeof(uint32_t), s)) != hipSuccess) return e;
            k_treelet_round<<<blocks_for(n, 256), 256, 0, s>>>(n, T, (uint32_t)kTreeletLeaves << r, flags, bounds + 6);
        }
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
