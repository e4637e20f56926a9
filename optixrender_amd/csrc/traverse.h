// BVH traversal on gfx950: the build's replacement for optixTrace
// (core/cuda/devicePrograms.cu:258-269 closest hit, :303-320/:389-405/:437-453
// any hit with TERMINATE_ON_FIRST_HIT).
//
// One ray per lane.  Each visited internal node costs one 32-byte fetch of
// its compressed record (2 x dwordx4: origin, children, per-axis scale and
// both children's boxes as 8-bit bounds; the vector-memory path, not the
// ALU, is what bounds this loop), leaves are tested at once in
// (left, right) order with a watertight ray/triangle test (Woop, Benthin,
// Wald 2013) on a 48-byte triangle record, then for a closest-hit ray the
// nearer hit internal child is descended and the farther pushed (any hit:
// the farther first).  The stack lives in LDS
// (S entries per lane, [entry][lane] so a wave touches consecutive banks) and
// spills the deeper entries to a per-lane HBM region.
// Closest-hit ties in t go to the lower triangle id, so the result does not
// depend on visit order.
#pragma once
#include "common.h"
#include "dmath.h"

namespace oxr {

struct RayPre {
    f3 o, inv;
    uint32_t axes;  // kx | ky << 2 | kz << 4: one register instead of three
    float Sx, Sy, Sz;
    OXR_D int kx() const { return (int)(axes & 3u); }
    OXR_D int ky() const { return (int)((axes >> 2) & 3u); }
    OXR_D int kz() const { return (int)((axes >> 4) & 3u); }
};

OXR_D float comp(f3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

OXR_D RayPre prep_ray(f3 o, f3 d) {
    RayPre r;
    r.o = o;
    // A zero component becomes +-1e-20 so that the slab products are never
    // 0 * inf = NaN for a ray lying in a box face.
    auto safe = [](float x) { return fabsf(x) > 1e-20f ? x : copysignf(1e-20f, x); };
    r.inv = f3(1.0f / safe(d.x), 1.0f / safe(d.y), 1.0f / safe(d.z));
    int kz = 0;
    if (fabsf(d.y) > fabsf(comp(d, kz))) kz = 1;
    if (fabsf(d.z) > fabsf(comp(d, kz))) kz = 2;
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    if (comp(d, kz) < 0.0f) { int t = kx; kx = ky; ky = t; }
    r.axes = (uint32_t)kx | (uint32_t)ky << 2 | (uint32_t)kz << 4;
    float dz = comp(d, kz);
    r.Sx = comp(d, kx) / dz;
    r.Sy = comp(d, ky) / dz;
    r.Sz = 1.0f / dz;
    return r;
}

// The sheared edge functions U, V, W, their sum det and the scaled distance
// T of the watertight test; the hit distance is T / det, the barycentrics
// V / det and W / det.
struct TriEdges {
    float V, W, det, T;
    bool inside;
};

OXR_D TriEdges tri_edges(const RayPre& r, float4 a, float4 b, float4 c) {
    f3 A = xyz(a) - r.o, B = xyz(b) - r.o, C = xyz(c) - r.o;
    const int kx = r.kx(), ky = r.ky(), kz = r.kz();
    float Akz = comp(A, kz), Bkz = comp(B, kz), Ckz = comp(C, kz);
    float Ax = comp(A, kx) - r.Sx * Akz, Ay = comp(A, ky) - r.Sy * Akz;
    float Bx = comp(B, kx) - r.Sx * Bkz, By = comp(B, ky) - r.Sy * Bkz;
    float Cx = comp(C, kx) - r.Sx * Ckz, Cy = comp(C, ky) - r.Sy * Ckz;
    float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    TriEdges e;
    e.V = V;
    e.W = W;
    e.inside = !((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f));
    e.det = U + V + W;
    e.T = U * (r.Sz * Akz) + V * (r.Sz * Bkz) + W * (r.Sz * Ckz);
    return e;
}

// Hit distance only: the traversal compares t and keeps the leaf; the
// barycentrics of the one hit it returns are derived afterwards by tri_uv
// from the same arithmetic, so two divisions run once per ray instead of
// once per candidate triangle (and never for an any-hit ray).
OXR_D bool tri_hit_t(const RayPre& r, float4 a, float4 b, float4 c, float tmin, float& t) {
    const TriEdges e = tri_edges(r, a, b, c);
    if (!e.inside || e.det == 0.0f) return false;
    t = e.T / e.det;
    return t > tmin;
}

OXR_D void tri_uv(const RayPre& r, float4 a, float4 b, float4 c, float& u, float& v) {
    const TriEdges e = tri_edges(r, a, b, c);
    u = e.V / e.det;
    v = e.W / e.det;
}

// Slab test of a box given by its six slab distances; the far distance is
// widened by 2 ulp.
OXR_D bool slab_test(float tx0, float tx1, float ty0, float ty1, float tz0, float tz1, float tmin, float tbest,
                     float& tn) {
    tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fminf(tz0, tz1));
    float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1)) * 1.00000024f;
    return tn <= tf && tf >= tmin && tn <= tbest;
}

// 8-bit bound k (0..3) of a word as a float (one v_cvt_f32_ubyte<k>)
OXR_D float qbyte(uint32_t w, int k) { return (float)((w >> (8 * k)) & 0xffu); }

// Both children of a compressed node record (layout in common.h): per axis
// the origin's distance A = (origin - o) * inv and the step's B = 2^e * inv;
// a bound q steps above the origin lies at fma(q, B, A).  A bound of 0 is the
// origin itself, so that face is tested exactly as an uncompressed box.
OXR_D void qnode_test(const RayPre& r, uint4 w0, uint4 w1, float tmin, float tbest, bool& h0, bool& h1, float& tn0,
                      float& tn1) {
    const float Ax = (__uint_as_float(w0.x) - r.o.x) * r.inv.x;
    const float Ay = (__uint_as_float(w0.y) - r.o.y) * r.inv.y;
    const float Az = (__uint_as_float(w0.z) - r.o.z) * r.inv.z;
    const float Bx = ldexpf(r.inv.x, (int)(int8_t)(w1.x & 0xffu));
    const float By = ldexpf(r.inv.y, (int)(int8_t)((w1.x >> 8) & 0xffu));
    const float Bz = ldexpf(r.inv.z, (int)(int8_t)((w1.x >> 16) & 0xffu));
    // bytes 0..11 of words y, z, w: child 0 lo.xyz hi.xyz, child 1 lo.xyz hi.xyz
    h0 = slab_test(fmaf(qbyte(w1.y, 0), Bx, Ax), fmaf(qbyte(w1.y, 3), Bx, Ax), fmaf(qbyte(w1.y, 1), By, Ay),
                   fmaf(qbyte(w1.z, 0), By, Ay), fmaf(qbyte(w1.y, 2), Bz, Az), fmaf(qbyte(w1.z, 1), Bz, Az), tmin,
                   tbest, tn0);
    h1 = slab_test(fmaf(qbyte(w1.z, 2), Bx, Ax), fmaf(qbyte(w1.w, 1), Bx, Ax), fmaf(qbyte(w1.z, 3), By, Ay),
                   fmaf(qbyte(w1.w, 2), By, Ay), fmaf(qbyte(w1.w, 0), Bz, Az), fmaf(qbyte(w1.w, 3), Bz, Az), tmin,
                   tbest, tn1);
}

struct Hit {
    float t;
    uint32_t leaf;  // sorted leaf position, kNoHit on miss
    uint32_t id;    // triangle id
    float u, v;
};

// A pointer into LDS.  Kept in its own address space so that the LDS and
// HBM halves of the stack stay distinct loads: with two generic pointers the
// compiler folds pop() into one flat load of a selected address, which
// waits on every outstanding memory operation of the wave.
typedef __attribute__((address_space(3))) uint32_t lds_u32;

// This lane's index in its wave, computed afresh at every use: the
// traversal loop runs at the 64-VGPR limit of 8 waves per SIMD, and a lane
// index (or a per-lane stack pointer) kept live across it gets spilled and
// reloaded from scratch around every stack access.  Two ALU instructions,
// no memory access.
OXR_D uint32_t lane_now() {
    uint32_t l;
    asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(l));
    return l;
}

// S entries per lane in LDS, the remaining kStackTotal - S in HBM.  Only
// wave-uniform bases are kept (in scalar registers); the lane's addresses are
// rebuilt from lane_now() at each access.
template <int S>
struct Stack {
    lds_u32* lds;       // the block's array: entry i of block lane t at i * kTraceBlock + t
    uint32_t* gbase;    // spill area: entry i of grid lane g at g + i * gstride
    uint32_t wave_lds;  // block lane of the wave's first lane
    uint32_t wave_glb;  // grid lane of the wave's first lane
    int64_t gstride;
    int sp;
    OXR_D Stack(uint32_t* block_lds, uint32_t* gbase_, int64_t gstride_)
        : lds((lds_u32*)block_lds), gbase(gbase_), gstride(gstride_), sp(0) {
        wave_lds = (uint32_t)__builtin_amdgcn_readfirstlane((int)threadIdx.x);
        wave_glb = blockIdx.x * blockDim.x + wave_lds;
    }
    OXR_D bool push(uint32_t x) {
        if (sp < S) lds[sp * kTraceBlock + wave_lds + lane_now()] = x;
        else if (sp < kStackTotal) gbase[(wave_glb + lane_now()) + (int64_t)(sp - S) * gstride] = x;
        else return false;
        sp++;
        return true;
    }
    OXR_D uint32_t pop() {
        sp--;
        if (sp < S) return lds[sp * kTraceBlock + wave_lds + lane_now()];
        return gbase[(wave_glb + lane_now()) + (int64_t)(sp - S) * gstride];
    }
};

// One lane's traversal between two node visits: the ray, the best hit so far
// and the record of the node to test next (its loads issued by the visit
// before).  traverse() runs a walk to its end; the trace kernel also keeps
// walks of many rays in flight in one wave, a node per lane at a time.
struct Walk {
    RayPre r;
    Hit h;
    uint4 w0, w1;
    float tmin, tmax;
    int budget;
};

// The triangle of sorted leaf k against the walk's ray; true when it is taken
// (any hit: it ends the walk).
template <bool ANY, bool COUNT>
OXR_D bool walk_leaf(const BvhView& b, Walk& w, uint32_t k, uint32_t& nt, int* overflow) {
    if (k >= (uint64_t)b.n) { *overflow = 2; return false; }
    const float4* tr = b.tris + 3 * (size_t)k;
    float4 a = tr[0], bb = tr[1], c = tr[2];
    if (COUNT) nt++;
    float t;
    if (!tri_hit_t(w.r, a, bb, c, w.tmin, t)) return false;
    uint32_t id = __float_as_uint(a.w);
    bool take = ANY ? (t < w.tmax) : (t < w.h.t || (t == w.h.t && w.h.leaf != kNoHit && id < w.h.id));
    if (take) { w.h.t = t; w.h.leaf = k; w.h.id = id; }
    return take;
}

// Starts a walk at the root; false when there is none to take (fewer than two
// triangles: the hit is then already final).
template <bool ANY, bool COUNT, class StackT>
OXR_D bool walk_start(const BvhView& b, Walk& w, f3 o, f3 d, float tmin, float tmax, StackT& st, uint32_t& nt,
                      int* overflow) {
    w.h.t = tmax;
    w.h.leaf = kNoHit;
    w.h.id = kNoHit;
    w.h.u = w.h.v = 0.0f;
    w.tmin = tmin;
    w.tmax = tmax;
    if (b.n <= 0) return false;
    w.r = prep_ray(o, d);
    if (b.n == 1) {
        walk_leaf<ANY, COUNT>(b, w, 0, nt, overflow);
        return false;
    }
    st.sp = 0;
    // A tree traversal visits each internal node at most once; the cap only
    // guards against a malformed hierarchy.
    w.budget = (int)b.n + 64 - 1;  // the root's visit
    w.w0 = b.qnodes[0];
    w.w1 = b.qnodes[1];
    return true;
}

// One node of a walk: tests the record in w0/w1, picks the next node and
// issues its record's loads before the triangle tests of this one's leaf
// children, so that the two fetches overlap; the nodes visited, their order
// and every hit are those of testing first and loading after.  False when
// the walk has ended.
template <bool ANY, bool COUNT, class StackT>
OXR_D bool walk_step(const BvhView& b, Walk& w, StackT& st, uint32_t& nn, uint32_t& nt, int* overflow) {
    if (COUNT) nn += 2;
    const uint4 w0 = w.w0, w1 = w.w1;
    const uint32_t g = w0.w & kQGamma;
    const uint32_t c0 = (w0.w & kQLeft) ? (kLeafBit | g) : g;
    const uint32_t c1 = (w0.w & kQRight) ? (kLeafBit | (g + 1)) : g + 1;
    float tn0, tn1;
    bool h0, h1;
    qnode_test(w.r, w0, w1, w.tmin, w.h.t, h0, h1, tn0, tn1);
    const bool l0 = h0 && (c0 & kLeafBit), l1 = h1 && (c1 & kLeafBit);
    h0 = h0 && !l0;
    h1 = h1 && !l1;
    // the next node: the nearer internal child (any hit: the farther -- a
    // shadow ray leaving a surface is mostly blocked by the far side of the
    // object it starts on, not by the boxes around its origin, about 40%
    // fewer node fetches on sun shadow rays; the answer does not depend on
    // the order), the only one, or the top of the stack
    uint32_t next = 0, far = 0;
    bool more = true, push = false;
    if (h0 && h1) {
        bool second = tn1 < tn0;
        if (ANY) second = !second;
        next = second ? c1 : c0;
        far = second ? c0 : c1;
        push = true;
    } else if (h0) {
        next = c0;
    } else if (h1) {
        next = c1;
    } else if (st.sp == 0) {
        more = false;
    } else {
        next = st.pop();
    }
    const bool bad = more && (--w.budget < 0 || next >= (uint32_t)(b.n - 1));
    if (more && !bad) {
        w.w0 = b.qnodes[2 * (size_t)next];
        w.w1 = b.qnodes[2 * (size_t)next + 1];
    }
    // Leaf children: child 0's triangle, then child 1's.  One block serves
    // every lane with a leaf on either side and a second only the lanes with
    // two, so a wave whose lanes reach leaves on different sides runs the
    // triangle test once instead of twice.
    if (l0 || l1) {
        if (walk_leaf<ANY, COUNT>(b, w, (l0 ? c0 : c1) & ~kLeafBit, nt, overflow) && ANY) return false;
        if (l0 && l1 && walk_leaf<ANY, COUNT>(b, w, c1 & ~kLeafBit, nt, overflow) && ANY) return false;
    }
    if (!more) return false;
    if (push && !st.push(far)) { *overflow = 1; return false; }
    if (bad) { *overflow = 2; return false; }
    return true;
}

// The ended walk's barycentrics of the closest hit, recomputed from its
// triangle record.
template <bool ANY>
OXR_D void walk_finish(const BvhView& b, Walk& w) {
    if (!ANY && w.h.leaf != kNoHit) {
        const float4* tr = b.tris + 3 * (size_t)w.h.leaf;
        tri_uv(w.r, tr[0], tr[1], tr[2], w.h.u, w.h.v);
    }
}

template <bool ANY, bool COUNT, class StackT>
OXR_D Hit traverse(const BvhView& b, f3 o, f3 d, float tmin, float tmax, StackT& st, uint32_t& nn, uint32_t& nt,
                   int* overflow) {
    Walk w;
    if (walk_start<ANY, COUNT>(b, w, o, d, tmin, tmax, st, nt, overflow))
        while (walk_step<ANY, COUNT>(b, w, st, nn, nt, overflow)) {
        }
    walk_finish<ANY>(b, w);
    return w.h;
}

}  // namespace oxr
