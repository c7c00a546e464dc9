// Closest-hit and any-hit traversal of the 4-wide BVH (bvh_build.h) on gfx950.
//
// Replaces QBVHScene::IntersectTriangles (src/liblightmetrica/scene.qbvh.cpp:837-960):
//   * one ray per lane, 64-lane wavefronts;
//   * the node (6 SoA box planes x 4 children + 4 child words) is fetched as seven
//     16-byte loads from one 128-byte line;
//   * children hit by the ray are sorted by entry distance; the nearest is descended
//     into and the others are pushed (closest hit), or all are pushed (any hit);
//   * the traversal stack lives in LDS, [level][lane] so the 64 lanes of a wave hit
//     distinct banks, and spills to per-lane private memory beyond kStackLDS levels;
//   * boxes are tested with the reciprocal-direction slab test and the exit distance
//     is scaled by 1 + 2*gamma(4) so rounding never culls a box the ray touches
//     (Ize, "Robust BVH Ray Traversal", JCGT 2013; one more gamma term for the 1-ulp
//     hardware reciprocal);
//   * triangles use the watertight test of Woop, Benthin and Wald (JCGT 2013), with
//     the FP64 fallback for edge functions that evaluate to exactly zero.
// Acceptance follows the reference's QBVH: t strictly inside (tmin, tmax), the first
// triangle found wins an exact tie (scene.qbvh.cpp:124-133).
#pragma once

#include "bvh_build.h"
#include "device_math.hpp"

namespace lmgpu {

constexpr int kTraceBlock = 256;
constexpr int kStackLDS = 16;
constexpr int kStackPriv = 80;
constexpr int kStackTotal = kStackLDS + kStackPriv;  // supports 4-wide trees of depth <= 32

struct RayIn {
    float ox, oy, oz, dx, dy, dz, tmin, tmax;
};

struct HitOut {
    float t;
    uint32_t rec;  // leaf-order triangle record of the winner
    uint32_t prim; // scene-order triangle index, 0xffffffff = miss
};

struct TraceCounters {
    uint32_t nodes;
    uint32_t tris;
};

// Per-ray constants of the watertight test.
struct Shear {
    int kx, ky, kz;
    float Sx, Sy, Sz;
    float ox, oy, oz;  // origin permuted into (kx, ky, kz)
};

__device__ __forceinline__ float pick3(float a, float b, float c, int k) { return k == 0 ? a : (k == 1 ? b : c); }

__device__ __forceinline__ Shear make_shear(const RayIn& r) {
    Shear s;
    const float ax = fabsf(r.dx), ay = fabsf(r.dy), az = fabsf(r.dz);
    s.kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    s.kx = s.kz == 2 ? 0 : s.kz + 1;
    s.ky = s.kx == 2 ? 0 : s.kx + 1;
    const float dkz = pick3(r.dx, r.dy, r.dz, s.kz);
    if (dkz < 0.f) {  // keep the winding of the projected triangle
        int t = s.kx;
        s.kx = s.ky;
        s.ky = t;
    }
    // The shear only has to be the same for every triangle of the ray, so the hardware
    // reciprocal (1 ulp) serves.
    const float inv = __builtin_amdgcn_rcpf(dkz);
    s.Sx = pick3(r.dx, r.dy, r.dz, s.kx) * inv;
    s.Sy = pick3(r.dx, r.dy, r.dz, s.ky) * inv;
    s.Sz = inv;
    s.ox = pick3(r.ox, r.oy, r.oz, s.kx);
    s.oy = pick3(r.ox, r.oy, r.oz, s.ky);
    s.oz = pick3(r.ox, r.oy, r.oz, s.kz);
    return s;
}

// Watertight ray/triangle test.  Returns true with t strictly inside (tmin, tmax);
// U, V, W are the scaled barycentric weights of v0, v1, v2 and det = U + V + W.
__device__ __forceinline__ bool watertight(const Shear& s, float4 v0, float4 v1, float4 v2, float tmin, float tmax,
                                           float& t, float& U, float& V, float& W, float& det) {
#pragma clang fp contract(off)
    const float Ax = pick3(v0.x, v0.y, v0.z, s.kx) - s.ox;
    const float Ay = pick3(v0.x, v0.y, v0.z, s.ky) - s.oy;
    const float Az = pick3(v0.x, v0.y, v0.z, s.kz) - s.oz;
    const float Bx = pick3(v1.x, v1.y, v1.z, s.kx) - s.ox;
    const float By = pick3(v1.x, v1.y, v1.z, s.ky) - s.oy;
    const float Bz = pick3(v1.x, v1.y, v1.z, s.kz) - s.oz;
    const float Cx = pick3(v2.x, v2.y, v2.z, s.kx) - s.ox;
    const float Cy = pick3(v2.x, v2.y, v2.z, s.ky) - s.oy;
    const float Cz = pick3(v2.x, v2.y, v2.z, s.kz) - s.oz;
    const float ax = Ax - s.Sx * Az, ay = Ay - s.Sy * Az;
    const float bx = Bx - s.Sx * Bz, by = By - s.Sy * Bz;
    const float cx = Cx - s.Sx * Cz, cy = Cy - s.Sy * Cz;
    U = cx * by - cy * bx;
    V = ax * cy - ay * cx;
    W = bx * ay - by * ax;
    if (U == 0.f || V == 0.f || W == 0.f) {
        // Exact products in FP64 decide the sign on the edge (Woop et al., section 3.3).
        U = float(double(cx) * double(by) - double(cy) * double(bx));
        V = float(double(ax) * double(cy) - double(ay) * double(cx));
        W = float(double(bx) * double(ay) - double(by) * double(ax));
    }
    if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return false;
    det = U + V + W;
    if (det == 0.f) return false;
    const float T = U * (s.Sz * Az) + V * (s.Sz * Bz) + W * (s.Sz * Cz);
    t = T * __builtin_amdgcn_rcpf(det);  // within 3 ulp of T / det
    return t > tmin && t < tmax;
}

// The winner re-evaluated in FP64 (Moeller-Trumbore form): t and the weights of v1, v2.
__device__ __forceinline__ void refine_fp64(const RayIn& r, float4 v0, float4 v1, float4 v2, float& t, float& b0,
                                            float& b1) {
    const double ox = r.ox, oy = r.oy, oz = r.oz, dx = r.dx, dy = r.dy, dz = r.dz;
    const double p0x = v0.x, p0y = v0.y, p0z = v0.z;
    const double e1x = double(v1.x) - p0x, e1y = double(v1.y) - p0y, e1z = double(v1.z) - p0z;
    const double e2x = double(v2.x) - p0x, e2y = double(v2.y) - p0y, e2z = double(v2.z) - p0z;
    const double s1x = dy * e2z - dz * e2y, s1y = dz * e2x - dx * e2z, s1z = dx * e2y - dy * e2x;
    const double div = s1x * e1x + s1y * e1y + s1z * e1z;
    if (div == 0.0) return;
    const double inv = 1.0 / div;
    const double qx = ox - p0x, qy = oy - p0y, qz = oz - p0z;
    const double u = (qx * s1x + qy * s1y + qz * s1z) * inv;
    const double s2x = qy * e1z - qz * e1y, s2y = qz * e1x - qx * e1z, s2z = qx * e1y - qy * e1x;
    const double v = (dx * s2x + dy * s2y + dz * s2z) * inv;
    const double tt = (e2x * s2x + e2y * s2y + e2z * s2z) * inv;
    t = float(tt);
    b0 = float(u);
    b1 = float(v);
}

// 1 + 2 gamma(4), gamma(n) = n u / (1 - n u), u = 2^-24: (plane - o) and * inv round once each
// and the hardware reciprocal is within 1 ulp (2u).
constexpr float kBoxRobust = 1.0f + 2.0f * 2.4e-7f;

__device__ __forceinline__ void cswap(float& ka, uint32_t& va, float& kb, uint32_t& vb) {
    const bool s = kb < ka;
    const float tk = s ? kb : ka;
    const uint32_t tv = s ? vb : va;
    kb = s ? ka : kb;
    vb = s ? va : vb;
    ka = tk;
    va = tv;
}

constexpr uint32_t kDone = 0xffffffffu;

// Slab tests of the four children of inner node `cur`.  Returns how many the ray enters
// within [tmin, tmax]; with kSort the hit children come first in (key, val), nearest first.
template <bool kSort>
__device__ __forceinline__ int node_children(const Node4* __restrict__ nodes, uint32_t cur, const RayIn& r, float ix,
                                             float iy, float iz, float tmin, float tmax, float key[4],
                                             uint32_t val[4]) {
    const float4* np = reinterpret_cast<const float4*>(nodes + cur);
    const float4 lox = np[0], hix = np[1], loy = np[2], hiy = np[3], loz = np[4], hiz = np[5];
    const uint4 ch = reinterpret_cast<const uint4*>(np)[6];
    const float lx[4] = {lox.x, lox.y, lox.z, lox.w}, hx[4] = {hix.x, hix.y, hix.z, hix.w};
    const float ly[4] = {loy.x, loy.y, loy.z, loy.w}, hy[4] = {hiy.x, hiy.y, hiy.z, hiy.w};
    const float lz[4] = {loz.x, loz.y, loz.z, loz.w}, hz[4] = {hiz.x, hiz.y, hiz.z, hiz.w};
    const uint32_t c[4] = {ch.x, ch.y, ch.z, ch.w};
    int nhit = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const float t0x = (lx[i] - r.ox) * ix, t1x = (hx[i] - r.ox) * ix;
        const float t0y = (ly[i] - r.oy) * iy, t1y = (hy[i] - r.oy) * iy;
        const float t0z = (lz[i] - r.oz) * iz, t1z = (hz[i] - r.oz) * iz;
        const float tn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), tmin));
        const float tf = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fminf(fmaxf(t0z, t1z), tmax));
        const bool hit = c[i] != kEmptyChild && tn <= tf * kBoxRobust;
        key[i] = hit ? tn : INFINITY;
        val[i] = c[i];
        nhit += hit ? 1 : 0;
    }
    if (kSort) {
        cswap(key[0], val[0], key[1], val[1]);
        cswap(key[2], val[2], key[3], val[3]);
        cswap(key[0], val[0], key[2], val[2]);
        cswap(key[1], val[1], key[3], val[3]);
        cswap(key[1], val[1], key[2], val[2]);
    }
    return nhit;
}

// Triangles of leaf reference `cur`.  Accepted hits shrink tmax and update h; with kAny the
// test stops at the first accepted triangle.  Returns whether any triangle was accepted.
template <bool kAny, bool kCount>
__device__ __forceinline__ bool leaf_hits(const TriRecord* __restrict__ tris, uint32_t cur, const Shear& sh,
                                          float tmin, float& tmax, HitOut& h, TraceCounters& cnt) {
    const uint32_t first = cur & kMaxLeafFirst;
    const uint32_t count = ((cur >> 27) & 0xfu) + 1u;
    bool any = false;
    for (uint32_t j = 0; j < count; j++) {
        if (kCount) cnt.tris++;
        const float4* tp = reinterpret_cast<const float4*>(tris + first + j);
        const float4 v0 = tp[0], v1 = tp[1], v2 = tp[2];
        float t, U, V, W, det;
        if (watertight(sh, v0, v1, v2, tmin, tmax, t, U, V, W, det)) {
            any = true;
            tmax = t;
            h.t = t;
            h.rec = first + j;
            h.prim = __float_as_uint(v0.w);
            if (kAny) break;
        }
    }
    return any;
}

// Hardware reciprocal (v_rcp_f32, within 1 ulp) of a direction component kept away from 0.
__device__ __forceinline__ float safe_rcp(float d) {
    return __builtin_amdgcn_rcpf(fabsf(d) > 1e-30f ? d : copysignf(1e-30f, d));
}

// The per-lane traversal stack: levels [0, kStackLDS) live in LDS (`lds` is this lane's
// column of the block's __shared__ array, lds[level * kTraceBlock]), deeper levels in
// private memory.  The builder bounds 3 * depth <= kStackTotal, so the guard never fires.
// The spill levels are 64-bit slots so the compiler cannot merge the two loads of pop()
// into one generic (flat) load through a selected pointer.
#define LMGPU_TRACE_STACK(lds)                                                  \
    uint64_t spill[kStackPriv];                                                 \
    int sp = 0;                                                                 \
    auto push = [&](uint32_t v) {                                               \
        if (sp < kStackLDS) (lds)[sp * kTraceBlock] = v;                        \
        else if (sp < kStackTotal) spill[sp - kStackLDS] = v;                   \
        else return;                                                            \
        sp++;                                                                   \
    };                                                                          \
    auto pop = [&]() -> uint32_t {                                              \
        sp--;                                                                   \
        if (sp < kStackLDS) return (lds)[sp * kTraceBlock];                     \
        return uint32_t(spill[sp - kStackLDS]);                                 \
    };

// Descends from the node on top: push the farther hit children (or all but one for any
// hit) and return the next reference to visit.
template <bool kAny, class Push, class Pop>
__device__ __forceinline__ uint32_t descend(int nhit, const float key[4], const uint32_t val[4], int sp, Push& push,
                                            Pop& pop) {
    if (nhit == 0) return sp > 0 ? pop() : kDone;
    if (!kAny) {
        if (nhit > 3) push(val[3]);
        if (nhit > 2) push(val[2]);
        if (nhit > 1) push(val[1]);
        return val[0];
    }
    uint32_t first = kDone;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        if (key[i] != INFINITY) {
            if (first == kDone) first = val[i];
            else push(val[i]);
        }
    }
    return first;
}

// Core traversal of one ray.  kAny: stop at the first accepted triangle.  Returns true on
// a hit; for closest hit, h.t/h.rec/h.prim describe the nearest triangle.
template <bool kAny, bool kCount>
__device__ __forceinline__ bool traverse(const Node4* __restrict__ nodes, const TriRecord* __restrict__ tris,
                                         const RayIn& r, uint32_t* lds, HitOut& h, TraceCounters& cnt) {
    LMGPU_TRACE_STACK(lds)
    const float ix = safe_rcp(r.dx), iy = safe_rcp(r.dy), iz = safe_rcp(r.dz);
    const Shear sh = make_shear(r);
    float tmax = r.tmax;
    const float tmin = r.tmin;
    bool found = false;
    h.prim = 0xffffffffu;
    h.rec = 0;
    h.t = tmax;
    uint32_t cur = 0;  // root node
    while (cur != kDone) {
        // Inner nodes: descend until a leaf reference is current, or the stack runs out.
        while (cur != kDone && !(cur & kLeafBit)) {
            if (kCount) cnt.nodes++;
            float key[4];
            uint32_t val[4];
            const int nhit = node_children<!kAny>(nodes, cur, r, ix, iy, iz, tmin, tmax, key, val);
            cur = descend<kAny>(nhit, key, val, sp, push, pop);
        }
        // Leaves: test their triangles, then continue with the next stack entry.
        while (cur != kDone && (cur & kLeafBit)) {
            if (leaf_hits<kAny, kCount>(tris, cur, sh, tmin, tmax, h, cnt)) {
                found = true;
                if (kAny) return true;
            }
            cur = sp > 0 ? pop() : kDone;
        }
    }
    return found;
}

// Lanes of a persistent traversal fetch new rays once this many lanes of the wave are idle
// (a build-time tuning knob).
#ifndef LMGPU_REFILL_IDLE
#define LMGPU_REFILL_IDLE 40
#endif
constexpr int kRefillIdle = LMGPU_REFILL_IDLE;

// Persistent traversal with per-lane ray refill (Aila & Laine, "Understanding the
// efficiency of ray traversal on GPUs", HPG 2009): every wave of the grid loops over
// the batch.  The first 64 rays of a wave are assigned statically (wave w takes
// [64 w, 64 w + 64)); after that a lane whose ray has finished fetches the next unclaimed
// ray, with one atomic per wave on `next` (counting from 64 x the number of waves), once
// kRefillIdle lanes are idle.  A batch smaller than the grid therefore costs no atomics.
// Leaves are processed speculatively (same paper): a lane that reaches a leaf postpones it
// and keeps descending until every lane still descending holds a leaf, so the inner-node
// and triangle loops each run with as many lanes as possible.
// `src(i)` returns ray i, `sink(i, found, h, r)` consumes its result.  Every wave exits
// once the batch is exhausted and its own lanes are done.  With kCount, `cnt` sums the
// lane's node visits and triangle tests over all rays it traced.
template <bool kAny, bool kCount, class Source, class Sink>
__device__ __forceinline__ void trace_persistent(const Node4* __restrict__ nodes, const TriRecord* __restrict__ tris,
                                                 uint32_t n, uint32_t* next, uint32_t* lds, Source& src, Sink& sink,
                                                 TraceCounters& cnt) {
    LMGPU_TRACE_STACK(lds)
    constexpr uint32_t kNone = 0xffffffffu;
    const uint32_t lane = __lane_id();
    uint32_t id = kNone;
    RayIn r{0.f, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f};
    float ix = 0.f, iy = 0.f, iz = 0.f, tmax = 0.f;
    Shear sh = make_shear(r);
    HitOut h{0.f, 0u, kNone};
    bool found = false;
    uint32_t cur = kDone;   // next node or leaf reference to visit
    uint32_t leaf = kDone;  // postponed leaf
    auto start = [&](uint32_t k) {
        id = k;
        r = src(k);
        ix = safe_rcp(r.dx);
        iy = safe_rcp(r.dy);
        iz = safe_rcp(r.dz);
        sh = make_shear(r);
        tmax = r.tmax;
        h.prim = kNone;
        h.rec = 0;
        h.t = tmax;
        found = false;
        cur = 0;
        leaf = kDone;
        sp = 0;
    };
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    const uint32_t static_total = waves * 64u;
    {
        const uint32_t k = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 64u + lane;
        if (k < n) start(k);
    }
    bool exhausted = static_total >= n;  // wave-uniform
    bool first = true;
    for (;;) {
        if (!exhausted && !first) {
            const bool need = id == kNone;
            const unsigned long long mask = __ballot(need);
            const uint32_t cnt_need = uint32_t(__popcll(mask));
            if (cnt_need > 0) {
                const int leader = __ffsll((long long)mask) - 1;
                uint32_t base = 0;
                if (int(lane) == leader) base = atomicAdd(next, cnt_need);
                base = __shfl(base, leader, 64) + static_total;
                if (base + cnt_need >= n) exhausted = true;
                if (need) {
                    const uint32_t k = base + uint32_t(__popcll(mask & ((1ull << lane) - 1ull)));
                    if (k < n) start(k);
                }
            }
        }
        first = false;
        if (__ballot(id != kNone) == 0ull) break;
        for (;;) {
            // Inner nodes, until every lane still descending holds a postponed leaf.
            while (cur != kDone && !(cur & kLeafBit)) {
                if (kCount) cnt.nodes++;
                float key[4];
                uint32_t val[4];
                const int nhit = node_children<!kAny>(nodes, cur, r, ix, iy, iz, r.tmin, tmax, key, val);
                cur = descend<kAny>(nhit, key, val, sp, push, pop);
                if (cur != kDone && (cur & kLeafBit) && leaf == kDone) {
                    leaf = cur;
                    cur = sp > 0 ? pop() : kDone;
                }
                if (__ballot(leaf == kDone) == 0ull) break;
            }
            // Postponed leaves, then any leaf that is next in line.
            while (leaf != kDone) {
                if (leaf_hits<kAny, kCount>(tris, leaf, sh, r.tmin, tmax, h, cnt)) {
                    found = true;
                    if (kAny) {
                        cur = kDone;
                        leaf = kDone;
                        break;
                    }
                }
                leaf = kDone;
                if (cur != kDone && (cur & kLeafBit)) {
                    leaf = cur;
                    cur = sp > 0 ? pop() : kDone;
                }
            }
            if (id != kNone && cur == kDone) {
                sink(id, found, h, r);
                id = kNone;
            }
            const uint32_t idle = uint32_t(__popcll(__ballot(id == kNone)));
            if (idle == 64u || (!exhausted && idle >= uint32_t(kRefillIdle))) break;
        }
    }
}

// Barycentrics of the recorded winner (weights of v1 and v2), FP32 or FP64.
__device__ __forceinline__ void winner_barycentrics(const TriRecord* __restrict__ tris, const RayIn& r,
                                                    const HitOut& h, bool fp64, float& t, float& b0, float& b1) {
    const float4* tp = reinterpret_cast<const float4*>(tris + h.rec);
    const float4 v0 = tp[0], v1 = tp[1], v2 = tp[2];
    t = h.t;
    if (fp64) {
        refine_fp64(r, v0, v1, v2, t, b0, b1);
        return;
    }
    const Shear sh = make_shear(r);
    float tt, U, V, W, det;
    // Same arithmetic as during traversal, so the result is bit-identical to the accepted hit.
    watertight(sh, v0, v1, v2, -INFINITY, INFINITY, tt, U, V, W, det);
    const float inv = 1.f / det;
    b0 = V * inv;
    b1 = W * inv;
}

}  // namespace lmgpu
