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
//     is scaled by 1 + 2*gamma(3) so rounding never culls a box the ray touches
//     (Ize, "Robust BVH Ray Traversal", JCGT 2013);
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
    s.Sx = pick3(r.dx, r.dy, r.dz, s.kx) / dkz;
    s.Sy = pick3(r.dx, r.dy, r.dz, s.ky) / dkz;
    s.Sz = 1.f / dkz;
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
    t = T / det;
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

struct Stack {
    uint32_t* lds;  // this block's LDS stack base
    int lane;       // thread index in the block
    int sp;
    bool overflow;
    uint32_t priv[kStackPriv];

    __device__ __forceinline__ void push(uint32_t v) {
        if (sp < kStackLDS) {
            lds[sp * kTraceBlock + lane] = v;
        } else if (sp < kStackTotal) {
            priv[sp - kStackLDS] = v;
        } else {
            overflow = true;  // unreachable for trees built with max_depth * 3 <= kStackTotal
            return;
        }
        sp++;
    }
    __device__ __forceinline__ uint32_t pop() {
        sp--;
        return sp < kStackLDS ? lds[sp * kTraceBlock + lane] : priv[sp - kStackLDS];
    }
};

constexpr float kBoxRobust = 1.0f + 2.0f * 1.8e-7f;  // 1 + 2*gamma(3), gamma(n) = n*u/(1-n*u), u = 2^-24

__device__ __forceinline__ void cswap(float& ka, uint32_t& va, float& kb, uint32_t& vb) {
    const bool s = kb < ka;
    const float tk = s ? kb : ka;
    const uint32_t tv = s ? vb : va;
    kb = s ? ka : kb;
    vb = s ? va : vb;
    ka = tk;
    va = tv;
}

// Core traversal.  kAny: stop at the first accepted triangle.  Returns true on a hit;
// for closest hit, h.t/h.rec/h.prim describe the nearest triangle.
template <bool kAny, bool kCount>
__device__ __forceinline__ bool traverse(const Node4* __restrict__ nodes, const TriRecord* __restrict__ tris,
                                         const RayIn& r, Stack& st, HitOut& h, TraceCounters& cnt) {
    const float ix = 1.f / (fabsf(r.dx) > 1e-30f ? r.dx : copysignf(1e-30f, r.dx));
    const float iy = 1.f / (fabsf(r.dy) > 1e-30f ? r.dy : copysignf(1e-30f, r.dy));
    const float iz = 1.f / (fabsf(r.dz) > 1e-30f ? r.dz : copysignf(1e-30f, r.dz));
    const Shear sh = make_shear(r);
    float tmax = r.tmax;
    const float tmin = r.tmin;
    bool found = false;
    h.prim = 0xffffffffu;
    h.rec = 0;
    h.t = tmax;

    st.sp = 0;
    uint32_t cur = 0;  // root node
    constexpr uint32_t kDone = 0xffffffffu;
    while (cur != kDone) {
        // Inner nodes: descend until a leaf reference is current, or the stack runs out.
        while (cur != kDone && !(cur & kLeafBit)) {
            if (kCount) cnt.nodes++;
            const float4* np = reinterpret_cast<const float4*>(nodes + cur);
            const float4 lox = np[0], hix = np[1], loy = np[2], hiy = np[3], loz = np[4], hiz = np[5];
            const uint4 ch = reinterpret_cast<const uint4*>(np)[6];
            const float lx[4] = {lox.x, lox.y, lox.z, lox.w}, hx[4] = {hix.x, hix.y, hix.z, hix.w};
            const float ly[4] = {loy.x, loy.y, loy.z, loy.w}, hy[4] = {hiy.x, hiy.y, hiy.z, hiy.w};
            const float lz[4] = {loz.x, loz.y, loz.z, loz.w}, hz[4] = {hiz.x, hiz.y, hiz.z, hiz.w};
            const uint32_t c[4] = {ch.x, ch.y, ch.z, ch.w};
            float key[4];
            uint32_t val[4];
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
            if (nhit == 0) {
                cur = st.sp > 0 ? st.pop() : kDone;
                continue;
            }
            if (!kAny) {
                cswap(key[0], val[0], key[1], val[1]);
                cswap(key[2], val[2], key[3], val[3]);
                cswap(key[0], val[0], key[2], val[2]);
                cswap(key[1], val[1], key[3], val[3]);
                cswap(key[1], val[1], key[2], val[2]);
                if (nhit > 3) st.push(val[3]);
                if (nhit > 2) st.push(val[2]);
                if (nhit > 1) st.push(val[1]);
                cur = val[0];
            } else {
                // Any order will do: compact the hit children, keep the first.
                uint32_t first = kDone;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    if (key[i] != INFINITY) {
                        if (first == kDone) first = val[i];
                        else st.push(val[i]);
                    }
                }
                cur = first;
            }
        }
        // Leaves: test their triangles, then continue with the next stack entry.
        while (cur != kDone && (cur & kLeafBit)) {
            const uint32_t first = cur & kMaxLeafFirst;
            const uint32_t count = ((cur >> 27) & 0xfu) + 1u;
            for (uint32_t j = 0; j < count; j++) {
                if (kCount) cnt.tris++;
                const float4* tp = reinterpret_cast<const float4*>(tris + first + j);
                const float4 v0 = tp[0], v1 = tp[1], v2 = tp[2];
                float t, U, V, W, det;
                if (watertight(sh, v0, v1, v2, tmin, tmax, t, U, V, W, det)) {
                    found = true;
                    tmax = t;
                    h.t = t;
                    h.rec = first + j;
                    h.prim = __float_as_uint(v0.w);
                    if (kAny) return true;
                }
            }
            cur = st.sp > 0 ? st.pop() : kDone;
        }
    }
    return found;
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
