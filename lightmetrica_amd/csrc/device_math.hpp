// Device-side vector math, counter-based RNG and sampling warps for gfx950.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lmgpu {

struct V3 {
    float x, y, z;
};

__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 v3s(float s) { return V3{s, s, s}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float length(V3 v) { return sqrtf(dot(v, v)); }
__device__ __forceinline__ V3 normalize(V3 v) { return v * rsqrtf(dot(v, v)); }
__device__ __forceinline__ bool is_zero(V3 v) { return v.x == 0.f && v.y == 0.f && v.z == 0.f; }
__device__ __forceinline__ float luminance(V3 v) { return 0.212671f * v.x + 0.715160f * v.y + 0.072169f * v.z; }

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.318309886183790671538f;
constexpr float kEps = 1e-4f;       // Math::Constants::Eps() for float (math.constants.inl:63-67)
constexpr float kEpsLarge = 1e-3f;  // EpsLarge() for float (math.constants.inl:75-79)
constexpr float kInf = 3.402823466e+38f;  // Inf() = FLT_MAX (math.constants.inl:51-55)

// ---------------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
// 1, 2, 3", SC'11).  One stream per path sample: key = seed, counter =
// (dimension block, sample lo, sample hi, 'LMPT').  Dimension d is word d & 3 of block d >> 2.

constexpr uint32_t kStreamTag = 0x4c4d5054u;

__device__ __forceinline__ void philox4x32_10(uint32_t k0, uint32_t k1, uint32_t c0, uint32_t c1, uint32_t c2,
                                              uint32_t c3, uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const uint32_t lo0 = 0xD2511F53u * c0;
        const uint32_t hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2);
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// Sequential view of one sample's stream, mirroring Sampler::Next()/NextVec2() call order.
struct SampleStream {
    uint32_t k0, k1;
    uint32_t s_lo, s_hi;
    uint32_t dim;
    uint32_t cached_block;
    uint32_t cache[4];

    __device__ __forceinline__ void init(uint64_t seed, uint64_t sample, uint32_t d) {
        k0 = uint32_t(seed);
        k1 = uint32_t(seed >> 32);
        s_lo = uint32_t(sample);
        s_hi = uint32_t(sample >> 32);
        dim = d;
        cached_block = 0xffffffffu;
    }
    __device__ __forceinline__ uint32_t next_u32() {
        const uint32_t b = dim >> 2;
        if (b != cached_block) {
            philox4x32_10(k0, k1, b, s_lo, s_hi, kStreamTag, cache);
            cached_block = b;
        }
        const uint32_t w = dim & 3u;
        dim++;
        return w == 0 ? cache[0] : (w == 1 ? cache[1] : (w == 2 ? cache[2] : cache[3]));
    }
    // Random::Next: Float(u * 2^-32) rounded once (random.inl:24-27).  Converting u to
    // float with round-to-nearest and scaling by 2^-32 is the same single rounding.
    __device__ __forceinline__ float next() { return __uint2float_rn(next_u32()) * 2.3283064365386963e-10f; }
};

// ---------------------------------------------------------------------------------
// Sampling warps (include/lightmetrica/math.stats.inl)

// Shirley-Chiu concentric map (math.stats.inl:25-63)
__device__ __forceinline__ void concentric_disk(float u0, float u1, float& x, float& y) {
    const float v1 = 2.f * u0 - 1.f;
    const float v2 = 2.f * u1 - 1.f;
    float r, th;
    const float q = kPi / 4.f;
    if (v1 == 0.f && v2 == 0.f) {
        r = 0.f;
        th = 0.f;
    } else if (v1 > -v2) {
        if (v1 > v2) { r = v1; th = q * v2 / v1; }
        else { r = v2; th = q * (2.f - v1 / v2); }
    } else {
        if (v1 < v2) { r = -v1; th = q * (4.f + v2 / v1); }
        else { r = -v2; th = q * (6.f - v1 / v2); }
    }
    float s, c;
    sincosf(th, &s, &c);
    x = r * c;
    y = r * s;
}

// Malley's method (math.stats.inl:71-76)
__device__ __forceinline__ V3 cosine_hemisphere(float u0, float u1) {
    float x, y;
    concentric_disk(u0, u1, x, y);
    return v3(x, y, sqrtf(fmaxf(0.f, 1.f - x * x - y * y)));
}

// math.stats.inl:105-112
__device__ __forceinline__ V3 uniform_sphere(float u0, float u1) {
    const float z = 1.f - 2.f * u0;
    const float r = sqrtf(fmaxf(0.f, 1.f - z * z));
    float s, c;
    sincosf(2.f * kPi * u1, &s, &c);
    return v3(r * c, r * s, z);
}

// math.stats.inl:120-125
__device__ __forceinline__ void uniform_triangle(float u0, float u1, float& b0, float& b1) {
    const float s = sqrtf(fmaxf(0.f, u0));
    b0 = 1.f - s;
    b1 = u1 * s;
}

// ---------------------------------------------------------------------------------
// Shading frame: OrthonormalBasis(sn, ss, st) (math.linalgebra.inl:26-33) and the
// world/shading conversions of SurfaceGeometry::ComputeTangentSpace (surfacegeometry.h:54-62).
struct Frame {
    V3 ss, st, sn;
    __device__ __forceinline__ void build(V3 n) {
        sn = n;
        st = fabsf(n.x) > fabsf(n.y) ? normalize(v3(n.z, 0.f, -n.x)) : normalize(v3(0.f, n.z, -n.y));
        ss = normalize(cross(st, sn));
    }
    __device__ __forceinline__ V3 to_local(V3 w) const { return v3(dot(ss, w), dot(st, w), dot(sn, w)); }
    __device__ __forceinline__ V3 to_world(V3 l) const {
        return v3(ss.x * l.x + st.x * l.y + sn.x * l.z, ss.y * l.x + st.y * l.y + sn.y * l.z,
                  ss.z * l.x + st.z * l.y + sn.z * l.z);
    }
};

// 4x4 column-major matrix times (x, y, z, w)
struct M4 {
    float m[16];  // m[col * 4 + row]
};
__device__ __forceinline__ void mul4(const M4& M, float x, float y, float z, float w, float o[4]) {
#pragma unroll
    for (int r = 0; r < 4; r++) o[r] = M.m[r] * x + M.m[4 + r] * y + M.m[8 + r] * z + M.m[12 + r] * w;
}

}  // namespace lmgpu
