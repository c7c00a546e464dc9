// Wavefront path tracer: the reference's per-sample loops restated as stage kernels
// over structure-of-arrays path queues in HBM.
//
//   gen     -- camera vertex of every path of a batch        (pathtrace.cpp:167-192,
//              pathtrace.direct.cpp:172-303 for the camera vertex of pt.direct / pt.mis)
//   extend  -- closest-hit traversal of the live queue        (Scene::Intersect, scene.cpp:67-72)
//   shade   -- emission, NEE, Russian roulette, BSDF sampling (pathtrace.cpp:203-262,
//              pathtrace.direct.cpp:188-336, pathtrace.mis.cpp:185-367); live paths are
//              compacted into the next queue with one atomic per wavefront
//   shadow  -- any-hit traversal of the NEE queue, then the film update
//              (RenderUtils::Visible, renderutils.cpp:58-70)
//   resolve -- film sums * W*H/N into the caller's film        (sched.mt.cpp:250-256)
//
// Each path sample owns a Philox4x32-10 stream keyed by (seed, global sample index) and
// draws from it in exactly the order the reference estimator calls Sampler::Next(), so
// the GPU traces the same path as the CPU restatement for the same sample index.
#include "context.hpp"
#include "traverse.hpp"

#include <algorithm>
#include <chrono>

namespace lmgpu {

constexpr int kPTBlock = 256;
constexpr uint32_t kMiss = 0xffffffffu;
constexpr uint32_t kBlackMat = 0xffffffffu;  // the black diffuse BSDF of the environment sphere
constexpr uint32_t kFlagDegenerate = 1u;     // previous vertex is the pinhole camera
constexpr uint32_t kFlagSpecular = 2u;       // previous BSDF sample was specular

// GeneralizedBSDFType bits (include/lightmetrica/generalizedbsdf.h:34-72)
constexpr int kDiffuseReflection = 1 << 0;
constexpr int kSpecularReflection = 1 << 2;

struct PTConsts {
    const Node4* nodes;
    const TriRecord* tris;
    const ShadeTri* shade;
    const MaterialDev* mats;
    const LightDev* lights;
    const LightTri* light_tris;
    uint32_t num_lights;
    CameraDev cam;
    EnvDev env;
    void* film;
    int film_fp64;
    uint64_t seed;
    uint64_t base;  // global index of the batch's first sample
    uint32_t batch;  // samples in this batch
    int max_pv;
    int rr_depth;
    int renderer;
};

// Path queue: structure of arrays of 16-byte records.
struct Queue {
    float4* ray_o;  // o.xyz, tmin
    float4* ray_d;  // d.xyz, pdf of the BSDF sample that produced the ray
    float4* thr;    // throughput.xyz
    float4* aux;    // pt: L.xyz;  pt.direct / pt.mis: shading normal of the ray origin
    uint4* meta;    // sample index in batch, vertex count n | flags << 16, stream dimension, pixel
};

struct ShadowQueue {
    float4* o;  // origin, tmax
    float4* d;  // direction, pixel (bits)
    float4* c;  // contribution
};

struct PTBuffers {
    Queue q[2];
    float4* hits;  // t, b0, b1, prim (bits)
    ShadowQueue sq;
    uint32_t* counters;  // [0, 64): live-queue sizes per segment, [64, 128): shadow sizes per segment,
                         // [128, 192): work counters of the extension launches, [192, 257): of the shadow launches
    unsigned long long* totals;  // closest rays, shadow rays
};

constexpr int kMaxSegments = 64;

// ---------------------------------------------------------------------------------

__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred) {
    const unsigned long long mask = __ballot(pred);
    const uint32_t cnt = uint32_t(__popcll(mask));
    const uint32_t lane = __lane_id();
    const uint32_t before = uint32_t(__popcll(mask & ((1ull << lane) - 1ull)));
    const int leader = cnt ? (__ffsll((long long)mask) - 1) : 0;
    uint32_t base = 0;
    if (cnt && int(lane) == leader) base = atomicAdd(counter, cnt);
    base = __shfl(base, leader, 64);
    return base + before;
}

// Block-aggregated queue append: one atomic per block of kPTBlock threads instead of one
// per wavefront (same-address atomics serialise in L2).  Every thread of the block must call
// it -- it synchronises the block -- and each call site of a kernel needs its own kSite.
template <int kSite>
__device__ __forceinline__ uint32_t block_append(uint32_t* counter, bool pred) {
    __shared__ uint32_t wave_off[kPTBlock / 64];
    __shared__ uint32_t base;
    const unsigned long long mask = __ballot(pred);
    const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
    if (lane == 0) wave_off[wave] = uint32_t(__popcll(mask));
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kPTBlock / 64; w++) {
            const uint32_t c = wave_off[w];
            wave_off[w] = t;
            t += c;
        }
        base = t ? atomicAdd(counter, t) : 0u;
    }
    __syncthreads();
    return base + wave_off[wave] + uint32_t(__popcll(mask & ((1ull << lane) - 1ull)));
}

__device__ __forceinline__ void film_add_raw(void* film, int fp64, uint32_t pixel, V3 v) {
    if (v.x == 0.f && v.y == 0.f && v.z == 0.f) return;
    if (fp64) {
        double* f = static_cast<double*>(film) + 3ull * pixel;
        unsafeAtomicAdd(f + 0, double(v.x));
        unsafeAtomicAdd(f + 1, double(v.y));
        unsafeAtomicAdd(f + 2, double(v.z));
    } else {
        float* f = static_cast<float*>(film) + 3ull * pixel;
        unsafeAtomicAdd(f + 0, v.x);
        unsafeAtomicAdd(f + 1, v.y);
        unsafeAtomicAdd(f + 2, v.z);
    }
}

__device__ __forceinline__ void film_add(const PTConsts& c, uint32_t pixel, V3 v) {
    if (v.x == 0.f && v.y == 0.f && v.z == 0.f) return;
    if (c.film_fp64) {
        double* f = static_cast<double*>(c.film) + 3ull * pixel;
        unsafeAtomicAdd(f + 0, double(v.x));
        unsafeAtomicAdd(f + 1, double(v.y));
        unsafeAtomicAdd(f + 2, double(v.z));
    } else {
        float* f = static_cast<float*>(c.film) + 3ull * pixel;
        unsafeAtomicAdd(f + 0, v.x);
        unsafeAtomicAdd(f + 1, v.y);
        unsafeAtomicAdd(f + 2, v.z);
    }
}

// HDRBitmapFilm pixel of a raster position (hdrfilm.cpp:160-166); y = 0 is the bottom row.
__device__ __forceinline__ uint32_t pixel_of(const CameraDev& cam, float rx, float ry) {
    const int px = min(max(int(rx * float(cam.width)), 0), cam.width - 1);
    const int py = min(max(int(ry * float(cam.height)), 0), cam.height - 1);
    return uint32_t(py) * uint32_t(cam.width) + uint32_t(px);
}

// PerspectiveCamera::EvaluateImportance (perspectivecamera.cpp:325-353)
__device__ __forceinline__ float cam_importance(const CameraDev& cam, float cos_theta) {
    if (cos_theta <= 0.f) return 0.f;
    if (fabsf(cos_theta - 1.f) < kEpsLarge) cos_theta = 1.f;
    const float ic = 1.f / cos_theta;
    return cam.inv_A * ic * ic * ic;
}

// PerspectiveCamera::SampleAndEstimateDirection (perspectivecamera.cpp:216-240): direction
// for a raster position; `pdf` is the importance in the projected-solid-angle measure.
__device__ __forceinline__ V3 cam_direction(const CameraDev& cam, float u0, float u1, float& pdf) {
    float d4[4];
    mul4(cam.inv_proj, u0 * 2.f - 1.f, u1 * 2.f - 1.f, 0.f, 1.f, d4);
    const V3 d3 = normalize(v3(d4[0] / d4[3], d4[1] / d4[3], d4[2] / d4[3]));
    float w4[4];
    mul4(cam.inv_view, d3.x, d3.y, d3.z, 0.f, w4);
    pdf = cam_importance(cam, -d3.z);
    return normalize(v3(w4[0], w4[1], w4[2]));
}

// PerspectiveCamera::RayToRasterPosition (perspectivecamera.cpp:304-323)
__device__ __forceinline__ bool cam_ray_to_raster(const CameraDev& cam, V3 p, V3 d, float& rx, float& ry) {
    float c4[4], n4[4];
    mul4(cam.view, p.x + d.x, p.y + d.y, p.z + d.z, 1.f, c4);
    mul4(cam.proj, c4[0], c4[1], c4[2], c4[3], n4);
    rx = (n4[0] / n4[3] + 1.f) / 2.f;
    ry = (n4[1] / n4[3] + 1.f) / 2.f;
    return !(rx < 0.f || rx > 1.f || ry < 0.f || ry > 1.f);
}

// PerspectiveCamera::EvaluateDirection / EvaluateDirectionPDF (perspectivecamera.cpp:276-302, :355-383)
__device__ __forceinline__ float cam_eval_direction(const CameraDev& cam, V3 p, V3 wo) {
    float c4[4], n4[4];
    mul4(cam.view, p.x + wo.x, p.y + wo.y, p.z + wo.z, 1.f, c4);
    mul4(cam.proj, c4[0], c4[1], c4[2], c4[3], n4);
    const float rx = (n4[0] / n4[3] + 1.f) / 2.f, ry = (n4[1] / n4[3] + 1.f) / 2.f;
    if (rx < 0.f || rx > 1.f || ry < 0.f || ry > 1.f) return 0.f;
    return cam_importance(cam, -normalize(v3(c4[0], c4[1], c4[2])).z);
}

// ---------------------------------------------------------------------------------
// Surfaces

struct Surf {
    V3 p, gn;
    Frame f;
    uint32_t mat;
    int32_t light;
};

// Scene::StoreIntersectionFromBarycentricCoords (scene.cpp:101-149) for a triangle hit,
// or the environment sphere when the triangles are missed (emittershape.sphere.cpp:117-176).
__device__ __forceinline__ bool surface_at(const PTConsts& c, V3 o, V3 d, float tmin, float4 hit, Surf& s) {
    const uint32_t prim = __float_as_uint(hit.w);
    if (prim != kMiss) {
        const float4* sp = reinterpret_cast<const float4*>(c.shade + prim);
        const float4 a = sp[0], b = sp[1], e = sp[2], g = sp[3];
        const float t = hit.x, b0 = hit.y, b1 = hit.z;
        s.p = o + d * t;
        s.gn = v3(a.x, a.y, a.z);
        s.mat = __float_as_uint(a.w);
        s.light = __float_as_int(b.w);
        const float w0 = 1.f - b0 - b1;
        const V3 n = v3(b.x, b.y, b.z) * w0 + v3(e.x, e.y, e.z) * b0 + v3(g.x, g.y, g.z) * b1;
        s.f.build(normalize(n));
        return true;
    }
    if (c.env.index < 0) return false;
    const V3 ctr = v3(c.env.center[0], c.env.center[1], c.env.center[2]);
    const V3 oc = o - ctr;
    const float A = dot(d, d), B = 2.f * dot(oc, d), C = dot(oc, oc) - c.env.radius * c.env.radius;
    const float det = B * B - 4.f * A * C;
    if (det < 0.f) return false;
    const float e = sqrtf(det), den = 2.f * A;
    const float t0 = (-B - e) / den, t1 = (-B + e) / den;
    if (t0 > kInf || t1 < tmin) return false;
    float t = t0;
    if (t < tmin) {
        t = t1;
        if (t > kInf) return false;
    }
    s.p = o + d * t;
    const V3 n = normalize(ctr - s.p);
    s.gn = n;
    s.f.build(n);
    s.mat = kBlackMat;
    s.light = c.env.index;
    return true;
}

// BSDF::ShadingNormalCorrectionFactor for EL transport (bsdf.h:57-83): 1 or 0.
__device__ __forceinline__ bool sides_agree(V3 gn, V3 wi, V3 wo, float lwi_z, float lwo_z) {
    return !(dot(wi, gn) * lwi_z <= 0.f || dot(wo, gn) * lwo_z <= 0.f);
}

struct MatInfo {
    uint32_t type;
    V3 R;
};
__device__ __forceinline__ MatInfo mat_info(const PTConsts& c, uint32_t m) {
    if (m == kBlackMat) return MatInfo{LMGPU_BSDF_DIFFUSE, v3s(0.f)};
    const float4 x = reinterpret_cast<const float4*>(c.mats)[m];
    return MatInfo{__float_as_uint(x.x), v3(x.y, x.z, x.w)};
}

// SampleAndEstimateDirection (EL): diffuse.cpp:133-157, mirror.cpp:84-107,
// diffusemirror.cpp:84-125.  Returns the estimate f/p; zero means the path ends.
__device__ __forceinline__ V3 bsdf_sample(const MatInfo& m, const Surf& s, V3 wi, float u0, float u1, float uc,
                                          V3& wo, float& pdf, int& type) {
    const V3 lwi = s.f.to_local(wi);
    if (lwi.z <= 0.f) return v3s(0.f);
    V3 lwo;
    V3 est;
    if (m.type == LMGPU_BSDF_DIFFUSE || (m.type == LMGPU_BSDF_DIFFUSEMIRROR && uc < 0.5f)) {
        lwo = cosine_hemisphere(u0, u1);
        type = kDiffuseReflection;
        pdf = m.type == LMGPU_BSDF_DIFFUSE ? kInvPi : kInvPi * 0.5f;
        est = m.type == LMGPU_BSDF_DIFFUSE ? m.R : m.R * 0.5f / 0.5f;
    } else {
        lwo = v3(-lwi.x, -lwi.y, lwi.z);  // ReflectZUp (math.vector.inl:532-536)
        type = kSpecularReflection;
        pdf = m.type == LMGPU_BSDF_MIRROR ? 1.f / lwo.z : 0.5f / fabsf(lwi.z);
        est = m.type == LMGPU_BSDF_MIRROR ? m.R : m.R * 0.5f / 0.5f;
    }
    wo = s.f.to_world(lwo);
    if (!sides_agree(s.gn, wi, wo, lwi.z, lwo.z)) return v3s(0.f);
    return est;
}

// EvaluateDirection / EvaluateDirectionPDF with type NonDelta (diffuse.cpp:192-220,
// mirror.cpp:141-215 -> 0, diffusemirror.cpp:188-274).
__device__ __forceinline__ V3 bsdf_eval_nondelta(const MatInfo& m, const Surf& s, V3 wi, V3 wo, float& pdf) {
    pdf = 0.f;
    if (m.type == LMGPU_BSDF_MIRROR) return v3s(0.f);
    const V3 lwi = s.f.to_local(wi), lwo = s.f.to_local(wo);
    if (lwi.z <= 0.f || lwo.z <= 0.f) return v3s(0.f);
    pdf = m.type == LMGPU_BSDF_DIFFUSE ? kInvPi : kInvPi * 0.5f;
    if (!sides_agree(s.gn, wi, wo, lwi.z, lwo.z)) return v3s(0.f);
    return m.type == LMGPU_BSDF_DIFFUSE ? m.R * kInvPi : m.R * kInvPi * 0.5f;
}

__device__ __forceinline__ bool mat_nondelta(const MatInfo& m) { return m.type != LMGPU_BSDF_MIRROR; }

// Emission toward -d at a surface: EvaluateDirection * EvaluatePosition = (1/pi) * (Le pi)
// on the front side (light.area.cpp:233-252, light.env.const.cpp:174-209).
__device__ __forceinline__ V3 light_Le(const PTConsts& c, int li) {
    const float4 x = reinterpret_cast<const float4*>(c.lights)[2 * li];
    return v3(x.y, x.z, x.w);
}

// ---------------------------------------------------------------------------------
// Next-event estimation at a vertex (pathtrace.direct.cpp:197-248, pathtrace.mis.cpp:194-262).

struct NEEVertex {
    V3 p;
    Surf s;           // valid when !camera
    bool camera;      // the pinhole camera vertex (degenerated, no cosine)
    MatInfo m;
    V3 wi;
};

struct ShadowRec {
    bool emit;
    float4 o, d, c;
};

// Returns the shadow-ray record of the connection (emit = false when the contribution is 0;
// the caller appends emitted records to the shadow queue).
__device__ __forceinline__ ShadowRec nee(const PTConsts& c, const NEEVertex& v, V3 T, int n, uint32_t pixel,
                                         SampleStream& rng, bool mis, bool active) {
    V3 contrib = v3s(0.f);
    V3 target = v.p;
    uint32_t pix = pixel;
    if (active) {
        // Scene::SampleLightSelection (scene.cpp:79-86)
        float lu0 = rng.next(), lu1 = rng.next();
        const int nl = int(c.num_lights);
        const int li = min(int(lu0 * float(nl)), nl - 1);
        lu0 = lu0 * float(nl) - float(li);
        const float sel_pdf = 1.f / float(nl);
        const float4 L0 = reinterpret_cast<const float4*>(c.lights)[2 * li];
        const float4 L1 = reinterpret_cast<const float4*>(c.lights)[2 * li + 1];
        const uint32_t ltype = __float_as_uint(L0.x);
        const V3 Le = v3(L0.y, L0.z, L0.w);
        V3 lp, ln;
        float pdf_pl;
        if (ltype == LMGPU_LIGHT_ENV_CONST) {  // light.env.const.cpp:196-204
            const V3 dir = uniform_sphere(lu0, lu1);
            lp = v3(c.env.center[0], c.env.center[1], c.env.center[2]) + dir * c.env.radius;
            ln = -dir;
            pdf_pl = c.env.inv_area;
        } else {  // light.area.cpp:146-179
            const uint32_t tb = __float_as_uint(L1.x), tc = __float_as_uint(L1.y);
            pdf_pl = L1.z;
            // upper_bound over the cumulative area fractions
            uint32_t lo = 0, hi = tc;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (c.light_tris[tb + mid].p2[3] <= lu1) lo = mid + 1;
                else hi = mid;
            }
            const uint32_t k = min(lo, tc - 1);
            const LightTri lt = c.light_tris[tb + k];
            const float ylo = lt.p1[3], yhi = lt.p2[3];
            const float y = (lu1 - ylo) / (yhi - ylo);
            float b0, b1;
            uniform_triangle(lu0, y, b0, b1);
            const V3 p1 = v3(lt.p1[0], lt.p1[1], lt.p1[2]), p2 = v3(lt.p2[0], lt.p2[1], lt.p2[2]),
                     p3 = v3(lt.p3[0], lt.p3[1], lt.p3[2]);
            lp = p1 * (1.f - b0 - b1) + p2 * b0 + p3 * b1;
            ln = normalize(cross(p2 - p1, p3 - p1));
        }
        pdf_pl *= sel_pdf;
        const V3 ppl = normalize(lp - v.p);
        bool vis = true;
        if (n == 1) {
            float rx, ry;
            vis = cam_ray_to_raster(c.cam, v.p, ppl, rx, ry);
            pix = pixel_of(c.cam, rx, ry);
        }
        if (vis) {
            V3 fsE;
            float pdf_bsdf;
            if (v.camera) {
                const float w = cam_eval_direction(c.cam, v.p, ppl);
                fsE = v3s(w);
                pdf_bsdf = w;
            } else {
                fsE = bsdf_eval_nondelta(v.m, v.s, v.wi, ppl, pdf_bsdf);
            }
            // Light frame: shading normal = geometric normal (light.area.cpp:170-172).
            const float fsL = dot(ln, -ppl) <= 0.f ? 0.f : kInvPi;
            // RenderUtils::GeneralizedGeometryTerm (renderutils.cpp:29-56)
            V3 dv = lp - v.p;
            const float l2 = dot(dv, dv);
            const float l = sqrtf(l2);
            dv = dv / l;
            float num = 1.f;
            if (!v.camera) num *= fabsf(dot(v.s.f.sn, dv));
            num *= fabsf(dot(ln, -dv));
            const float G = fabsf(l) < kEps ? 0.f : num / l2;
            const V3 LeP = Le * kPi;
            if (!mis) {
                contrib = T * fsE * G * fsL * LeP / pdf_pl;
            } else if (G != 0.f) {
                const float pdf_dl = pdf_pl / G;
                const float w = pdf_dl / (pdf_dl + pdf_bsdf);
                contrib = T * w * fsE * G * fsL * LeP / pdf_pl;
            }
            target = lp;
        }
    }
    ShadowRec rec;
    rec.emit = active && !is_zero(contrib);
    if (rec.emit) {
        // RenderUtils::Visible: d = (p2 - p1) / |p2 - p1|, (Eps, |p2 - p1| (1 - Eps))
        const V3 dv = target - v.p;
        const float len = length(dv);
        const V3 d = dv / len;
        rec.o = make_float4(v.p.x, v.p.y, v.p.z, len * (1.f - kEps));
        rec.d = make_float4(d.x, d.y, d.z, __uint_as_float(pix));
        rec.c = make_float4(contrib.x, contrib.y, contrib.z, 0.f);
    }
    return rec;
}

__device__ __forceinline__ void put_shadow(const ShadowQueue& sq, uint32_t slot, const ShadowRec& r) {
    sq.o[slot] = r.o;
    sq.d[slot] = r.d;
    sq.c[slot] = r.c;
}

// ---------------------------------------------------------------------------------
// Kernels

__device__ __forceinline__ void store_path(const Queue& q, uint32_t slot, V3 o, float tmin, V3 d, float pdf, V3 T,
                                           V3 aux, uint32_t sid, uint32_t n_flags, uint32_t dim, uint32_t pixel) {
    q.ray_o[slot] = make_float4(o.x, o.y, o.z, tmin);
    q.ray_d[slot] = make_float4(d.x, d.y, d.z, pdf);
    q.thr[slot] = make_float4(T.x, T.y, T.z, 0.f);
    q.aux[slot] = make_float4(aux.x, aux.y, aux.z, 0.f);
    q.meta[slot] = make_uint4(sid, n_flags, dim, pixel);
}

// Camera vertex.  pt: pathtrace.cpp:167-192.  pt.direct / pt.mis: the first loop
// iteration of pathtrace.direct.cpp:172-303 (NEE from the camera, RR, camera sampling).
__global__ __launch_bounds__(kPTBlock) void k_gen(PTConsts c, Queue out, uint32_t* live_count, ShadowQueue sq,
                                                  uint32_t* shadow_count) {
    const uint32_t i = blockIdx.x * kPTBlock + threadIdx.x;
    const bool valid = i < c.batch;
    SampleStream rng;
    rng.init(c.seed, c.base + i, 0);
    const V3 cam_p = v3(c.cam.pos[0], c.cam.pos[1], c.cam.pos[2]);
    const bool start = !(c.max_pv != -1 && 1 >= c.max_pv);
    if (c.renderer == LMGPU_PT) {
        // Every path of pt survives its camera vertex: slot = sample index, no compaction.
        if (!valid || !start) return;
        if (i == 0) *live_count = c.batch;
        const float r0 = rng.next(), r1 = rng.next();
        rng.next();  // Camera::SamplePosition draws a Vec2 (pathtrace.cpp:173)
        rng.next();
        float pdf;
        const V3 d = cam_direction(c.cam, r0, r1, pdf);
        store_path(out, i, cam_p, 0.f, d, pdf, v3s(1.f), v3s(0.f), i, 1u, rng.dim, pixel_of(c.cam, r0, r1));
        return;
    }
    const bool mis = c.renderer == LMGPU_PT_MIS;
    rng.next();  // Camera::SamplePosition (pathtrace.direct.cpp:175)
    rng.next();
    V3 T = v3s(1.f);
    bool alive = valid && start;
    NEEVertex v;
    v.p = cam_p;
    v.camera = true;
    v.wi = v3s(0.f);
    const ShadowRec rec = nee(c, v, T, 1, 0u, rng, mis, alive);
    const uint32_t sslot = block_append<0>(shadow_count, rec.emit);
    if (rec.emit) put_shadow(sq, sslot, rec);
    if (alive && c.rr_depth != -1 && 1 >= c.rr_depth) {
        const float q = fminf(0.5f, luminance(T));
        if (rng.next() > q) alive = false;
        else T = T / q;
    }
    V3 d = v3s(0.f);
    float pdf = 0.f;
    uint32_t pixel = 0;
    if (alive) {
        const float u0 = rng.next(), u1 = rng.next();
        rng.next();  // uComp
        d = cam_direction(c.cam, u0, u1, pdf);
        float rx, ry;
        if (!cam_ray_to_raster(c.cam, cam_p, d, rx, ry)) alive = false;
        pixel = pixel_of(c.cam, rx, ry);
    }
    const uint32_t slot = block_append<1>(live_count, alive);
    if (alive)
        store_path(out, slot, cam_p, kEps, d, pdf, T, v3s(0.f), i, 1u | (kFlagDegenerate << 16), rng.dim, pixel);
}

// Extension rays: closest hit of every live path (persistent, refilling lanes).
struct QueueSource {
    Queue q;
    __device__ __forceinline__ RayIn operator()(uint32_t i) const {
        const float4 o = q.ray_o[i], d = q.ray_d[i];
        return RayIn{o.x, o.y, o.z, d.x, d.y, d.z, o.w, kInf};
    }
};

struct HitSink {
    const TriRecord* tris;
    float4* hits;
    __device__ __forceinline__ void operator()(uint32_t i, bool found, const HitOut& h, const RayIn& r) const {
        float4 out = make_float4(0.f, 0.f, 0.f, __uint_as_float(kMiss));
        if (found) {
            float t, b0, b1;
            winner_barycentrics(tris, r, h, false, t, b0, b1);
            out = make_float4(t, b0, b1, __uint_as_float(h.prim));
        }
        hits[i] = out;
    }
};

__global__ __launch_bounds__(kTraceBlock) void k_extend(const Node4* __restrict__ nodes,
                                                        const TriRecord* __restrict__ tris, Queue q,
                                                        const uint32_t* __restrict__ count, uint32_t* next,
                                                        float4* __restrict__ hits) {
    __shared__ uint32_t lds_stack[kStackLDS * kTraceBlock];
    const uint32_t n = *count;
    QueueSource src{q};
    HitSink sink{tris, hits};
    TraceCounters cnt{0, 0};
    trace_persistent<false, false>(nodes, tris, n, next, lds_stack + threadIdx.x, src, sink, cnt);
}

// pt: pathtrace.cpp:203-262 for one segment.
__global__ __launch_bounds__(kPTBlock) void k_shade_pt(PTConsts c, Queue in, const uint32_t* __restrict__ in_count,
                                                       const float4* __restrict__ hits, Queue out,
                                                       uint32_t* out_count) {
    const uint32_t i = blockIdx.x * kPTBlock + threadIdx.x;
    const bool valid = i < *in_count;
    const uint32_t j = valid ? i : 0u;  // invalid lanes read slot 0 and discard it
    const float4 ro = in.ray_o[j], rd = in.ray_d[j], th = in.thr[j], ax = in.aux[j];
    const uint4 meta = in.meta[j];
    const V3 o = v3(ro.x, ro.y, ro.z), d = v3(rd.x, rd.y, rd.z);
    V3 T = v3(th.x, th.y, th.z), L = v3(ax.x, ax.y, ax.z);
    int n = int(meta.y & 0xffffu);
    bool alive = false;
    Surf s;
    V3 wo = v3s(0.f);
    float pdf = 0.f;
    SampleStream rng;
    rng.init(c.seed, c.base + meta.x, meta.z);
    if (valid && surface_at(c, o, d, ro.w, hits[j], s)) {
        if (s.light >= 0 && dot(s.f.sn, -d) > 0.f) L = L + T * v3s(kInvPi) * (light_Le(c, s.light) * kPi);
        const float u0 = rng.next(), u1 = rng.next(), uc = rng.next();
        int type;
        const V3 fs = bsdf_sample(mat_info(c, s.mat), s, -d, u0, u1, uc, wo, pdf, type);
        if (!is_zero(fs)) {
            T = T * fs;
            alive = true;
            if (c.rr_depth != -1 && n >= c.rr_depth) {
                const float q = fminf(0.5f, luminance(T));
                if (rng.next() > q) alive = false;
                else T = T / q;
            }
            n++;
            if (c.max_pv != -1 && n >= c.max_pv) alive = false;
        }
    }
    const uint32_t slot = block_append<0>(out_count, alive);
    if (alive) store_path(out, slot, s.p, kEps, wo, pdf, T, L, meta.x, uint32_t(n), rng.dim, meta.w);
    else if (valid) film_add(c, meta.w, L);
}

// pt.direct / pt.mis: the hit half of one loop iteration (pathtrace.direct.cpp:304-336,
// pathtrace.mis.cpp:318-367) followed by the next iteration up to the next ray
// (NEE :197-248 / :194-262, RR, BSDF sampling).
__global__ __launch_bounds__(kPTBlock) void k_shade_nee(PTConsts c, Queue in, const uint32_t* __restrict__ in_count,
                                                        const float4* __restrict__ hits, Queue out,
                                                        uint32_t* out_count, ShadowQueue sq, uint32_t* shadow_count) {
    const uint32_t i = blockIdx.x * kPTBlock + threadIdx.x;
    const bool valid = i < *in_count;
    const uint32_t j = valid ? i : 0u;  // invalid lanes read slot 0 and discard it
    const bool mis = c.renderer == LMGPU_PT_MIS;
    const float4 ro = in.ray_o[j], rd = in.ray_d[j], th = in.thr[j], ax = in.aux[j];
    const uint4 meta = in.meta[j];
    const V3 o = v3(ro.x, ro.y, ro.z), d = v3(rd.x, rd.y, rd.z);
    V3 T = v3(th.x, th.y, th.z);
    int n = int(meta.y & 0xffffu);
    const uint32_t flags = meta.y >> 16;
    SampleStream rng;
    rng.init(c.seed, c.base + meta.x, meta.z);
    Surf s;
    bool alive = valid && surface_at(c, o, d, ro.w, hits[j], s);
    if (alive && s.light >= 0) {
        const float LeD = dot(s.f.sn, -d) > 0.f ? kInvPi : 0.f;
        const V3 LeP = light_Le(c, s.light) * kPi;
        if (flags & kFlagSpecular) {
            film_add(c, meta.w, T * v3s(LeD) * LeP);
        } else if (mis) {
            // GeneralizedGeometryTerm(previous vertex, hit) and the light-sampling pdf
            V3 dv = s.p - o;
            const float l2 = dot(dv, dv);
            const float l = sqrtf(l2);
            dv = dv / l;
            float num = 1.f;
            if (!(flags & kFlagDegenerate)) num *= fabsf(dot(v3(ax.x, ax.y, ax.z), dv));
            num *= fabsf(dot(s.f.sn, -dv));
            const float G = fabsf(l) < kEps ? 0.f : num / l2;
            const float pos_pdf = s.light == c.env.index ? c.env.inv_area
                                                         : reinterpret_cast<const float4*>(c.lights)[2 * s.light + 1].z;
            const float pdf_dl = G == 0.f ? 0.f : (1.f / float(c.num_lights)) * pos_pdf / G;
            const float w = rd.w / (rd.w + pdf_dl);
            film_add(c, meta.w, T * v3s(LeD) * LeP * w);
        }
    }
    if (alive) {
        n++;
        if (c.max_pv != -1 && n >= c.max_pv) alive = false;
    }
    NEEVertex v;
    v.p = s.p;
    v.s = s;
    v.camera = false;
    v.wi = -d;
    v.m = alive ? mat_info(c, s.mat) : MatInfo{LMGPU_BSDF_MIRROR, v3s(0.f)};
    const ShadowRec rec = nee(c, v, T, n, meta.w, rng, mis, alive && mat_nondelta(v.m));
    const uint32_t sslot = block_append<0>(shadow_count, rec.emit);
    if (rec.emit) put_shadow(sq, sslot, rec);
    if (alive && c.rr_depth != -1 && n >= c.rr_depth) {
        const float q = fminf(0.5f, luminance(T));
        if (rng.next() > q) alive = false;
        else T = T / q;
    }
    V3 wo = v3s(0.f);
    float pdf = 0.f;
    int type = 0;
    if (alive) {
        const float u0 = rng.next(), u1 = rng.next(), uc = rng.next();
        const V3 fs = bsdf_sample(v.m, s, -d, u0, u1, uc, wo, pdf, type);
        if (is_zero(fs)) alive = false;
        else T = T * fs;
    }
    const uint32_t slot = block_append<1>(out_count, alive);
    if (alive) {
        const uint32_t nf = uint32_t(n) | ((type & (1 << 2)) ? (kFlagSpecular << 16) : 0u);
        store_path(out, slot, s.p, kEps, wo, pdf, T, s.f.sn, meta.x, nf, rng.dim, meta.w);
    }
}

// NEE shadow rays: any hit over the triangles, then the film update when unoccluded.
struct ShadowSource {
    ShadowQueue sq;
    __device__ __forceinline__ RayIn operator()(uint32_t i) const {
        const float4 o = sq.o[i], d = sq.d[i];
        return RayIn{o.x, o.y, o.z, d.x, d.y, d.z, kEps, o.w};
    }
};

struct ShadowSink {
    ShadowQueue sq;
    void* film;
    int fp64;
    __device__ __forceinline__ void operator()(uint32_t i, bool found, const HitOut&, const RayIn&) const {
        if (found) return;
        const float4 v = sq.c[i];
        film_add_raw(film, fp64, __float_as_uint(sq.d[i].w), v3(v.x, v.y, v.z));
    }
};

__global__ __launch_bounds__(kTraceBlock) void k_shadow(PTConsts c, ShadowQueue sq, const uint32_t* __restrict__ count,
                                                        uint32_t* next) {
    __shared__ uint32_t lds_stack[kStackLDS * kTraceBlock];
    const uint32_t n = *count;
    ShadowSource src{sq};
    ShadowSink sink{sq, c.film, c.film_fp64};
    TraceCounters cnt{0, 0};
    trace_persistent<true, false>(c.nodes, c.tris, n, next, lds_stack + threadIdx.x, src, sink, cnt);
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

__global__ void k_zero_bytes(uint4* p, size_t n16) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n16) p[i] = make_uint4(0u, 0u, 0u, 0u);
}

__global__ void k_accumulate_totals(const uint32_t* counters, int segments, unsigned long long* totals) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        unsigned long long a = 0, b = 0;
        for (int k = 0; k < segments; k++) {
            a += counters[k];
            b += counters[kMaxSegments + k];
        }
        totals[0] += a;
        totals[1] += b;
    }
}

// HDRBitmapFilm::Rescale after the master-film reduction (sched.mt.cpp:250-256):
// weight = Float(W*H) / Float(N), data *= weight, in the film's arithmetic type.
template <class Acc, class Out>
__global__ void k_resolve(const Acc* __restrict__ acc, Out* __restrict__ out, size_t n, Acc weight) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = Out(acc[i] * weight);
}

// ---------------------------------------------------------------------------------
// Host driver

static size_t queue_bytes(uint32_t cap) { return size_t(cap) * 16 * 5; }

static void carve(void* base, uint32_t cap, PTBuffers& b) {
    char* p = static_cast<char*>(base);
    for (int k = 0; k < 2; k++) {
        b.q[k].ray_o = reinterpret_cast<float4*>(p); p += size_t(cap) * 16;
        b.q[k].ray_d = reinterpret_cast<float4*>(p); p += size_t(cap) * 16;
        b.q[k].thr = reinterpret_cast<float4*>(p); p += size_t(cap) * 16;
        b.q[k].aux = reinterpret_cast<float4*>(p); p += size_t(cap) * 16;
        b.q[k].meta = reinterpret_cast<uint4*>(p); p += size_t(cap) * 16;
    }
    b.hits = reinterpret_cast<float4*>(p); p += size_t(cap) * 16;
    b.sq.o = reinterpret_cast<float4*>(p); p += size_t(cap) * 16;
    b.sq.d = reinterpret_cast<float4*>(p); p += size_t(cap) * 16;
    b.sq.c = reinterpret_cast<float4*>(p); p += size_t(cap) * 16;
    b.counters = reinterpret_cast<uint32_t*>(p); p += 4096;
    b.totals = reinterpret_cast<unsigned long long*>(p);
}

static size_t pt_bytes(uint32_t cap) { return 2 * queue_bytes(cap) + size_t(cap) * 16 * 4 + 4096 + 64; }

lmgpu_status run_pt(lmgpu_ctx* ctx, const lmgpu_pt_params* p, lmgpu_film* film, lmgpu_pt_stats* stats) {
    if (!ctx->has_bvh) return fail(ctx, LMGPU_ERR_NO_BVH, "lmgpu_pt_render: build the BVH first");
    if (!ctx->has_shading || !ctx->has_camera)
        return fail(ctx, LMGPU_ERR_NO_SCENE, "lmgpu_pt_render: scene has no materials or camera");
    if (p->renderer < LMGPU_PT || p->renderer > LMGPU_PT_MIS)
        return fail(ctx, LMGPU_ERR_INVALID_ARGUMENT, "lmgpu_pt_render: unknown renderer");
    if (p->renderer != LMGPU_PT && ctx->num_lights == 0)
        return fail(ctx, LMGPU_ERR_INVALID_ARGUMENT, "lmgpu_pt_render: pt.direct / pt.mis need at least one light");
    if (p->film_precision != LMGPU_FILM_FP32 && p->film_precision != LMGPU_FILM_FP64)
        return fail(ctx, LMGPU_ERR_UNSUPPORTED, "lmgpu_pt_render: film precision not supported");
    if (p->num_samples == 0) return fail(ctx, LMGPU_ERR_INVALID_ARGUMENT, "lmgpu_pt_render: num_samples is 0");
    if (p->max_path_vertices == 0 || p->max_path_vertices < -1)
        return fail(ctx, LMGPU_ERR_INVALID_ARGUMENT, "lmgpu_pt_render: max_path_vertices must be -1 or >= 1");
    if (p->max_path_vertices == -1 || p->max_path_vertices > kMaxSegments)
        return fail(ctx, LMGPU_ERR_UNSUPPORTED, "lmgpu_pt_render: max_path_vertices must be in [1, 64]");
    if (!film || !film->data || film->width != uint32_t(ctx->cam.width) || film->height != uint32_t(ctx->cam.height))
        return fail(ctx, LMGPU_ERR_INVALID_ARGUMENT, "lmgpu_pt_render: film size differs from the camera's");
    if (film->precision != LMGPU_FILM_FP32 && film->precision != LMGPU_FILM_FP64)
        return fail(ctx, LMGPU_ERR_INVALID_ARGUMENT, "lmgpu_pt_render: output film precision");
    if (3u * ctx->bvh_max_depth > uint32_t(kStackTotal))
        return fail(ctx, LMGPU_ERR_BVH_TOO_DEEP, "lmgpu_pt_render: BVH deeper than the traversal stack");

    uint32_t cap = p->wavefront_width ? p->wavefront_width : (1u << 21);
    cap = uint32_t(std::min<uint64_t>(cap, p->num_samples));
    cap = std::max<uint32_t>(cap, 64u);
    const size_t need = pt_bytes(cap);
    if (need > ctx->pt_buf_bytes) {
        dev_free(ctx->pt_buf);
        ctx->pt_buf = nullptr;
        ctx->pt_buf_bytes = 0;
        lmgpu_status s = dev_alloc(ctx, &ctx->pt_buf, need, "path queues");
        if (s != LMGPU_OK) return s;
        ctx->pt_buf_bytes = need;
    }
    ctx->pt_capacity = cap;
    const size_t npix = size_t(ctx->cam.width) * size_t(ctx->cam.height);
    const size_t acc_elem = p->film_precision == LMGPU_FILM_FP64 ? 8 : 4;
    const size_t acc_bytes = (npix * 3 * acc_elem + 15) / 16 * 16;
    if (acc_bytes > ctx->film_acc_bytes) {
        dev_free(ctx->film_acc);
        ctx->film_acc = nullptr;
        ctx->film_acc_bytes = 0;
        lmgpu_status s = dev_alloc(ctx, &ctx->film_acc, acc_bytes, "film accumulator");
        if (s != LMGPU_OK) return s;
        ctx->film_acc_bytes = acc_bytes;
    }
    const size_t out_elem = film->precision == LMGPU_FILM_FP64 ? 8 : 4;
    const size_t out_bytes = npix * 3 * out_elem;
    void* out_dev = film->data;
    if (!film->device) {
        if (out_bytes > ctx->film_out_staging_bytes) {
            dev_free(ctx->film_out_staging);
            ctx->film_out_staging = nullptr;
            ctx->film_out_staging_bytes = 0;
            lmgpu_status s = dev_alloc(ctx, &ctx->film_out_staging, out_bytes, "film staging");
            if (s != LMGPU_OK) return s;
            ctx->film_out_staging_bytes = out_bytes;
        }
        out_dev = ctx->film_out_staging;
    }

    PTBuffers b;
    carve(ctx->pt_buf, cap, b);
    PTConsts c;
    c.nodes = ctx->nodes.ptr;
    c.tris = ctx->tris.ptr;
    c.shade = ctx->shade.ptr;
    c.mats = ctx->mats.ptr;
    c.lights = ctx->lights.ptr;
    c.light_tris = ctx->light_tris.ptr;
    c.num_lights = ctx->num_lights;
    c.cam = ctx->cam;
    c.env = ctx->env;
    c.film = ctx->film_acc;
    c.film_fp64 = p->film_precision == LMGPU_FILM_FP64;
    c.seed = p->seed;
    c.max_pv = p->max_path_vertices;
    c.rr_depth = p->rr_depth;
    c.renderer = p->renderer;

    hipStream_t s = ctx->stream;
    hipError_t e;
    int occ_e = 0, occ_s = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_e, k_extend, kTraceBlock, 0) != hipSuccess || occ_e < 1) occ_e = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_s, k_shadow, kTraceBlock, 0) != hipSuccess || occ_s < 1) occ_s = 1;
    const uint32_t trace_grid = uint32_t(std::max(ctx->num_cus, 1)) * uint32_t(std::min(occ_e, occ_s));
    hipEvent_t r0, r1;
    (void)hipEventCreate(&r0);
    (void)hipEventCreate(&r1);
    (void)hipEventRecord(r0, s);
    hipLaunchKernelGGL(k_zero_bytes, dim3(uint32_t((acc_bytes / 16 + 255) / 256)), dim3(256), 0, s,
                       static_cast<uint4*>(ctx->film_acc), acc_bytes / 16);
    hipLaunchKernelGGL(k_zero_u32, dim3(5), dim3(256), 0, s, b.counters, (4096u + 64u) / 4u);
    // Segments traced per path: at most max_path_vertices - 1 (pathtrace.cpp:196-199).
    const int segments = p->max_path_vertices - 1;
    const bool is_pt = p->renderer == LMGPU_PT;
    for (uint64_t base = 0; base < p->num_samples; base += cap) {
        const uint32_t nb = uint32_t(std::min<uint64_t>(cap, p->num_samples - base));
        c.base = p->sample_offset + base;
        c.batch = nb;
        hipLaunchKernelGGL(k_zero_u32, dim3(2), dim3(256), 0, s, b.counters, 512u);
        const dim3 grid_b((nb + kPTBlock - 1) / kPTBlock), grid_c((cap + kPTBlock - 1) / kPTBlock);
        // Persistent traversal grids: the device fill, never more than the queue capacity needs.
        const dim3 grid_t(std::min<uint32_t>(grid_c.x, trace_grid));
        // counters[k]: queue entering segment k; counters[kMaxSegments + k]: shadow rays emitted before segment k.
        e = timed(ctx, kKGen, [&]() {
            hipLaunchKernelGGL(k_gen, grid_b, dim3(kPTBlock), 0, s, c, b.q[0], b.counters + 0, b.sq,
                               b.counters + kMaxSegments + 0);
        });
        if (e != hipSuccess) return hip_fail(ctx, e, "k_gen");
        if (!is_pt) {
            e = timed(ctx, kKShadow, [&]() {
                hipLaunchKernelGGL(k_shadow, grid_t, dim3(kTraceBlock), 0, s, c, b.sq, b.counters + kMaxSegments + 0,
                                   b.counters + 3 * kMaxSegments + 0);
            });
            if (e != hipSuccess) return hip_fail(ctx, e, "k_shadow");
        }
        for (int k = 0; k < segments; k++) {
            const Queue& qi = b.q[k & 1];
            const Queue& qo = b.q[(k + 1) & 1];
            e = timed(ctx, kKExtend, [&]() {
                hipLaunchKernelGGL(k_extend, grid_t, dim3(kTraceBlock), 0, s, c.nodes, c.tris, qi, b.counters + k,
                                   b.counters + 2 * kMaxSegments + k, b.hits);
            });
            if (e != hipSuccess) return hip_fail(ctx, e, "k_extend");
            e = timed(ctx, kKShade, [&]() {
                if (is_pt)
                    hipLaunchKernelGGL(k_shade_pt, grid_c, dim3(kPTBlock), 0, s, c, qi, b.counters + k, b.hits, qo,
                                       b.counters + k + 1);
                else
                    hipLaunchKernelGGL(k_shade_nee, grid_c, dim3(kPTBlock), 0, s, c, qi, b.counters + k, b.hits, qo,
                                       b.counters + k + 1, b.sq, b.counters + kMaxSegments + k + 1);
            });
            if (e != hipSuccess) return hip_fail(ctx, e, "k_shade");
            if (!is_pt) {
                e = timed(ctx, kKShadow, [&]() {
                    hipLaunchKernelGGL(k_shadow, grid_t, dim3(kTraceBlock), 0, s, c, b.sq,
                                       b.counters + kMaxSegments + k + 1, b.counters + 3 * kMaxSegments + k + 1);
                });
                if (e != hipSuccess) return hip_fail(ctx, e, "k_shadow");
            }
        }
        // Paths still alive after the last segment end there; pt adds their radiance.
        if (is_pt && segments >= 0) {
            // k_shade_pt never enqueues a path whose vertex count reached max_path_vertices,
            // so the final queue is empty; nothing to flush.
        }
        hipLaunchKernelGGL(k_accumulate_totals, dim3(1), dim3(64), 0, s, b.counters, segments + 1, b.totals);
    }
    // Resolve: sums * W*H/N into the output film.
    const size_t nval = npix * 3;
    const dim3 rg(uint32_t((nval + 255) / 256));
    if (p->film_precision == LMGPU_FILM_FP64) {
        const double w = double(npix) / double(p->num_samples);
        if (film->precision == LMGPU_FILM_FP64)
            hipLaunchKernelGGL((k_resolve<double, double>), rg, dim3(256), 0, s, static_cast<const double*>(ctx->film_acc),
                               static_cast<double*>(out_dev), nval, w);
        else
            hipLaunchKernelGGL((k_resolve<double, float>), rg, dim3(256), 0, s, static_cast<const double*>(ctx->film_acc),
                               static_cast<float*>(out_dev), nval, w);
    } else {
        const float w = float(npix) / float(p->num_samples);
        if (film->precision == LMGPU_FILM_FP64)
            hipLaunchKernelGGL((k_resolve<float, double>), rg, dim3(256), 0, s, static_cast<const float*>(ctx->film_acc),
                               static_cast<double*>(out_dev), nval, w);
        else
            hipLaunchKernelGGL((k_resolve<float, float>), rg, dim3(256), 0, s, static_cast<const float*>(ctx->film_acc),
                               static_cast<float*>(out_dev), nval, w);
    }
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(ctx, e, "k_resolve");
    (void)hipEventRecord(r1, s);
    unsigned long long totals[2] = {0, 0};
    e = hipMemcpyAsync(totals, b.totals, sizeof(totals), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return hip_fail(ctx, e, "totals readback");
    if (!film->device) {
        e = hipMemcpyAsync(film->data, out_dev, out_bytes, hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) return hip_fail(ctx, e, "film readback");
    }
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(ctx, e, "pt synchronize");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, r0, r1);
    (void)hipEventDestroy(r0);
    (void)hipEventDestroy(r1);
    if (stats) {
        stats->closest_rays = totals[0];
        stats->shadow_rays = totals[1];
        stats->samples = p->num_samples;
        stats->seconds = double(ms) / 1000.0;
    }
    return LMGPU_OK;
}

}  // namespace lmgpu
