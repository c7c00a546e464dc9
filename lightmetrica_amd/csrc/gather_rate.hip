// gather_rate: the memory system's rate for the traversal's access pattern -- every active
// lane of a wave gathers one random, aligned record (a 128-byte node line, or a 64-byte
// record) as 16-byte loads, as k_trace_closest fetches its nodes.  Prints one JSON line per
// case: buffer size, record size, active lanes per wave, records per second.
//
//   gather_rate [iterations]
//
// A measurement tool (built by lightmetrica_amd/build.py), not part of the library.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

namespace {

__device__ __forceinline__ uint32_t mix(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// Lanes [0, active) of every wave load record mix(...) & mask, kRecordF4 float4 each.
template <int kRecordF4>
__global__ __launch_bounds__(256) void k_gather(const float4* __restrict__ buf, uint32_t mask, int active, int iters,
                                                float* __restrict__ sink) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = blockIdx.x * 4u + (threadIdx.x >> 6);
    float acc = 0.f;
    for (int it = 0; it < iters; it++) {
        if (int(lane) < active) {
            const uint32_t rec = mix(wave * 7919u + uint32_t(it) * 64u + lane) & mask;
            const float4* p = buf + size_t(rec) * kRecordF4;
#pragma unroll
            for (int j = 0; j < kRecordF4; j++) {
                const float4 v = p[j];
                acc += v.x + v.w;
            }
        }
    }
    if (acc == 1.0f) sink[0] = acc;  // the buffer is zero, so this never stores
}

bool check(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    std::fprintf(stderr, "gather_rate: %s: %s\n", what, hipGetErrorString(e));
    return false;
}

}  // namespace

int main(int argc, char** argv) {
    const int iters = argc > 1 ? std::max(1, std::atoi(argv[1])) : 2000;
    const size_t big = size_t(128) << 20, small = size_t(2) << 20;
    float4* buf = nullptr;
    float* sink = nullptr;
    if (!check(hipMalloc(&buf, big), "hipMalloc") || !check(hipMalloc(&sink, sizeof(float)), "hipMalloc") ||
        !check(hipMemset(buf, 0, big), "hipMemset"))
        return 1;
    int cus = 0;
    if (!check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0), "attribute")) return 1;
    const int blocks = cus * 6;  // the traversal kernel's residency: 6 blocks of 4 waves per CU
    hipEvent_t e0, e1;
    if (!check(hipEventCreate(&e0), "event") || !check(hipEventCreate(&e1), "event")) return 1;
    for (const size_t bytes : {big, small}) {
        for (const int record : {128, 64}) {
            // record count is a power of two, so `mask` keeps every index inside the buffer
            const uint32_t mask = uint32_t(bytes / size_t(record)) - 1u;
            for (const int active : {8, 16, 32, 64}) {
                auto launch = [&](int n) {
                    if (record == 128)
                        hipLaunchKernelGGL(k_gather<8>, dim3(blocks), dim3(256), 0, 0, buf, mask, active, n, sink);
                    else
                        hipLaunchKernelGGL(k_gather<4>, dim3(blocks), dim3(256), 0, 0, buf, mask, active, n, sink);
                };
                launch(20);  // warm-up
                if (!check(hipEventRecord(e0), "record")) return 1;
                launch(iters);
                if (!check(hipEventRecord(e1), "record") || !check(hipEventSynchronize(e1), "sync")) return 1;
                float ms = 0.f;
                if (!check(hipEventElapsedTime(&ms, e0, e1), "elapsed")) return 1;
                const double recs = double(blocks) * 4.0 * iters * active;
                std::printf("{\"buffer_MiB\": %zu, \"record_bytes\": %d, \"active_lanes\": %d, \"ms\": %.3f, "
                            "\"G_records_per_s\": %.2f, \"GB_per_s\": %.1f}\n",
                            bytes >> 20, record, active, ms, recs / ms / 1e6, recs * record / ms / 1e6);
            }
        }
    }
    if (!check(hipGetLastError(), "launch")) return 1;
    return 0;
}
