/*
 * gq_gather_bench.hip -- random-access roofline microbenchmark (measurement
 * only; bench.py loads it from lib/libgq_bench.so, the planner never does).
 *
 * Reads `n_probes` records at uniformly random, 64-byte-aligned slots of a
 * buffer of `footprint` bytes.  A record is `record_bytes` long (16: one
 * dense-memo entry; 40: an 8-byte key plus a 32-byte JoinRelation, the memo
 * record the reference probes).  The rate is record bytes delivered per
 * second, the figure memo probes are compared against.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return x;
}

template <int REC>
__global__ void __launch_bounds__(256) k_gather(const uint4 *buf, uint64_t slots, uint32_t per_thread,
                                                uint64_t seed, uint32_t *sink) {
    uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t acc = 0;
    for (uint32_t i = 0; i < per_thread; i++) {
        uint64_t slot = mix64(seed ^ (tid * per_thread + i)) % slots;
        const uint4 *p = buf + slot * 4;          /* 64-byte slot = 4 x uint4 */
        uint4 a = p[0];
        acc ^= a.x ^ a.y ^ a.z ^ a.w;
        if (REC > 16) {
            uint4 b = p[1];
            uint2 c = *(const uint2 *)(p + 2);
            acc ^= b.x ^ b.y ^ b.z ^ b.w ^ c.x ^ c.y;
        }
    }
    if (acc == 0x9e3779b9u) sink[tid & 1023] = acc;   /* keeps the loads live */
}

extern "C" int gq_gather_bench(uint64_t footprint, uint32_t record_bytes, uint64_t n_probes, double *gbps,
                               double *ms_out) {
    if (footprint < 64 || (record_bytes != 16 && record_bytes != 40)) return 1;
    uint64_t slots = footprint / 64;
    uint4 *buf = nullptr;
    uint32_t *sink = nullptr;
    if (hipMalloc(&buf, slots * 64) != hipSuccess) return 3;
    if (hipMalloc(&sink, 4096) != hipSuccess) { (void)hipFree(buf); return 3; }
    (void)hipMemset(buf, 0x5a, slots * 64);
    const uint32_t per = 16;
    uint64_t threads = (n_probes + per - 1) / per;
    uint32_t blocks = (uint32_t)((threads + 255) / 256);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    /* warm-up pass, then the timed one */
    for (int rep = 0; rep < 2; rep++) {
        (void)hipEventRecord(e0, 0);
        if (record_bytes == 16) hipLaunchKernelGGL(k_gather<16>, dim3(blocks), dim3(256), 0, 0, buf, slots, per, 0x1234ull + rep, sink);
        else hipLaunchKernelGGL(k_gather<40>, dim3(blocks), dim3(256), 0, 0, buf, slots, per, 0x1234ull + rep, sink);
        (void)hipEventRecord(e1, 0);
        (void)hipEventSynchronize(e1);
    }
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(buf);
    (void)hipFree(sink);
    double bytes = (double)blocks * 256.0 * per * record_bytes;
    *gbps = bytes / (ms * 1e-3) / 1e9;
    *ms_out = ms;
    return 0;
}

/* Plain streaming copy of `bytes` (a multiple of 16): the calibration point
 * for hardware byte counters -- its algorithmic traffic is exactly `bytes`
 * read and `bytes` written per launch. */
__global__ void __launch_bounds__(256) k_copy(const uint4 *src, uint4 *dst, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

extern "C" int gq_copy_bench(uint64_t bytes, int reps, double *gbps, double *ms_out) {
    if (bytes < 16 || (bytes & 15) || reps < 1) return 1;
    const uint64_t n = bytes / 16;
    uint4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess) return 3;
    if (hipMalloc(&b, bytes) != hipSuccess) { (void)hipFree(a); return 3; }
    (void)hipMemset(a, 0x3c, bytes);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const uint32_t blocks = 4096;
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, 0, a, b, n);   /* warm-up */
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(e1, 0);
    (void)hipEventSynchronize(e1);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    *ms_out = ms / reps;
    *gbps = 2.0 * (double)bytes / (*ms_out * 1e-3) / 1e9;
    return 0;
}
