// calibrate.hip -- link microbenchmarks: HBM copy rate, zero-copy read rate
// of pinned host memory (the host link the uncached chunks are read over) and
// random 64-byte line read rates of both.
// Also the known-size copy kernel used to scale rocprof FETCH_SIZE/WRITE_SIZE.
#include "engine.hpp"

namespace dcg {

__global__ __launch_bounds__(256) void k_stream_copy(const uint4 *src, uint4 *dst, uint64_t n16) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Sum of 16-byte words (the result is written so the reads are not elided).
__global__ __launch_bounds__(256) void k_stream_read(const uint4 *src, uint64_t n16, unsigned long long *sink) {
    uint32_t acc = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint4 v = src[i];
        acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    if (acc == 0x12345678u) atomicAdd(sink, 1ull);
}

// Each quarter-wave reads `reads` random 64-byte lines of buf[0, lines * 16).
__global__ __launch_bounds__(256) void k_random_lines(const uint32_t *buf, uint64_t lines, int reads,
                                                      unsigned long long *sink) {
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    uint32_t acc = 0;
    for (int r = 0; r < reads; ++r) {
        const uint64_t line = mix64((t >> 4) * 1000003ull + (uint64_t)r) % lines;
        acc ^= buf[line * 16 + (t & 15)];
    }
    if (acc == 0x12345678u) atomicAdd(sink, 1ull);
}

}  // namespace dcg

extern "C" int dcg_calibrate_lines(int32_t device, uint64_t bytes, double *host_lines_per_s, double *hbm_lines_per_s) {
    using namespace dcg;
    if (!host_lines_per_s || !hbm_lines_per_s || bytes < (1u << 20)) return DCG_ERR_ARG;
    try {
        DCG_HIP(hipSetDevice(device));
        hipStream_t s;
        DCG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        hipDeviceProp_t prop;
        DCG_HIP(hipGetDeviceProperties(&prop, device));
        const uint64_t words = bytes / 64 * 16, lines = words / 16;
        const uint32_t grid = (uint32_t)prop.multiProcessorCount * 16;
        const int reads = 16;
        DevBuf<uint32_t> dev;
        DevBuf<unsigned long long> sink;
        dev.alloc(words);
        sink.alloc(1);
        fill<uint32_t>(s, dev.p, words, 1u);
        fill<unsigned long long>(s, sink.p, 1, 0ull);
        uint32_t *h = nullptr;
        DCG_HIP(hipHostMalloc((void **)&h, words * 4, hipHostMallocDefault));
        std::memset(h, 1, words * 4);
        hipEvent_t e0, e1;
        DCG_HIP(hipEventCreate(&e0));
        DCG_HIP(hipEventCreate(&e1));
        const double total = (double)grid * 256 / 16 * reads;
        for (int mem = 0; mem < 2; ++mem) {
            float best = 1e30f;
            for (int it = 0; it < 4; ++it) {
                DCG_HIP(hipEventRecord(e0, s));
                hipLaunchKernelGGL(k_random_lines, dim3(grid), dim3(256), 0, s, mem ? dev.p : h, lines, reads, sink.p);
                DCG_HIP(hipEventRecord(e1, s));
                DCG_HIP(hipEventSynchronize(e1));
                float ms = 0;
                DCG_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (it > 0 && ms < best) best = ms;
            }
            (mem ? *hbm_lines_per_s : *host_lines_per_s) = total / (best * 1e-3);
        }
        DCG_HIP(hipHostFree(h));
        DCG_HIP(hipEventDestroy(e0));
        DCG_HIP(hipEventDestroy(e1));
        DCG_HIP(hipStreamDestroy(s));
    } catch (const Error &e) {
        return e.status;
    }
    return DCG_OK;
}

extern "C" int dcg_calibrate(int32_t device, uint64_t bytes, double *hbm_copy_gbs, double *host_read_gbs) {
    using namespace dcg;
    if (!hbm_copy_gbs || !host_read_gbs || bytes < (1u << 20)) return DCG_ERR_ARG;
    try {
        DCG_HIP(hipSetDevice(device));
        hipStream_t s;
        DCG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        hipDeviceProp_t prop;
        DCG_HIP(hipGetDeviceProperties(&prop, device));
        const uint32_t grid = (uint32_t)prop.multiProcessorCount * 8;
        const uint64_t n16 = bytes / 16;
        DevBuf<uint4> a, b;
        DevBuf<unsigned long long> sink;
        a.alloc(n16);
        b.alloc(n16);
        sink.alloc(1);
        fill<unsigned long long>(s, sink.p, 1, 0ull);
        fill<uint32_t>(s, reinterpret_cast<uint32_t *>(a.p), n16 * 4, 1u);
        hipEvent_t e0, e1;
        DCG_HIP(hipEventCreate(&e0));
        DCG_HIP(hipEventCreate(&e1));
        float best = 1e30f;
        for (int it = 0; it < 6; ++it) {
            DCG_HIP(hipEventRecord(e0, s));
            hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, s, a.p, b.p, n16);
            DCG_HIP(hipEventRecord(e1, s));
            DCG_HIP(hipEventSynchronize(e1));
            float ms = 0;
            DCG_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (it > 0 && ms < best) best = ms;
        }
        *hbm_copy_gbs = 2.0 * (double)(n16 * 16) / (best * 1e-3) / 1e9;
        uint4 *h = nullptr;
        DCG_HIP(hipHostMalloc((void **)&h, n16 * 16, hipHostMallocDefault));
        std::memset(h, 1, n16 * 16);
        best = 1e30f;
        for (int it = 0; it < 4; ++it) {
            DCG_HIP(hipEventRecord(e0, s));
            hipLaunchKernelGGL(k_stream_read, dim3(grid), dim3(256), 0, s, h, n16, sink.p);
            DCG_HIP(hipEventRecord(e1, s));
            DCG_HIP(hipEventSynchronize(e1));
            float ms = 0;
            DCG_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (it > 0 && ms < best) best = ms;
        }
        *host_read_gbs = (double)(n16 * 16) / (best * 1e-3) / 1e9;
        DCG_HIP(hipHostFree(h));
        DCG_HIP(hipEventDestroy(e0));
        DCG_HIP(hipEventDestroy(e1));
        DCG_HIP(hipStreamDestroy(s));
    } catch (const Error &e) {
        return e.status;
    }
    return DCG_OK;
}
