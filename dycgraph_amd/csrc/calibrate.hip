// calibrate.hip -- link microbenchmarks: HBM copy rate and zero-copy read rate
// of pinned host memory (the host link the uncached chunks are read over).
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

}  // namespace dcg

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
