// gen_gpu.hip -- GPU RMAT generator: the same draws and the same edge set as
// dcg_gen_rmat (gen.cpp), for scales where a host hash set is too slow.
//
//   1. draw M candidates (key = u<<scale | v; self-loops get a key above every
//      valid one), payload = draw index;
//   2. stable radix sort by key: the head of each equal-key run carries the
//      pair's first draw index;
//   3. mark first draws, prefix-sum the marks over draw order, and find T, the
//      draw at which the num_edges-th distinct pair appeared (M grows by 25%
//      and the pass repeats if M draws hold too few distinct pairs);
//   4. keep the run heads whose first draw is <= T: already sorted by (u, v).

#include "engine.hpp"
#include "gen_common.hpp"
#include "primitives.hpp"

namespace dcg {

__global__ void k_rmat_keys(RmatParams p, uint64_t count, uint64_t *keys, uint32_t *idx) {
    const uint64_t invalid = 1ull << (2 * p.scale);
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = rmat_draw(p, i);
        const uint32_t u = (uint32_t)(k >> 32), v = (uint32_t)k;
        keys[i] = u == v ? invalid : (((uint64_t)u << p.scale) | v);
        idx[i] = (uint32_t)i;
    }
}

__global__ void k_mark_first(const uint64_t *keys, const uint32_t *idx, uint64_t count, uint64_t invalid,
                             uint32_t *mark) {
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < count; j += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[j];
        if (k != invalid && (j == 0 || keys[j - 1] != k)) mark[idx[j]] = 1u;
    }
}

__global__ void k_find_cut(const uint32_t *mark, const uint32_t *scan, uint64_t count, uint32_t want,
                           unsigned long long *cut) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
        if (mark[i] && scan[i] == want - 1) *cut = i;
}

__global__ void k_emit(const uint64_t *keys, const uint32_t *idx, uint64_t count, uint64_t invalid, uint64_t cut,
                       uint32_t scale, const uint32_t *pos, uint32_t *src, uint32_t *dst) {
    const uint64_t mask = (1ull << scale) - 1;
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < count; j += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[j];
        if (k == invalid || (j > 0 && keys[j - 1] == k) || idx[j] > cut) continue;
        const uint32_t o = pos[j];
        src[o] = (uint32_t)(k >> scale);
        dst[o] = (uint32_t)(k & mask);
    }
}

__global__ void k_keep_flags(const uint64_t *keys, const uint32_t *idx, uint64_t count, uint64_t invalid, uint64_t cut,
                             uint32_t *flag) {
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < count; j += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[j];
        flag[j] = (k != invalid && (j == 0 || keys[j - 1] != k) && idx[j] <= cut) ? 1u : 0u;
    }
}

}  // namespace dcg

extern "C" int dcg_gen_rmat_gpu(int32_t device, uint32_t scale, uint64_t num_edges, double a, double b, double c,
                                uint64_t seed, uint32_t *src, uint32_t *dst) {
    using namespace dcg;
    if (scale == 0 || scale > 31 || !src || !dst || a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return DCG_ERR_ARG;
    if (num_edges == 0) return DCG_OK;
    if (num_edges >= (1ull << 31)) return DCG_ERR_UNSUPPORTED;
    const double nn = std::ldexp(1.0, (int)scale);
    if ((double)num_edges > nn * (nn - 1) * 0.5) return DCG_ERR_ARG;
    const RmatParams p = rmat_params(scale, a, b, c, seed);
    const uint64_t invalid = 1ull << (2 * scale);
    const int end_bit = (int)(2 * scale + 1);
    try {
        DCG_HIP(hipSetDevice(device));
        hipStream_t s;
        DCG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        uint64_t M = num_edges + num_edges / 8 + 4096;
        for (int attempt = 0; attempt < 40; ++attempt) {
            if (M >= (1ull << 31)) M = (1ull << 31) - 1;
            DevBuf<uint64_t> k0, k1;
            DevBuf<uint32_t> i0, i1, mark, scan;
            DevBuf<uint8_t> tmp;
            k0.alloc(M);
            k1.alloc(M);
            i0.alloc(M);
            i1.alloc(M);
            hipLaunchKernelGGL(k_rmat_keys, dim3(grid_for(M)), dim3(256), 0, s, p, M, k0.p, i0.p);
            size_t tb = radix_sort_temp_bytes<uint64_t>(M, true);
            tmp.alloc(tb);
            radix_sort<uint64_t>(s, k0.p, k1.p, i0.p, i1.p, M, 0, end_bit, tmp.p, tb);
            k0.release();
            i0.release();
            mark.alloc(M);
            fill<uint32_t>(s, mark.p, M, 0u);
            hipLaunchKernelGGL(k_mark_first, dim3(grid_for(M)), dim3(256), 0, s, k1.p, i1.p, M, invalid, mark.p);
            scan.alloc(M);
            tb = scan_temp_bytes<uint32_t>(M);
            tmp.alloc(tb);
            exclusive_sum<uint32_t>(s, mark.p, scan.p, M, tmp.p, tb);
            uint32_t last_scan = 0, last_mark = 0;
            DCG_HIP(hipMemcpyAsync(&last_scan, scan.p + M - 1, 4, hipMemcpyDeviceToHost, s));
            DCG_HIP(hipMemcpyAsync(&last_mark, mark.p + M - 1, 4, hipMemcpyDeviceToHost, s));
            DCG_HIP(hipStreamSynchronize(s));
            const uint64_t distinct = (uint64_t)last_scan + last_mark;
            if (distinct < num_edges) {
                M = M + M / 4 + 4096;
                continue;
            }
            DevBuf<unsigned long long> cut;
            cut.alloc(1);
            fill<unsigned long long>(s, cut.p, 1, ~0ull);
            hipLaunchKernelGGL(k_find_cut, dim3(grid_for(M)), dim3(256), 0, s, mark.p, scan.p, M, (uint32_t)num_edges,
                               cut.p);
            unsigned long long T = 0;
            DCG_HIP(hipMemcpyAsync(&T, cut.p, 8, hipMemcpyDeviceToHost, s));
            DCG_HIP(hipStreamSynchronize(s));
            // output positions of the kept heads (sorted order)
            hipLaunchKernelGGL(k_keep_flags, dim3(grid_for(M)), dim3(256), 0, s, k1.p, i1.p, M, invalid, (uint64_t)T,
                               mark.p);
            tb = scan_temp_bytes<uint32_t>(M);
            tmp.alloc(tb);
            exclusive_sum<uint32_t>(s, mark.p, scan.p, M, tmp.p, tb);
            DevBuf<uint32_t> ds, dd;
            ds.alloc(num_edges);
            dd.alloc(num_edges);
            hipLaunchKernelGGL(k_emit, dim3(grid_for(M)), dim3(256), 0, s, k1.p, i1.p, M, invalid, (uint64_t)T, scale,
                               scan.p, ds.p, dd.p);
            DCG_HIP(hipGetLastError());
            DCG_HIP(hipMemcpyAsync(src, ds.p, num_edges * 4, hipMemcpyDeviceToHost, s));
            DCG_HIP(hipMemcpyAsync(dst, dd.p, num_edges * 4, hipMemcpyDeviceToHost, s));
            DCG_HIP(hipStreamSynchronize(s));
            DCG_HIP(hipStreamDestroy(s));
            return DCG_OK;
        }
        DCG_HIP(hipStreamDestroy(s));
        return DCG_ERR_CAPACITY;
    } catch (const Error &e) {
        return e.status;
    }
}
