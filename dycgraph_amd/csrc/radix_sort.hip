// radix_sort.hip -- stable LSD radix sort of 64-bit keys with optional 32-bit
// values, hand-written for gfx950 (wave64).
//
// One pass per 8-bit digit, three launches each:
//   k_rs_hist     a 4,096-key tile per 256-thread workgroup counts its digits in
//                 LDS and stores the 256 counts digit-major (hist[digit][tile]);
//   k_rs_scan     one workgroup per digit turns its row into exclusive offsets
//                 over the tiles and stores the digit's total;
//   k_rs_scatter  every tile ranks its keys in 16 rounds of 256: inside a wave
//                 the lanes holding the same digit are found with 8 ballots
//                 (multisplit), their rank is the popcount of the lower peers;
//                 waves, rounds and tiles follow in index order, so the sort is
//                 stable.  A key's place = digits below it (scan of the totals)
//                 + its tile's offset + the keys of its digit before it in the tile.
#include "radix_sort.hpp"

namespace dcg {

constexpr uint32_t kRsBlock = 256, kRsRounds = 16, kRsTile = kRsBlock * kRsRounds, kRsBins = 256;

__global__ __launch_bounds__(256) void k_rs_hist(const uint64_t *keys, uint64_t n, uint32_t shift, uint32_t mask,
                                                 uint32_t *hist, uint32_t ntiles) {
    __shared__ uint32_t h[kRsBins];
    const uint32_t t = threadIdx.x;
    h[t] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kRsTile;
    for (uint32_t r = 0; r < kRsRounds; ++r) {
        const uint64_t i = base + r * kRsBlock + t;
        if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & mask], 1u);
    }
    __syncthreads();
    hist[(uint64_t)t * ntiles + blockIdx.x] = h[t];
}

// Inclusive sum over the 1,024 threads of a workgroup (16 waves); `wsum` holds
// 16 words of LDS.  Every thread gets its inclusive prefix and the block total.
__device__ __forceinline__ uint32_t block_incl_scan(uint32_t x, uint32_t *wsum, uint32_t &total) {
    const uint32_t w = threadIdx.x / kWave, nw = blockDim.x / kWave;
    const uint32_t incl = wave_incl_scan(x);
    if (lane_id() == kWave - 1) wsum[w] = incl;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (uint32_t q = 0; q < nw; ++q) {
        const uint32_t v = wsum[q];
        before += q < w ? v : 0u;
        all += v;
    }
    __syncthreads();   // wsum is reused by the next call
    total = all;
    return before + incl;
}

__global__ __launch_bounds__(1024) void k_rs_scan(uint32_t *hist, uint32_t ntiles, uint32_t *totals) {
    __shared__ uint32_t wsum[16];
    uint32_t *row = hist + (uint64_t)blockIdx.x * ntiles;
    uint32_t running = 0;
    for (uint32_t c0 = 0; c0 < ntiles; c0 += blockDim.x) {
        const uint32_t i = c0 + threadIdx.x;
        const uint32_t x = i < ntiles ? row[i] : 0u;
        uint32_t total;
        const uint32_t incl = block_incl_scan(x, wsum, total);
        if (i < ntiles) row[i] = running + incl - x;
        running += total;
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = running;
}

template <bool VALS>
__global__ __launch_bounds__(256) void k_rs_scatter(const uint64_t *keys_in, uint64_t *keys_out, const uint32_t *vals_in,
                                                    uint32_t *vals_out, uint64_t n, uint32_t shift, uint32_t mask,
                                                    const uint32_t *hist, const uint32_t *totals, uint32_t ntiles) {
    __shared__ uint32_t base[kRsBins];
    __shared__ uint32_t wcnt[kRsBlock / kWave][kRsBins];
    __shared__ uint32_t woff[kRsBlock / kWave][kRsBins];
    __shared__ uint32_t wsum[kRsBlock / kWave];
    const uint32_t t = threadIdx.x, w = t / kWave;
    // digit bases: the totals of the digits below + this tile's offset
    {
        const uint32_t x = totals[t];
        const uint32_t incl = wave_incl_scan(x);
        if (lane_id() == kWave - 1) wsum[w] = incl;
        __syncthreads();
        uint32_t before = 0;
        for (uint32_t q = 0; q < w; ++q) before += wsum[q];
        base[t] = before + incl - x + hist[(uint64_t)t * ntiles + blockIdx.x];
    }
    const uint64_t tile0 = (uint64_t)blockIdx.x * kRsTile;
    for (uint32_t r = 0; r < kRsRounds; ++r) {
        for (uint32_t j = t; j < (kRsBlock / kWave) * kRsBins; j += kRsBlock) (&wcnt[0][0])[j] = 0;
        __syncthreads();   // also: every thread finished the previous round's writes
        const uint64_t i = tile0 + r * kRsBlock + t;
        const bool valid = i < n;
        const uint64_t key = valid ? keys_in[i] : 0ull;
        const uint32_t d = (uint32_t)(key >> shift) & mask;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (uint32_t b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(valid && bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t rank = lanes_below(peers);
        if (valid && rank == 0) wcnt[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        {
            uint32_t run = base[t];
#pragma unroll
            for (uint32_t q = 0; q < kRsBlock / kWave; ++q) {
                woff[q][t] = run;
                run += wcnt[q][t];
            }
            base[t] = run;
        }
        __syncthreads();
        if (valid) {
            const uint32_t pos = woff[w][d] + rank;
            keys_out[pos] = key;
            if (VALS) vals_out[pos] = vals_in[i];
        }
    }
}

static uint64_t tiles_of(uint64_t n) { return (n + kRsTile - 1) / kRsTile; }

size_t radix_sort_temp_bytes(uint64_t n, bool with_values) {
    const uint64_t nt = tiles_of(n) ? tiles_of(n) : 1;
    size_t b = n * sizeof(uint64_t);
    if (with_values) b += n * sizeof(uint32_t);
    b = (b + 255) & ~(size_t)255;
    b += ((size_t)kRsBins * nt + kRsBins) * sizeof(uint32_t);
    return b;
}

void radix_sort(hipStream_t s, const uint64_t *keys_in, uint64_t *keys_out, const uint32_t *vals_in, uint32_t *vals_out,
                uint64_t n, int begin_bit, int end_bit, void *temp, size_t temp_bytes) {
    const bool vals = vals_in && vals_out;
    if (n == 0) return;
    if (n > 0xFFFFFFFFull) throw Error(DCG_ERR_UNSUPPORTED, "radix sort of more than 2^32 - 1 keys");
    if (keys_in == keys_out || (vals && vals_in == vals_out)) throw Error(DCG_ERR_ARG, "radix sort needs distinct buffers");
    if (temp_bytes < radix_sort_temp_bytes(n, vals)) throw Error(DCG_ERR_ARG, "radix sort scratch too small");
    const int passes = end_bit > begin_bit ? (end_bit - begin_bit + 7) / 8 : 0;
    if (passes == 0) {
        DCG_HIP(hipMemcpyAsync(keys_out, keys_in, n * 8, hipMemcpyDeviceToDevice, s));
        if (vals) DCG_HIP(hipMemcpyAsync(vals_out, vals_in, n * 4, hipMemcpyDeviceToDevice, s));
        return;
    }
    const uint32_t nt = (uint32_t)tiles_of(n);
    uint8_t *p = static_cast<uint8_t *>(temp);
    uint64_t *alt_k = reinterpret_cast<uint64_t *>(p);
    p += n * sizeof(uint64_t);
    uint32_t *alt_v = vals ? reinterpret_cast<uint32_t *>(p) : nullptr;
    if (vals) p += n * sizeof(uint32_t);
    p = static_cast<uint8_t *>(temp) + ((size_t)(p - static_cast<uint8_t *>(temp)) + 255 & ~(size_t)255);
    uint32_t *hist = reinterpret_cast<uint32_t *>(p);
    uint32_t *totals = hist + (size_t)kRsBins * nt;
    const uint64_t *src_k = keys_in;
    const uint32_t *src_v = vals_in;
    for (int pass = 0; pass < passes; ++pass) {
        const uint32_t shift = (uint32_t)(begin_bit + 8 * pass);
        const int width = end_bit - (int)shift < 8 ? end_bit - (int)shift : 8;
        const uint32_t mask = (1u << width) - 1u;
        // the last pass writes keys_out
        const bool to_out = ((passes - 1 - pass) % 2) == 0;
        uint64_t *dst_k = to_out ? keys_out : alt_k;
        uint32_t *dst_v = vals ? (to_out ? vals_out : alt_v) : nullptr;
        hipLaunchKernelGGL(k_rs_hist, dim3(nt), dim3(kRsBlock), 0, s, src_k, n, shift, mask, hist, nt);
        hipLaunchKernelGGL(k_rs_scan, dim3(kRsBins), dim3(1024), 0, s, hist, nt, totals);
        if (vals)
            hipLaunchKernelGGL(k_rs_scatter<true>, dim3(nt), dim3(kRsBlock), 0, s, src_k, dst_k, src_v, dst_v, n, shift,
                               mask, hist, totals, nt);
        else
            hipLaunchKernelGGL(k_rs_scatter<false>, dim3(nt), dim3(kRsBlock), 0, s, src_k, dst_k,
                               (const uint32_t *)nullptr, (uint32_t *)nullptr, n, shift, mask, hist, totals, nt);
        src_k = dst_k;
        src_v = dst_v;
    }
    DCG_HIP(hipGetLastError());
}

}  // namespace dcg

// Sort caller keys (and values) on the GPU: host arrays, sorted in place.  No
// reference counterpart (the reference sorts with CUB); exposed so the sort
// the engine uses can be checked against a CPU sort.
extern "C" int dcg_radix_sort(int32_t device, uint64_t *keys, uint32_t *vals, uint64_t n, int begin_bit, int end_bit) {
    using namespace dcg;
    if ((!keys && n) || begin_bit < 0 || end_bit > 64 || begin_bit > end_bit) return DCG_ERR_ARG;
    if (n == 0) return DCG_OK;
    try {
        DCG_HIP(hipSetDevice(device));
        hipStream_t s;
        DCG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        DevBuf<uint64_t> ki, ko;
        DevBuf<uint32_t> vi, vo;
        DevBuf<uint8_t> tmp;
        ki.alloc(n);
        ko.alloc(n);
        if (vals) {
            vi.alloc(n);
            vo.alloc(n);
        }
        tmp.alloc(radix_sort_temp_bytes(n, vals != nullptr));
        DCG_HIP(hipMemcpyAsync(ki.p, keys, n * 8, hipMemcpyHostToDevice, s));
        if (vals) DCG_HIP(hipMemcpyAsync(vi.p, vals, n * 4, hipMemcpyHostToDevice, s));
        radix_sort(s, ki.p, ko.p, vals ? vi.p : nullptr, vals ? vo.p : nullptr, n, begin_bit, end_bit, tmp.p, tmp.n);
        DCG_HIP(hipMemcpyAsync(keys, ko.p, n * 8, hipMemcpyDeviceToHost, s));
        if (vals) DCG_HIP(hipMemcpyAsync(vals, vo.p, n * 4, hipMemcpyDeviceToHost, s));
        DCG_HIP(hipStreamSynchronize(s));
        DCG_HIP(hipStreamDestroy(s));
    } catch (const Error &e) {
        return e.status;
    }
    return DCG_OK;
}
