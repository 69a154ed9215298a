// primitives.hip -- device-wide radix sort, exclusive scan and flagged
// selection for gfx950 (wave64), in caller scratch.
//
// All three work on tiles of 4,096 elements: a 256-thread workgroup walks its
// tile in 16 rounds of 256 consecutive elements, so every round's loads and
// stores are coalesced and the tile count (n / 4,096) stays small enough for a
// single 1,024-thread workgroup to scan the per-tile figures.
//
// Radix sort: one pass per 8-bit digit (the last one narrower when the bit
// range is not a multiple of 8), three launches each:
//   k_rs_hist     a tile counts its digits in LDS and stores the 256 counts
//                 digit-major (hist[digit][tile]);
//   k_scan_rows   one workgroup per digit turns its row into exclusive offsets
//                 over the tiles and stores the digit's total;
//   k_rs_scatter  every wave ranks its 1,024 keys (a contiguous quarter of the
//                 tile) round by round: inside a wave the lanes holding the
//                 same digit are found with 8 ballots (multisplit), their rank
//                 is the popcount of the lower peers; lanes, rounds, waves and
//                 tiles follow in index order, so the sort is stable.  A key's
//                 place = digits below it (scan of the totals) + its tile's
//                 offset + the keys of its digit before it in the tile.
// Scan and selection: per-tile sums (or flag counts), one workgroup scanning
// them, then every tile rescans itself from its offset (each wave its own
// contiguous quarter of the tile, with wave scans).
#include "primitives.hpp"

namespace dcg {

constexpr uint32_t kTileThreads = 256, kTileRounds = 16, kTile = kTileThreads * kTileRounds, kBins = 256;
constexpr uint32_t kTileWaves = kTileThreads / kWave;

static uint64_t tiles_of(uint64_t n) { return (n + kTile - 1) / kTile; }
static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

template <typename T>
__device__ __forceinline__ T wave_scan_incl(T x) {
    const uint32_t l = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const T y = __shfl_up(x, d, kWave);
        if (l >= (uint32_t)d) x += y;
    }
    return x;
}

// Inclusive scan over the workgroup (any multiple of 64 threads up to 1,024);
// `wsum` holds one word per wave.  Every thread gets its inclusive prefix and
// the workgroup total.  All threads must call it.
template <typename T>
__device__ __forceinline__ T block_scan_incl(T x, T *wsum, T &total) {
    const uint32_t w = threadIdx.x / kWave, nw = blockDim.x / kWave;
    const T incl = wave_scan_incl(x);
    if (lane_id() == kWave - 1) wsum[w] = incl;
    __syncthreads();
    T before = 0, all = 0;
    for (uint32_t q = 0; q < nw; ++q) {
        const T v = wsum[q];
        before += q < w ? v : (T)0;
        all += v;
    }
    __syncthreads();   // wsum is reused by the next call
    total = all;
    return before + incl;
}

template <typename T>
__device__ __forceinline__ T block_sum(T x, T *wsum) {
    T total;
    block_scan_incl(x, wsum, total);
    return total;
}

// ---- scan of per-tile figures ---------------------------------------------
// One 1,024-thread workgroup per row: row b of v (len entries from
// v + b * len) becomes its exclusive prefix sums; the row's total goes to
// totals[b] when totals is given.
template <typename T, typename TT>
__global__ __launch_bounds__(1024) void k_scan_rows(T *v, uint64_t len, TT *totals) {
    __shared__ T wsum[1024 / kWave];
    T *row = v + (uint64_t)blockIdx.x * len;
    T running = 0;
    for (uint64_t c0 = 0; c0 < len; c0 += blockDim.x) {
        const uint64_t i = c0 + threadIdx.x;
        const T x = i < len ? row[i] : (T)0;
        T tot;
        const T incl = block_scan_incl(x, wsum, tot);
        if (i < len) row[i] = running + incl - x;
        running += tot;
    }
    if (totals && threadIdx.x == 0) totals[blockIdx.x] = (TT)running;
}

// ---- radix sort -------------------------------------------------------------
// Lanes of the wave holding the same digit d (8 ballots: wave64 multisplit).
__device__ __forceinline__ uint64_t digit_peers(uint32_t d, bool valid) {
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (uint32_t b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bal = __ballot(valid && bit);
        peers &= bit ? bal : ~bal;
    }
    return peers;
}

// Wave w of a tile owns its 1,024 consecutive keys: round r holds keys
// tile0 + w * 1024 + r * 64 + lane, so every load is a coalesced 64-key run
// and the wave meets its keys in index order.
__device__ __forceinline__ uint64_t wave_first(uint64_t tile0) {
    return tile0 + (uint64_t)(threadIdx.x / kWave) * (kTileRounds * kWave) + lane_id();
}

template <typename K>
__global__ __launch_bounds__(256) void k_rs_hist(const K *keys, uint64_t n, uint32_t shift, uint32_t mask,
                                                 uint32_t *hist, uint32_t ntiles) {
    __shared__ uint32_t h[kBins];
    const uint32_t t = threadIdx.x;
    h[t] = 0;
    const uint64_t first = wave_first((uint64_t)blockIdx.x * kTile);
    K key[kTileRounds];
#pragma unroll
    for (uint32_t r = 0; r < kTileRounds; ++r) {
        const uint64_t i = first + r * kWave;
        key[r] = i < n ? keys[i] : (K)0;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < kTileRounds; ++r) {
        const bool valid = first + r * kWave < n;
        const uint32_t d = (uint32_t)(key[r] >> shift) & mask;
        const uint64_t peers = digit_peers(d, valid);
        // one LDS atomic per distinct digit of the wave
        if (valid && lanes_below(peers) == 0) atomicAdd(&h[d], (uint32_t)__popcll(peers));
    }
    __syncthreads();
    hist[(uint64_t)t * ntiles + blockIdx.x] = h[t];
}

// Scatter of one tile: every wave ranks its 16 rounds against a wave-private
// digit count in LDS (no workgroup barrier between rounds), then one pass
// over the 256 digits turns the four waves' counts into output bases: the
// digits below (scan of the totals) + the tile's offset + the waves before.
template <typename K, bool VALS>
__global__ __launch_bounds__(256) void k_rs_scatter(const K *keys_in, K *keys_out, const uint32_t *vals_in,
                                                    uint32_t *vals_out, uint64_t n, uint32_t shift, uint32_t mask,
                                                    const uint32_t *hist, const uint32_t *totals, uint32_t ntiles) {
    __shared__ uint32_t whist[kTileWaves][kBins];
    __shared__ uint32_t wsum[kTileWaves];
    const uint32_t t = threadIdx.x, w = t / kWave;
    const uint64_t first = wave_first((uint64_t)blockIdx.x * kTile);
    K key[kTileRounds];
    uint32_t val[kTileRounds];
#pragma unroll
    for (uint32_t r = 0; r < kTileRounds; ++r) {
        const uint64_t i = first + r * kWave;
        key[r] = i < n ? keys_in[i] : (K)0;
        if (VALS) val[r] = i < n ? vals_in[i] : 0u;
    }
#pragma unroll
    for (uint32_t q = 0; q < kTileWaves; ++q) whist[q][t] = 0;
    // digit t's base: the totals of the digits below + this tile's offset
    uint32_t dbase;
    {
        const uint32_t x = totals[t];
        uint32_t all;
        dbase = block_scan_incl(x, wsum, all) - x + hist[(uint64_t)t * ntiles + blockIdx.x];
    }
    // (block_scan_incl's barriers also order the zeroing of whist)
    uint32_t local[kTileRounds];
#pragma unroll
    for (uint32_t r = 0; r < kTileRounds; ++r) {
        const bool valid = first + r * kWave < n;
        const uint32_t d = (uint32_t)(key[r] >> shift) & mask;
        const uint64_t peers = digit_peers(d, valid);
        const uint32_t rank = lanes_below(peers);
        const uint32_t before = whist[w][d];   // every lane reads before the leader's update below
        local[r] = before + rank;
        if (valid && rank == 0) whist[w][d] = before + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    {
        uint32_t run = dbase;
#pragma unroll
        for (uint32_t q = 0; q < kTileWaves; ++q) {
            const uint32_t c = whist[q][t];
            whist[q][t] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < kTileRounds; ++r) {
        if (first + r * kWave < n) {
            const uint32_t d = (uint32_t)(key[r] >> shift) & mask;
            const uint32_t pos = whist[w][d] + local[r];
            keys_out[pos] = key[r];
            if (VALS) vals_out[pos] = val[r];
        }
    }
}

template <typename K>
size_t radix_sort_temp_bytes(uint64_t n, bool with_values) {
    const uint64_t nt = tiles_of(n) ? tiles_of(n) : 1;
    size_t b = n * sizeof(K);
    if (with_values) b += n * sizeof(uint32_t);
    return align256(b) + ((size_t)kBins * nt + kBins) * sizeof(uint32_t);
}

template <typename K>
void radix_sort(hipStream_t s, const K *keys_in, K *keys_out, const uint32_t *vals_in, uint32_t *vals_out, uint64_t n,
                int begin_bit, int end_bit, void *temp, size_t temp_bytes) {
    const bool vals = vals_in && vals_out;
    if (n == 0) return;
    if (n > 0xFFFFFFFFull) throw Error(DCG_ERR_UNSUPPORTED, "radix sort of more than 2^32 - 1 keys");
    if (begin_bit < 0 || end_bit > (int)(8 * sizeof(K)) || begin_bit > end_bit)
        throw Error(DCG_ERR_ARG, "radix sort bit range outside the key");
    if (keys_in == keys_out || (vals && vals_in == vals_out)) throw Error(DCG_ERR_ARG, "radix sort needs distinct buffers");
    if (temp_bytes < radix_sort_temp_bytes<K>(n, vals)) throw Error(DCG_ERR_ARG, "radix sort scratch too small");
    const int passes = (end_bit - begin_bit + 7) / 8;
    if (passes == 0) {
        DCG_HIP(hipMemcpyAsync(keys_out, keys_in, n * sizeof(K), hipMemcpyDeviceToDevice, s));
        if (vals) DCG_HIP(hipMemcpyAsync(vals_out, vals_in, n * 4, hipMemcpyDeviceToDevice, s));
        return;
    }
    const uint32_t nt = (uint32_t)tiles_of(n);
    uint8_t *const t0 = static_cast<uint8_t *>(temp);
    K *alt_k = reinterpret_cast<K *>(t0);
    uint32_t *alt_v = vals ? reinterpret_cast<uint32_t *>(t0 + n * sizeof(K)) : nullptr;
    uint32_t *hist = reinterpret_cast<uint32_t *>(t0 + align256(n * sizeof(K) + (vals ? n * 4 : 0)));
    uint32_t *totals = hist + (size_t)kBins * nt;
    const K *src_k = keys_in;
    const uint32_t *src_v = vals_in;
    for (int pass = 0; pass < passes; ++pass) {
        const uint32_t shift = (uint32_t)(begin_bit + 8 * pass);
        const int width = end_bit - (int)shift < 8 ? end_bit - (int)shift : 8;
        const uint32_t mask = (1u << width) - 1u;
        // the last pass writes keys_out
        const bool to_out = ((passes - 1 - pass) % 2) == 0;
        K *dst_k = to_out ? keys_out : alt_k;
        uint32_t *dst_v = vals ? (to_out ? vals_out : alt_v) : nullptr;
        hipLaunchKernelGGL(k_rs_hist<K>, dim3(nt), dim3(kTileThreads), 0, s, src_k, n, shift, mask, hist, nt);
        hipLaunchKernelGGL((k_scan_rows<uint32_t, uint32_t>), dim3(kBins), dim3(1024), 0, s, hist, (uint64_t)nt,
                           totals);
        if (vals)
            hipLaunchKernelGGL((k_rs_scatter<K, true>), dim3(nt), dim3(kTileThreads), 0, s, src_k, dst_k, src_v, dst_v,
                               n, shift, mask, hist, totals, nt);
        else
            hipLaunchKernelGGL((k_rs_scatter<K, false>), dim3(nt), dim3(kTileThreads), 0, s, src_k, dst_k,
                               (const uint32_t *)nullptr, (uint32_t *)nullptr, n, shift, mask, hist, totals, nt);
        src_k = dst_k;
        src_v = dst_v;
    }
    DCG_HIP(hipGetLastError());
}

// ---- exclusive scan -------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_scan_reduce(const T *in, uint64_t n, T *sums) {
    __shared__ T wsum[kTileWaves];
    const uint64_t base = (uint64_t)blockIdx.x * kTile;
    T x = 0;
    for (uint32_t r = 0; r < kTileRounds; ++r) {
        const uint64_t i = base + r * kTileThreads + threadIdx.x;
        if (i < n) x += in[i];
    }
    const T tot = block_sum(x, wsum);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Each wave scans its 1,024 keys with wave scans (no workgroup barrier), then
// the four wave totals give every wave its start.
template <typename T>
__global__ __launch_bounds__(256) void k_scan_apply(const T *in, T *out, uint64_t n, const T *offs) {
    __shared__ T wtot[kTileWaves];
    const uint32_t w = threadIdx.x / kWave;
    const uint64_t first = wave_first((uint64_t)blockIdx.x * kTile);
    T x[kTileRounds];
#pragma unroll
    for (uint32_t r = 0; r < kTileRounds; ++r) {
        const uint64_t i = first + r * kWave;
        x[r] = i < n ? in[i] : (T)0;
    }
    T run = 0;
#pragma unroll
    for (uint32_t r = 0; r < kTileRounds; ++r) {
        const T incl = wave_scan_incl(x[r]);
        const T sum = __shfl(incl, kWave - 1, kWave);
        x[r] = run + incl - x[r];   // exclusive prefix inside the wave's run
        run += sum;
    }
    if (lane_id() == 0) wtot[w] = run;
    __syncthreads();
    T base = offs[blockIdx.x];
    for (uint32_t q = 0; q < w; ++q) base += wtot[q];
#pragma unroll
    for (uint32_t r = 0; r < kTileRounds; ++r) {
        const uint64_t i = first + r * kWave;
        if (i < n) out[i] = base + x[r];
    }
}

template <typename T>
size_t scan_temp_bytes(uint64_t n) {
    return align256((tiles_of(n) ? tiles_of(n) : 1) * sizeof(T));
}

template <typename T>
void exclusive_sum(hipStream_t s, const T *in, T *out, uint64_t n, void *temp, size_t temp_bytes) {
    if (n == 0) return;
    if (in == out) throw Error(DCG_ERR_ARG, "exclusive sum needs distinct buffers");
    if (temp_bytes < scan_temp_bytes<T>(n)) throw Error(DCG_ERR_ARG, "scan scratch too small");
    const uint64_t nt = tiles_of(n);
    T *sums = static_cast<T *>(temp);
    hipLaunchKernelGGL(k_scan_reduce<T>, dim3((uint32_t)nt), dim3(kTileThreads), 0, s, in, n, sums);
    hipLaunchKernelGGL((k_scan_rows<T, T>), dim3(1), dim3(1024), 0, s, sums, nt, (T *)nullptr);
    hipLaunchKernelGGL(k_scan_apply<T>, dim3((uint32_t)nt), dim3(kTileThreads), 0, s, in, out, n, sums);
    DCG_HIP(hipGetLastError());
}

// ---- flagged selection ------------------------------------------------------
__global__ __launch_bounds__(256) void k_sel_count(const uint8_t *flags, uint64_t n, uint32_t *counts) {
    __shared__ uint32_t wsum[kTileWaves];
    const uint64_t base = (uint64_t)blockIdx.x * kTile;
    uint32_t c = 0;
    for (uint32_t r = 0; r < kTileRounds; ++r) {
        const uint64_t i = base + r * kTileThreads + threadIdx.x;
        c += (i < n && flags[i]) ? 1u : 0u;
    }
    const uint32_t tot = block_sum(c, wsum);
    if (threadIdx.x == 0) counts[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void k_sel_write(const uint8_t *flags, uint64_t n, const uint32_t *offs,
                                                   uint32_t *out) {
    __shared__ uint32_t wtot[kTileWaves];
    const uint32_t w = threadIdx.x / kWave;
    const uint64_t first = wave_first((uint64_t)blockIdx.x * kTile);
    bool f[kTileRounds];
#pragma unroll
    for (uint32_t r = 0; r < kTileRounds; ++r) {
        const uint64_t i = first + r * kWave;
        f[r] = i < n && flags[i];
    }
    uint32_t pos[kTileRounds], run = 0;
#pragma unroll
    for (uint32_t r = 0; r < kTileRounds; ++r) {
        const uint64_t m = __ballot(f[r]);
        pos[r] = run + lanes_below(m);
        run += (uint32_t)__popcll(m);
    }
    if (lane_id() == 0) wtot[w] = run;
    __syncthreads();
    uint32_t base = offs[blockIdx.x];
    for (uint32_t q = 0; q < w; ++q) base += wtot[q];
#pragma unroll
    for (uint32_t r = 0; r < kTileRounds; ++r)
        if (f[r]) out[base + pos[r]] = (uint32_t)(first + r * kWave);
}

size_t select_temp_bytes(uint64_t n) { return align256((tiles_of(n) ? tiles_of(n) : 1) * sizeof(uint32_t)); }

void select_flagged(hipStream_t s, const uint8_t *flags, uint64_t n, uint32_t *out, unsigned long long *count,
                    void *temp, size_t temp_bytes) {
    if (n > 0xFFFFFFFFull) throw Error(DCG_ERR_UNSUPPORTED, "selection over more than 2^32 - 1 flags");
    if (temp_bytes < select_temp_bytes(n)) throw Error(DCG_ERR_ARG, "selection scratch too small");
    if (n == 0) {
        fill<unsigned long long>(s, count, 1, 0ull);
        return;
    }
    const uint64_t nt = tiles_of(n);
    uint32_t *counts = static_cast<uint32_t *>(temp);
    hipLaunchKernelGGL(k_sel_count, dim3((uint32_t)nt), dim3(kTileThreads), 0, s, flags, n, counts);
    hipLaunchKernelGGL((k_scan_rows<uint32_t, unsigned long long>), dim3(1), dim3(1024), 0, s, counts, nt, count);
    hipLaunchKernelGGL(k_sel_write, dim3((uint32_t)nt), dim3(kTileThreads), 0, s, flags, n, counts, out);
    DCG_HIP(hipGetLastError());
}

// ---- descending float keys ----------------------------------------------------
__global__ void k_float_desc_keys(const float *in, uint32_t *out, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t u = __float_as_uint(in[i]);
        if (u == 0x80000000u) u = 0u;                          // -0 ranks as +0
        u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);        // ascending order of the floats
        out[i] = ~u;                                          // descending
    }
}

void float_desc_keys(hipStream_t s, const float *in, uint32_t *out, uint64_t n) {
    if (!n) return;
    hipLaunchKernelGGL(k_float_desc_keys, dim3(grid_for(n)), dim3(256), 0, s, in, out, n);
    DCG_HIP(hipGetLastError());
}

template size_t radix_sort_temp_bytes<uint32_t>(uint64_t, bool);
template size_t radix_sort_temp_bytes<uint64_t>(uint64_t, bool);
template void radix_sort<uint32_t>(hipStream_t, const uint32_t *, uint32_t *, const uint32_t *, uint32_t *, uint64_t,
                                   int, int, void *, size_t);
template void radix_sort<uint64_t>(hipStream_t, const uint64_t *, uint64_t *, const uint32_t *, uint32_t *, uint64_t,
                                   int, int, void *, size_t);
template size_t scan_temp_bytes<uint32_t>(uint64_t);
template size_t scan_temp_bytes<uint64_t>(uint64_t);
template void exclusive_sum<uint32_t>(hipStream_t, const uint32_t *, uint32_t *, uint64_t, void *, size_t);
template void exclusive_sum<uint64_t>(hipStream_t, const uint64_t *, uint64_t *, uint64_t, void *, size_t);

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
        tmp.alloc(radix_sort_temp_bytes<uint64_t>(n, vals != nullptr));
        DCG_HIP(hipMemcpyAsync(ki.p, keys, n * 8, hipMemcpyHostToDevice, s));
        if (vals) DCG_HIP(hipMemcpyAsync(vi.p, vals, n * 4, hipMemcpyHostToDevice, s));
        radix_sort<uint64_t>(s, ki.p, ko.p, vals ? vi.p : nullptr, vals ? vo.p : nullptr, n, begin_bit, end_bit, tmp.p,
                             tmp.n);
        DCG_HIP(hipMemcpyAsync(keys, ko.p, n * 8, hipMemcpyDeviceToHost, s));
        if (vals) DCG_HIP(hipMemcpyAsync(vals, vo.p, n * 4, hipMemcpyDeviceToHost, s));
        DCG_HIP(hipStreamSynchronize(s));
        DCG_HIP(hipStreamDestroy(s));
    } catch (const Error &e) {
        return e.status;
    }
    return DCG_OK;
}
