// cache_policy.hip -- device chunk cache over the pinned host chunk store.
//
// The cache holds whole 65,536-edge chunks in HBM frames.  Kernels resolve a
// slot through chunk_frame[]: a resident chunk is read from its frame, any
// other chunk zero-copy from pinned host memory.  Writes go through to both
// copies (store_slot), so the host copy is always current and an eviction
// only unmaps the frame.
//
// HOTNESS policy (the reference's plain policy, driver.cuh:486-579 and
// framework.cuh:1264-1567, at chunk granularity): score = 64-byte lines read
// over the last four snapshots (row passes and delete windows; the host link's
// cost is per line), ties by most recent use (LRU), then by chunk id; the
// top `frames` chunks are admitted.  A refresh happens only when the admitted
// set differs from the resident set (cache_refresh_gate.h).
#include <algorithm>
#include <numeric>

#include "engine.hpp"
#include "mlp.hpp"
#include "primitives.hpp"

namespace dcg {

// Load whole chunks from pinned host memory into HBM frames: one workgroup
// per chunk, 16-byte vector accesses.
__global__ __launch_bounds__(256) void k_move_chunks(uint32_t *frames, const uint32_t *host, const int64_t *list,
                                                     uint32_t count) {
    for (uint32_t i = blockIdx.x; i < count; i += gridDim.x) {
        const int64_t frame = list[2 * i], chunk = list[2 * i + 1];
        uint4 *f = reinterpret_cast<uint4 *>(frames + ((uint64_t)frame << kChunkLog2));
        const uint4 *h = reinterpret_cast<const uint4 *>(host + ((uint64_t)chunk << kChunkLog2));
        constexpr uint32_t kVec = kChunkEdges / 4;
        for (uint32_t j = threadIdx.x; j < kVec; j += blockDim.x) f[j] = h[j];
    }
}

// hist[c] = (reads this snapshot, previous three), reads reset afterwards.
__global__ void k_shift_hist(float *hist, float *reads, uint64_t chunks, float *score) {
    for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < chunks; c += (uint64_t)gridDim.x * blockDim.x) {
        float4 h = reinterpret_cast<float4 *>(hist)[c];
        h.w = h.z;
        h.z = h.y;
        h.y = h.x;
        h.x = reads[c];
        reinterpret_cast<float4 *>(hist)[c] = h;
        score[c] = h.x + h.y + h.z + h.w;
    }
}

void Engine::cache_init() {
    const uint64_t C = cache.chunks;
    cache.d_chunk_frame.alloc(C);
    cache.chunk_frame.assign(C, -1);
    cache.frame_chunk.assign(cache.frames, -1);
    cache.last_use.assign(C, 0);
    cache.d_reads.alloc(2 * C);
    fill<float>(stream, cache.d_reads.p, 2 * C, 0.f);
    cache.d_hist.alloc(4 * C);
    fill<float>(stream, cache.d_hist.p, 4 * C, 0.f);
    cache.d_touch.alloc(C);
    fill<float>(stream, cache.d_touch.p, C, 0.f);
    cache.d_score.alloc(C);
    cache.d_rows_last.alloc(C);
    fill<float>(stream, cache.d_rows_last.p, C, 0.f);
    if (cfg.placement == DCG_PLACE_DEVICE) {
        for (uint64_t c = 0; c < C; ++c) cache.chunk_frame[c] = (int32_t)c;
        cache.frame_chunk.resize(C);
        for (uint64_t c = 0; c < C; ++c) cache.frame_chunk[c] = (int64_t)c;
    } else if (cache.frames) {
        // Initial admission: the lowest chunks.  The load places rows by
        // descending capacity, so these hold the highest-degree rows (the
        // reference's initial LoadCache admits by hotness after Start; the
        // first maintenance call replaces this with the measured ranking).
        std::vector<int64_t> list;
        for (uint64_t f = 0; f < cache.frames; ++f) {
            cache.chunk_frame[f] = (int32_t)f;
            cache.frame_chunk[f] = (int64_t)f;
            list.push_back((int64_t)f);
            list.push_back((int64_t)f);
        }
        DevBuf<int64_t> dl;
        dl.alloc(list.size());
        DCG_HIP(hipMemcpyAsync(dl.p, list.data(), list.size() * 8, hipMemcpyHostToDevice, stream));
        hipLaunchKernelGGL(k_move_chunks, dim3(grid_for(cache.frames, 1, 4096)), dim3(256), 0, stream,
                           cache.d_frames.p, h_slab, dl.p, (uint32_t)cache.frames);
        DCG_HIP(hipGetLastError());
        cache.h2d += cache.frames * kChunkEdges * 4;
        sync();
    }
    DCG_HIP(hipMemcpyAsync(cache.d_chunk_frame.p, cache.chunk_frame.data(), C * 4, hipMemcpyHostToDevice, stream));
    sync();
}

// Make the resident set equal to `desired` (chunk ids, at most `frames`).
void Engine::cache_apply_plan(const std::vector<uint64_t> &desired, dcg_cache_stats *st) {
    const uint64_t C = cache.chunks;
    std::vector<uint8_t> want(C, 0);
    for (uint64_t c : desired) want[c] = 1;
    std::vector<int64_t> evict_list, free_frames;
    for (uint64_t f = 0; f < cache.frames; ++f) {
        const int64_t c = cache.frame_chunk[f];
        if (c < 0) {
            free_frames.push_back((int64_t)f);
        } else if (!want[c]) {
            evict_list.push_back((int64_t)f);
            evict_list.push_back(c);
        }
    }
    std::vector<int64_t> load_list;
    std::vector<uint64_t> to_load;
    for (uint64_t c : desired)
        if (cache.chunk_frame[c] < 0) to_load.push_back(c);
    if (evict_list.empty() && to_load.empty()) return;
    // the host copy is current (write-through): evicting only unmaps the frame
    for (size_t i = 0; i < evict_list.size(); i += 2) {
        const int64_t f = evict_list[i], c = evict_list[i + 1];
        cache.chunk_frame[c] = -1;
        cache.frame_chunk[f] = -1;
        free_frames.push_back(f);
    }
    size_t k = 0;
    for (uint64_t c : to_load) {
        if (k >= free_frames.size()) break;
        const int64_t f = free_frames[k++];
        cache.chunk_frame[c] = (int32_t)f;
        cache.frame_chunk[f] = (int64_t)c;
        load_list.push_back(f);
        load_list.push_back((int64_t)c);
    }
    if (!load_list.empty()) {
        // the list buffer is kept between decisions (the stream synchronises below before it is reused)
        cache.d_load.ensure(load_list.size());
        DCG_HIP(hipMemcpyAsync(cache.d_load.p, load_list.data(), load_list.size() * 8, hipMemcpyHostToDevice,
                               stream));
        hipLaunchKernelGGL(k_move_chunks, dim3(grid_for(load_list.size() / 2, 1, 4096)), dim3(256), 0, stream,
                           cache.d_frames.p, h_slab, cache.d_load.p, (uint32_t)(load_list.size() / 2));
        DCG_HIP(hipGetLastError());
    }
    DCG_HIP(hipMemcpyAsync(cache.d_chunk_frame.p, cache.chunk_frame.data(), C * 4, hipMemcpyHostToDevice, stream));
    sync();
    const uint64_t ln = load_list.size() / 2;
    cache.evicted += evict_list.size() / 2;
    cache.loaded += ln;
    cache.h2d += ln * kChunkEdges * 4;
    if (st) {
        st->evicted_chunks += evict_list.size() / 2;
        st->loaded_chunks += ln;
        st->h2d_bytes += ln * kChunkEdges * 4;
    }
}

static void fill_common(Engine &e, dcg_cache_stats *st) {
    if (!st) return;
    if (e.cache.mlp) mlp_stats(e, st);
    st->frames = e.cache.frames;
    st->total_chunks = e.cache.chunks;
    uint64_t r = 0;
    for (uint64_t c = 0; c < e.cache.chunks; ++c) r += e.cache.chunk_frame[c] >= 0;
    st->resident_chunks = r;
}

// Admissions per decision: 1/128 of the frames (at least 4).  A load moves a
// whole 256 KiB chunk over the host link while the repair kernels read only
// what they touch (and the delete search reads one line per target), so churn
// near the ranking boundary costs more link time than the hits it buys; a
// changed ranking is followed over several snapshots instead.
uint64_t Engine::move_budget() const { return std::max<uint64_t>(4, cache.frames / 128); }

// Admission plan shared by both policies.  Candidates are ranked by score
// (ties: most recent use, then chunk id).  A chunk is admitted only if its
// score reaches `admit_min` line reads, it displaces a resident chunk only if
// its score exceeds the resident's by `margin`, and at most move_budget()
// chunks are admitted, so only the budget strongest candidates and the budget
// weakest residents are ranked (partial selection, then a sort of those).
std::vector<uint64_t> Engine::plan_admission(const std::vector<float> &score, double admit_min, double margin) {
    const uint64_t C = cache.chunks, F = cache.frames;
    const uint64_t budget = move_budget();
    auto better = [&](uint64_t a, uint64_t b) {
        if (score[a] != score[b]) return score[a] > score[b];
        if (cache.last_use[a] != cache.last_use[b]) return cache.last_use[a] > cache.last_use[b];
        return a < b;
    };
    auto weaker = [&](uint64_t a, uint64_t b) { return better(b, a); };
    // resident chunks: the `budget` weakest, weakest first
    std::vector<uint64_t> res;
    res.reserve(F);
    for (uint64_t f = 0; f < F; ++f)
        if (cache.frame_chunk[f] >= 0) res.push_back((uint64_t)cache.frame_chunk[f]);
    const size_t rk = std::min<size_t>(budget, res.size());
    if (rk < res.size()) std::nth_element(res.begin(), res.begin() + (ptrdiff_t)rk, res.end(), weaker);
    std::sort(res.begin(), res.begin() + (ptrdiff_t)rk, weaker);
    // candidates: the `budget` strongest, strongest first
    std::vector<uint64_t> cand;
    for (uint64_t c = 0; c < C; ++c)
        if (cache.chunk_frame[c] < 0 && score[c] >= admit_min) cand.push_back(c);
    if (budget < cand.size()) {
        std::nth_element(cand.begin(), cand.begin() + (ptrdiff_t)budget, cand.end(), better);
        cand.resize(budget);
    }
    std::sort(cand.begin(), cand.end(), better);
    const uint64_t free_frames = F - res.size();
    std::vector<uint8_t> drop(rk, 0);
    std::vector<uint64_t> admit;
    size_t ci = 0, ri = 0;
    for (; ci < cand.size() && admit.size() < free_frames && admit.size() < budget; ++ci) admit.push_back(cand[ci]);
    for (; ci < cand.size() && ri < rk && admit.size() < budget; ++ci, ++ri) {
        if (!(score[cand[ci]] > margin * std::max(score[res[ri]], 1.0f))) break;
        drop[ri] = 1;
        admit.push_back(cand[ci]);
    }
    std::vector<uint64_t> keep;
    keep.reserve(F);
    for (size_t i = 0; i < res.size(); ++i)
        if (i >= rk || !drop[i]) keep.push_back(res[i]);
    for (uint64_t c : admit) keep.push_back(c);
    return keep;
}

// After the static convergence: admit the chunks with the most line reads in
// it, without the per-decision budget (the reference's LoadCache after Start
// admits by hotness; framework.cuh:1321-1350).  Ties keep resident chunks, then
// the lower chunk id.  The reads stay in the window of the first decision.
void Engine::cache_admit_measured() {
    if (cfg.placement == DCG_PLACE_DEVICE || cache.frames == 0 || cache.frames >= cache.chunks) return;
    const uint64_t C = cache.chunks, F = cache.frames;
    std::vector<float> reads(C);
    DCG_HIP(hipMemcpyAsync(reads.data(), cache.d_reads.p, C * 4, hipMemcpyDeviceToHost, stream));
    sync();
    std::vector<uint64_t> order(C);
    for (uint64_t c = 0; c < C; ++c) order[c] = c;
    auto better = [&](uint64_t a, uint64_t b) {
        if (reads[a] != reads[b]) return reads[a] > reads[b];
        const bool ra = cache.chunk_frame[a] >= 0, rb = cache.chunk_frame[b] >= 0;
        if (ra != rb) return ra;
        return a < b;
    };
    std::nth_element(order.begin(), order.begin() + (ptrdiff_t)F, order.end(), better);
    order.resize(F);
    cache_apply_plan(order, nullptr);
}

void Engine::cache_maintain(dcg_cache_stats *st) {
    const double t0 = now_ms();
    if (st) *st = dcg_cache_stats{};
    snapshots++;
    if (cfg.placement == DCG_PLACE_DEVICE || cache.frames == 0) {
        // a device store tallies no reads (its kernels get no tally array)
        if (cfg.placement != DCG_PLACE_DEVICE) fill<float>(stream, cache.d_reads.p, 2 * cache.chunks, 0.f);
        fill_common(*this, st);
        if (st) st->ms = now_ms() - t0;
        return;
    }
    const uint64_t C = cache.chunks;
    std::vector<float> reads(C);
    DCG_HIP(hipMemcpyAsync(reads.data(), cache.d_reads.p, C * 4, hipMemcpyDeviceToHost, stream));
    hipLaunchKernelGGL(k_shift_hist, dim3(grid_for(C)), dim3(256), 0, stream, cache.d_hist.p, cache.d_reads.p, C,
                       cache.d_score.p);
    std::vector<float> score(C);
    DCG_HIP(hipMemcpyAsync(score.data(), cache.d_score.p, C * 4, hipMemcpyDeviceToHost, stream));
    sync();
    if (cfg.cache_policy == DCG_CACHE_LEARNED) cache.hot_score = score;   // the next decision may take the hotness plan
    for (uint64_t c = 0; c < C; ++c)
        if (reads[c] > 0) cache.last_use[c] = (uint32_t)snapshots;
    if (cfg.cache_policy == DCG_CACHE_LEARNED && cache.mlp) {
        mlp_observe(*this);   // online training on this snapshot's labels
    }
    DCG_HIP(hipMemcpyAsync(cache.d_rows_last.p, cache.d_reads.p + C, C * 4, hipMemcpyDeviceToDevice, stream));
    // the learned policy admits at dcg_cache_prefetch, before the next batch
    if (cfg.cache_policy == DCG_CACHE_HOTNESS) {
        // window of 4 snapshots: admit chunks whose reads would repay a load every snapshot
        cache_apply_plan(plan_admission(score, 4.0 * kLoadLines, 1.5), st);
    }
    fill<float>(stream, cache.d_reads.p, 2 * C, 0.f);
    fill_common(*this, st);
    if (st) st->ms = now_ms() - t0;
}

void Engine::cache_prefetch(uint64_t ab, uint64_t ac, uint64_t db, uint64_t dc, dcg_cache_stats *st) {
    const double t0 = now_ms();
    if (st) *st = dcg_cache_stats{};
    if (cfg.placement == DCG_PLACE_DEVICE || cache.frames == 0 || cfg.cache_policy != DCG_CACHE_LEARNED) {
        fill_common(*this, st);
        return;
    }
    if (!cache.mlp) mlp_create(*this);
    std::vector<uint64_t> desired = mlp_plan(*this, ab, ac, db, dc);
    cache_apply_plan(desired, st);
    fill_common(*this, st);
    if (st) st->ms = now_ms() - t0;
}

}  // namespace dcg

// ---- the reference's vertex hotness ranking ------------------------------------
namespace dcg {

// Device placement (no read list): fold the snapshot's out-row expansions
// into the window's current byte.
__global__ void k_window_fold(uint32_t *hot_out, uint32_t *win, uint32_t n) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        const uint32_t h = hot_out[v];
        if (h) {
            win[v] = window_add(win[v], h);
            hot_out[v] = 0;
        }
    }
}

// comp_hotness_sssp (driver.cuh:495-517): score = the window's four counters
// summed, 0 where the query has not reached the vertex; the sort key
// (1023 - score) << idbits | v orders by descending score, then ascending id.
__global__ void k_hot_keys(const uint32_t *win, const VSlot *st, uint32_t n, uint32_t idbits, uint64_t *keys) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        const uint32_t w = win[v];
        uint32_t score = (w & 0xFFu) + ((w >> 8) & 0xFFu) + ((w >> 16) & 0xFFu) + (w >> 24);
        if (st && (uint32_t)(st[v].st >> 32) == kInf) score = 0;
        keys[v] = ((uint64_t)(1023u - score) << idbits) | v;
    }
}

constexpr uint32_t kRankChunk = 1024;   // ranked entries per degree-sum block

// extract_vtx_degree (driver.cuh:548-560): live out-degree of each ranked
// vertex, and the degree sum of every kRankChunk ranked entries.
__global__ __launch_bounds__(256) void k_hot_degrees(const uint64_t *sorted, uint64_t idmask, const RowDesc *out,
                                                     uint32_t n, uint32_t *deg, unsigned long long *bsum) {
    __shared__ unsigned long long part[256 / kWave];
    const uint64_t base = (uint64_t)blockIdx.x * kRankChunk;
    unsigned long long s = 0;
    for (uint32_t j = threadIdx.x; j < kRankChunk; j += blockDim.x) {
        const uint64_t i = base + j;
        if (i >= n) break;
        const uint32_t d = row_live(out[(uint32_t)(sorted[i] & idmask)]);
        deg[i] = d;
        s += d;
    }
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, kWave);
    if (lane_id() == 0) part[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (uint32_t w = 0; w < blockDim.x / kWave; ++w) t += part[w];
        bsum[blockIdx.x] = t;
    }
}

// reset_hotness (driver.cuh:562-579): every counter moves one snapshot back.
__global__ void k_window_shift(uint32_t *win, uint32_t n) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) win[v] <<= 8;
}

// A host store folds its listed rows' counters into the window after every
// stage; a device store lists no rows, so its counters are folded here.
void Engine::fold_window() {
    if (cfg.placement == DCG_PLACE_DEVICE)
        hipLaunchKernelGGL(k_window_fold, dim3(grid_for(n)), dim3(256), 0, stream, d_hot_out.p, d_hot_window.p, n);
}

uint64_t Engine::hot_candidates(uint32_t query, uint64_t cap, uint32_t *ids, uint64_t ids_len) {
    fold_window();
    uint32_t idbits = 1;
    while (idbits < 32 && (1ull << idbits) < n) ++idbits;
    DevBuf<uint64_t> keys, sorted;
    DevBuf<uint32_t> deg;
    DevBuf<unsigned long long> bsum;
    DevBuf<uint8_t> tmp;
    keys.alloc(n);
    sorted.alloc(n);
    deg.alloc(n);
    const uint64_t nb = (n + kRankChunk - 1) / kRankChunk;
    bsum.alloc(nb);
    tmp.alloc(radix_sort_temp_bytes<uint64_t>(n, false));
    const VSlot *st = app == DCG_APP_PR ? nullptr : d_vs.p + (uint64_t)query * n;
    hipLaunchKernelGGL(k_hot_keys, dim3(grid_for(n)), dim3(256), 0, stream, d_hot_window.p, st, n, idbits, keys.p);
    radix_sort<uint64_t>(stream, keys.p, sorted.p, nullptr, nullptr, n, 0, (int)idbits + 10, tmp.p, tmp.n);
    const uint64_t idmask = (1ull << idbits) - 1;
    hipLaunchKernelGGL(k_hot_degrees, dim3((uint32_t)nb), dim3(256), 0, stream, sorted.p, idmask, d_out.p, n, deg.p,
                       bsum.p);
    DCG_HIP(hipGetLastError());
    std::vector<unsigned long long> hb(nb);
    DCG_HIP(hipMemcpyAsync(hb.data(), bsum.p, nb * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
    sync();
    // CacheCandidateCount (framework.cuh:1321-1350): the ranked vertices whose
    // inclusive degree prefix stays below the cap; all of them when the total
    // fits, none without a cache
    uint64_t count = 0;
    if (cap > 0) {
        unsigned long long run = 0;
        uint64_t b = 0;
        while (b < nb && run + hb[b] < cap) run += hb[b++];
        if (b == nb) {
            count = n;
        } else {
            const uint64_t lo = b * kRankChunk, len = std::min<uint64_t>(kRankChunk, n - lo);
            std::vector<uint32_t> hd(len);
            DCG_HIP(hipMemcpyAsync(hd.data(), deg.p + lo, len * 4, hipMemcpyDeviceToHost, stream));
            sync();
            count = lo;
            for (uint64_t j = 0; j < len && run + hd[j] < cap; ++j) {
                run += hd[j];
                ++count;
            }
        }
    }
    // the caller's ids; with a cache-candidate trace open, the whole admitted
    // prefix and its degrees for the next trace record
    const uint64_t k = std::min<uint64_t>(count, f1_on ? n : ids_len);
    if (k) {
        std::vector<uint64_t> hk(k);
        DCG_HIP(hipMemcpyAsync(hk.data(), sorted.p, k * 8, hipMemcpyDeviceToHost, stream));
        if (f1_on) {
            rank_deg.resize(count);
            DCG_HIP(hipMemcpyAsync(rank_deg.data(), deg.p, count * 4, hipMemcpyDeviceToHost, stream));
        }
        sync();
        for (uint64_t i = 0; i < std::min<uint64_t>(k, ids_len); ++i) ids[i] = (uint32_t)(hk[i] & idmask);
        if (f1_on) {
            rank_ids.resize(count);
            for (uint64_t i = 0; i < count; ++i) rank_ids[i] = (uint32_t)(hk[i] & idmask);
        }
    } else if (f1_on) {
        rank_ids.clear();
        rank_deg.clear();
    }
    rank_cap = cap;
    hipLaunchKernelGGL(k_window_shift, dim3(grid_for(n)), dim3(256), 0, stream, d_hot_window.p, n);
    DCG_HIP(hipGetLastError());
    sync();
    return count;
}

}  // namespace dcg
