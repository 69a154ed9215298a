// store.hip -- chunked-CSR topology store: load, GPU mutation, export.
//
// Rows are power-of-two blocks (minimum 4 slots) like the reference's
// source-local chunk store (include/groute/graphs/source_local_chunk_store.h:133-164),
// but the mutation itself runs on the GPU: one wave64 per touched row.
//  * delete: remove the first occurrence per request, survivors keep their order
//    (source_local_chunk_store.h:685-821); requests with no remaining occurrence
//    are counted as missing_deletes.
//  * add: append in request order (:809-815); a row that outgrows its block moves
//    to a fresh block of the next power of two, carved from the store tail by a
//    per-class bump allocator so blocks stay aligned to their capacity.
// Both the out-rows and the in-rows (reverse index, the role of
// include/framework/dynamic_reverse_index.h) live in the same slot space.

#include "engine.hpp"
#include "primitives.hpp"

namespace dcg {

enum Counter : int {
    C_MISSING = 0,
    C_INVALID = 1,
    C_GROUPS_OUT = 2,
    C_GROUPS_IN = 3,
    C_CHANGED = 4,
    C_RELOCATED = 5,
    C_BAD_ID = 6,
    C_CACHED_READS = 7,
    C_HOST_READS = 8,
    C_SLOTS_READ = 9,
    C_SLOTS_WRITTEN = 10,
    C_NBIG = 11,
    C_W_READ = 12,       // wave-path delete: slots read, slots written, rows
    C_W_WRITTEN = 13,
    C_W_ROWS = 14,
    C_B_ROWS = 15,       // workgroup-path delete rows
    C_W_HOST_LINES = 16, // wave-path delete: 64-byte lines read from pinned host memory
    C_NBIG_IN = 17,      // large-row list of the in direction (C_NBIG: out direction)
    C_NUM = 18
};
static_assert(C_NUM <= Engine::kCounterSlots, "pinned counter mirror too small");

// d_tail layout
// Store tail block (u64 words): bump pointer, overflow flag, and per capacity
// class: relocation demand, bump cursor, rank counter, blocks taken from the
// free pool, pool segment start, pool blocks; then the retired-block cursor.
constexpr int T_TAIL = 0, T_OVERFLOW = 1, T_COUNT = 2, T_CURSOR = 2 + kClasses, T_RANK = 2 + 2 * kClasses,
              T_TAKE = 2 + 3 * kClasses, T_PSTART = 2 + 4 * kClasses, T_FREE = 2 + 5 * kClasses,
              T_RETIRED = 2 + 6 * kClasses, T_NUM = 3 + 6 * kClasses;
static_assert(T_NUM == Engine::kTailWords, "pinned tail mirror size");
constexpr uint64_t kPoolOff = (1ull << 40) - 1;   // pool key = class << 40 | slot offset
constexpr int kPoolKeyBits = 46;                   // class < 64; taken entries (~0) sort last
static_assert(kClasses < 64, "pool classes fit 6 bits below the taken marker");

StoreView Engine::view() {
    StoreView v;
    v.out = d_out.p;
    v.in = d_in.p;
    v.host = h_slab;
    v.frames = cache.d_frames.p;
    v.chunk_frame = cache.d_chunk_frame.p;
    v.hot_out = d_hot_out.p;
    v.hot_in = d_hot_in.p;
    v.hot_window = d_hot_window.p;
    v.hot_list = h_slab ? d_hot_list.p : nullptr;
    v.hot_count = d_hot_count.p;
    return v;
}

__device__ __forceinline__ uint32_t log2_u64(uint64_t x) { return 63u - (uint32_t)__clzll((long long)x); }

// ---------------------------------------------------------------- load kernels
__global__ void k_check_ids(const uint32_t *src, const uint32_t *dst, uint64_t m, uint32_t n,
                            unsigned long long *counters, uint32_t *deg) {
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < m;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t u = src[e], v = dst[e];
        if (u >= n || v >= n) {
            atomicAdd(&counters[C_BAD_ID], 1ull);
            continue;
        }
        atomicAdd(&deg[u], 1u);
        atomicAdd(&deg[(uint64_t)n + v], 1u);
    }
}

// Placement key: larger classes first, then row id (out rows 0..n-1, in rows n..2n-1).
__global__ void k_place_keys(const uint32_t *deg, uint64_t rows, uint64_t *keys) {
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < rows;
         r += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t d = deg[r];
        if (!d) {
            keys[r] = ~0ull;
            continue;
        }
        uint64_t c = d < kMinRowCap ? kMinRowCap : d;
        const uint32_t cls = 64u - (uint32_t)__clzll((long long)(c - 1));  // ceil log2
        keys[r] = ((uint64_t)(63u - cls) << 36) | r;
    }
}

__global__ void k_place_caps(const uint64_t *keys, uint64_t rows, uint64_t *caps) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < rows;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[i];
        caps[i] = k == ~0ull ? 0ull : (1ull << (63u - (uint32_t)(k >> 36)));
    }
}

__global__ void k_place_desc(const uint64_t *keys, const uint64_t *offs, const uint32_t *deg, uint64_t rows,
                             uint32_t n, RowDesc *out, RowDesc *in) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < rows;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[i];
        if (k == ~0ull) continue;
        const uint64_t r = k & ((1ull << 36) - 1);
        RowDesc d;
        d.off = offs[i];
        d.ext = deg[r];
        d.meta = make_meta(63u - (uint32_t)(k >> 36), 0);
        if (r < n) out[r] = d;
        else in[r - n] = d;
    }
}

__global__ void k_widen(const uint32_t *a, uint64_t *b, uint64_t k) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < k; i += (uint64_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

__global__ void k_zero_desc(RowDesc *d, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        d[i] = RowDesc{0, 0, 0};
}

// Scatter a row-sorted edge list into the slot space (row order = input order
// because the sort is stable).
__global__ void k_scatter_rows(const uint32_t *keys, const uint32_t *vals, uint64_t m, const uint64_t *rp,
                               const RowDesc *desc, uint32_t *slots) {
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < m;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t r = keys[e];
        slots[desc[r].off + (e - rp[r])] = vals[e];
    }
}

constexpr uint32_t kFenceStride = 16;   // prefix slots per fence (one 64-byte line)
constexpr uint32_t kFenceMin = 16;      // prefixes shorter than this (one line or less) get no fences

// Length of each row's longest non-decreasing prefix (initialised to the degree):
// the first descent position of a row-sorted edge list.
__global__ void k_sorted_prefix(const uint32_t *keys, const uint32_t *vals, uint64_t m, const uint64_t *rp,
                                uint32_t *sorted) {
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < m;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t r = keys[e];
        if (e > rp[r] && vals[e] < vals[e - 1]) atomicMin(&sorted[r], (uint32_t)(e - rp[r]));
    }
}

// Fence capacity of each row: one fence per 16 slots of a sorted prefix of at
// least kFenceMin entries (shorter prefixes are scanned, not searched).
__global__ void k_fence_caps(const uint32_t *sorted, uint32_t n, uint32_t *cap) {
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r <= n; r += (uint64_t)gridDim.x * blockDim.x)
        cap[r] = r < n && sorted[r] >= kFenceMin ? (sorted[r] + kFenceStride - 1) / kFenceStride : 0u;
}

// Fence j of row r = the id at slot 16j (row-sorted edge list, before placement).
__global__ void k_fence_fill(const uint32_t *keys, const uint32_t *vals, uint64_t m, const uint64_t *rp,
                             const uint32_t *fbase, uint32_t *fence) {
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < m;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t r = keys[e];
        const uint64_t i = e - rp[r];
        if (i % kFenceStride) continue;
        const uint64_t j = i / kFenceStride;
        if (j < fbase[r + 1] - fbase[r]) fence[fbase[r] + j] = vals[e];
    }
}

__global__ void k_max_u32(const uint32_t *a, uint64_t k, uint32_t *out) {
    uint32_t m = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < k; i += (uint64_t)gridDim.x * blockDim.x)
        m = a[i] > m ? a[i] : m;
    for (int d = 32; d >= 1; d >>= 1) {
        const uint32_t o = __shfl_xor(m, d, kWave);
        m = o > m ? o : m;
    }
    if (lane_id() == 0 && m) atomicMax(out, m);
}

__global__ void k_copy_u32(const uint32_t *a, uint32_t *b, uint64_t k) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < k; i += (uint64_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

static int bits_for(uint64_t maxval) {
    int b = 1;
    while (b < 64 && (maxval >> b)) b++;
    return b;
}

void Engine::build_fences(const uint32_t *keys, const uint32_t *vals, uint64_t m, const uint64_t *rp,
                          const uint32_t *sorted, DevBuf<uint32_t> &fbase, DevBuf<uint32_t> &fence) {
    DevBuf<uint32_t> caps;
    caps.alloc((uint64_t)n + 1);
    fbase.alloc((uint64_t)n + 1);
    hipLaunchKernelGGL(k_fence_caps, dim3(grid_for((uint64_t)n + 1)), dim3(256), 0, stream, sorted, n, caps.p);
    const size_t tb = scan_temp_bytes<uint32_t>((uint64_t)n + 1);
    exclusive_sum<uint32_t>(stream, caps.p, fbase.p, (uint64_t)n + 1, temp(tb), tb);
    uint32_t total = 0;
    DCG_HIP(hipMemcpyAsync(&total, fbase.p + n, 4, hipMemcpyDeviceToHost, stream));
    sync();
    fence.alloc(total ? total : 1);
    hipLaunchKernelGGL(k_fence_fill, dim3(grid_for(m)), dim3(256), 0, stream, keys, vals, m, rp, fbase.p, fence.p);
    DCG_HIP(hipGetLastError());
}

void Engine::load_edges_device(uint32_t nn, const uint32_t *d_src, const uint32_t *d_dst, uint64_t m) {
    if (m >= (1ull << 31)) throw Error(DCG_ERR_UNSUPPORTED, "edge count must be below 2^31 per load");
    if (nn >= 0x7FFFFFFFu) throw Error(DCG_ERR_UNSUPPORTED, "vertex count must be below 2^31 - 1");
    n = nn;
    num_edges = m;
    const uint64_t rows = 2ull * n;
    d_counters.alloc(C_NUM);
    fill<unsigned long long>(stream, d_counters.p, C_NUM, 0ull);
    DevBuf<uint32_t> deg;
    deg.alloc(rows ? rows : 1);
    fill<uint32_t>(stream, deg.p, rows, 0u);
    hipLaunchKernelGGL(k_check_ids, dim3(grid_for(m)), dim3(256), 0, stream, d_src, d_dst, m, n, d_counters.p, deg.p);
    DevBuf<uint32_t> dmax;
    dmax.alloc(1);
    fill<uint32_t>(stream, dmax.p, 1, 0u);
    hipLaunchKernelGGL(k_max_u32, dim3(grid_for(n)), dim3(256), 0, stream, deg.p, (uint64_t)n, dmax.p);
    DCG_HIP(hipGetLastError());
    unsigned long long bad = 0;
    DCG_HIP(hipMemcpyAsync(&bad, d_counters.p + C_BAD_ID, sizeof(bad), hipMemcpyDeviceToHost, stream));
    sync();
    if (bad) throw Error(DCG_ERR_ARG, "graph has " + std::to_string(bad) + " edges with an id >= num_nodes");
    DCG_HIP(hipMemcpyAsync(&max_out_deg, dmax.p, 4, hipMemcpyDeviceToHost, stream));
    sync();

    // Row pointers of the out- and in-CSR (64-bit exclusive scans of degrees).
    DevBuf<uint64_t> rp;
    rp.alloc(rows + 2);
    DevBuf<uint64_t> deg64;
    deg64.alloc(rows + 1);
    hipLaunchKernelGGL(k_widen, dim3(grid_for(rows)), dim3(256), 0, stream, deg.p, deg64.p, rows);
    // out rows: rp[0..n), in rows: rp[n+1 .. 2n+1) (each half scanned separately)
    size_t tb = scan_temp_bytes<uint64_t>(n);
    exclusive_sum<uint64_t>(stream, deg64.p, rp.p, n, temp(tb), tb);
    exclusive_sum<uint64_t>(stream, deg64.p + n, rp.p + n + 1, n, temp(tb), tb);

    // Stable sorts: out-CSR by src (payload dst), in-CSR by dst (payload src).
    DevBuf<uint32_t> ks, vs;
    ks.alloc(m ? m : 1);
    vs.alloc(m ? m : 1);
    const int kb = bits_for(n ? n - 1 : 0);

    // placement first (needs only degrees)
    DevBuf<uint64_t> pkeys, pkeys2, caps, offs;
    pkeys.alloc(rows ? rows : 1);
    pkeys2.alloc(rows ? rows : 1);
    caps.alloc(rows ? rows : 1);
    offs.alloc(rows ? rows : 1);
    hipLaunchKernelGGL(k_place_keys, dim3(grid_for(rows)), dim3(256), 0, stream, deg.p, rows, pkeys.p);
    // key = (63 - class) << 36 | row with rows < 2^32, ~0 for empty rows: bits [0, 42)
    tb = radix_sort_temp_bytes<uint64_t>(rows, false);
    radix_sort<uint64_t>(stream, pkeys.p, pkeys2.p, nullptr, nullptr, rows, 0, 42, temp(tb), tb);
    hipLaunchKernelGGL(k_place_caps, dim3(grid_for(rows)), dim3(256), 0, stream, pkeys2.p, rows, caps.p);
    tb = scan_temp_bytes<uint64_t>(rows);
    exclusive_sum<uint64_t>(stream, caps.p, offs.p, rows, temp(tb), tb);
    uint64_t last_off = 0, last_cap = 0;
    if (rows) {
        DCG_HIP(hipMemcpyAsync(&last_off, offs.p + rows - 1, 8, hipMemcpyDeviceToHost, stream));
        DCG_HIP(hipMemcpyAsync(&last_cap, caps.p + rows - 1, 8, hipMemcpyDeviceToHost, stream));
    }
    sync();
    used_slots = last_off + last_cap;
    pool_len = pool_taken = 0;
    double slack = cfg.store_slack < 0.05 ? 0.05 : cfg.store_slack;
    uint64_t want = used_slots + (uint64_t)((double)used_slots * slack) + 8 * kChunkEdges;
    total_slots = (want + kChunkEdges - 1) / kChunkEdges * kChunkEdges;

    d_out.alloc(n ? n : 1);
    d_in.alloc(n ? n : 1);
    hipLaunchKernelGGL(k_zero_desc, dim3(grid_for(n)), dim3(256), 0, stream, d_out.p, (uint64_t)n);
    hipLaunchKernelGGL(k_zero_desc, dim3(grid_for(n)), dim3(256), 0, stream, d_in.p, (uint64_t)n);
    hipLaunchKernelGGL(k_place_desc, dim3(grid_for(rows)), dim3(256), 0, stream, pkeys2.p, offs.p, deg.p, rows, n,
                       d_out.p, d_in.p);
    pkeys.release();
    pkeys2.release();
    caps.release();
    offs.release();

    DevBuf<uint32_t> slots;
    slots.alloc(total_slots);
    // out rows
    tb = radix_sort_temp_bytes<uint32_t>(m, true);
    void *t = temp(tb);
    radix_sort<uint32_t>(stream, d_src, ks.p, d_dst, vs.p, m, 0, kb, t, tb);
    hipLaunchKernelGGL(k_scatter_rows, dim3(grid_for(m)), dim3(256), 0, stream, ks.p, vs.p, m, rp.p, d_out.p, slots.p);
    d_sorted_out.alloc(n ? n : 1);
    d_sorted_in.alloc(n ? n : 1);
    hipLaunchKernelGGL(k_copy_u32, dim3(grid_for(n)), dim3(256), 0, stream, deg.p, d_sorted_out.p, (uint64_t)n);
    hipLaunchKernelGGL(k_copy_u32, dim3(grid_for(n)), dim3(256), 0, stream, deg.p + n, d_sorted_in.p, (uint64_t)n);
    hipLaunchKernelGGL(k_sorted_prefix, dim3(grid_for(m)), dim3(256), 0, stream, ks.p, vs.p, m, rp.p,
                       d_sorted_out.p);
    build_fences(ks.p, vs.p, m, rp.p, d_sorted_out.p, d_fbase_out, d_fence_out);
    // in rows
    radix_sort<uint32_t>(stream, d_dst, ks.p, d_src, vs.p, m, 0, kb, t, tb);
    hipLaunchKernelGGL(k_scatter_rows, dim3(grid_for(m)), dim3(256), 0, stream, ks.p, vs.p, m, rp.p + n + 1, d_in.p,
                       slots.p);
    hipLaunchKernelGGL(k_sorted_prefix, dim3(grid_for(m)), dim3(256), 0, stream, ks.p, vs.p, m, rp.p + n + 1,
                       d_sorted_in.p);
    build_fences(ks.p, vs.p, m, rp.p + n + 1, d_sorted_in.p, d_fbase_in, d_fence_in);
    DCG_HIP(hipGetLastError());
    ks.release();
    vs.release();
    rp.release();
    deg64.release();
    deg.release();

    d_tail.alloc(T_NUM);
    fill<uint64_t>(stream, d_tail.p, T_NUM, 0ull);
    DCG_HIP(hipMemcpyAsync(d_tail.p + T_TAIL, &used_slots, 8, hipMemcpyHostToDevice, stream));
    d_hot_out.alloc(n ? n : 1);
    d_hot_in.alloc(n ? n : 1);
    d_hot_window.alloc(n ? n : 1);
    fill<uint32_t>(stream, d_hot_out.p, n, 0u);
    fill<uint32_t>(stream, d_hot_in.p, n, 0u);
    fill<uint32_t>(stream, d_hot_window.p, n, 0u);
    d_hot_count.alloc(1);
    fill<unsigned long long>(stream, d_hot_count.p, 1, 0ull);
    // each row enters the list at most once per fold: capacity 2n
    if (cfg.placement != DCG_PLACE_DEVICE) d_hot_list.alloc(2ull * (n ? n : 1));
    else d_hot_list.release();

    cache.chunks = total_slots / kChunkEdges;
    if (cfg.placement == DCG_PLACE_DEVICE) {
        // every chunk is resident: the slot space itself is the frame array
        cache.frames = cache.chunks;
        std::swap(cache.d_frames.p, slots.p);
        std::swap(cache.d_frames.n, slots.n);
    } else {
        DCG_HIP(hipHostMalloc((void **)&h_slab, total_slots * sizeof(uint32_t), hipHostMallocDefault));
        DCG_HIP(hipMemcpyAsync(h_slab, slots.p, used_slots * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        sync();
        slots.release();
        const uint64_t frame_bytes = kChunkEdges * sizeof(uint32_t);
        cache.frames = cfg.cache_bytes / frame_bytes;
        if (cache.frames > cache.chunks) cache.frames = cache.chunks;
        if (cache.frames) cache.d_frames.alloc(cache.frames * kChunkEdges);
    }
    sync();
    cache_init();
    sync();
    d_temp.release();   // the load's sort scratch (~8 bytes per edge); batches regrow it to their size
    loaded = true;
}

// ------------------------------------------------------------- mutation kernels
// Keys of both directions in one array, sorted together.  A packed key holds
// the row (and, for deletes, the column) in just the bits the vertex count
// needs, with a direction bit above the row field, so out keys sort below in
// keys and one sort over [0, direction bit] orders each direction and keeps
// them apart in as few 8-bit passes as the graph allows.  The row field has a
// spare top bit: a record with an id out of range keys all ones there (the top
// of its direction) and counts as missing / invalid.  The split pass after the
// sort restores the layout the mutation kernels read (row << 32 | column for
// deletes, the row for appends) and ~0 for the out-of-range records.
//
// Delete keys (nb = bits of a vertex id): column in [0, nb), row in
// [nb, 2 nb + 1), direction bit 2 nb + 1.
__host__ __device__ __forceinline__ uint32_t del_dir_bit(uint32_t nb) { return 2 * nb + 1; }
// Append keys: row in [0, dir), direction bit dir = min(nb + 1, 31) (vertex
// ids stay below 2^31 - 1, so a 31-bit row field keeps its all-ones top).
__host__ __device__ __forceinline__ uint32_t add_dir_bit(uint32_t nb) { return nb + 1 < 31 ? nb + 1 : 31; }

__global__ void k_del_keys(const dcg_edge *recs, uint64_t count, uint32_t n, uint32_t nb, uint64_t *keys,
                           unsigned long long *counters) {
    const uint64_t dirbit = 1ull << del_dir_bit(nb);
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const dcg_edge e = recs[i];
        if (e.u >= n || e.v >= n) {
            keys[i] = dirbit - 1;
            keys[count + i] = (dirbit << 1) - 1;
            atomicAdd(&counters[C_MISSING], 1ull);
        } else {
            keys[i] = ((uint64_t)e.u << nb) | e.v;
            keys[count + i] = dirbit | ((uint64_t)e.v << nb) | e.u;
        }
    }
}

__global__ void k_split_del_keys(uint64_t *keys, uint64_t count, uint32_t nb) {
    const uint64_t dirbit = 1ull << del_dir_bit(nb), colmask = (1ull << nb) - 1;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < 2 * count;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t x = keys[i];
        const uint64_t top = i < count ? dirbit - 1 : (dirbit << 1) - 1;
        keys[i] = x == top ? ~0ull : (((x & (dirbit - 1)) >> nb) << 32) | (x & colmask);
    }
}

// Append pairs: row u with payload v (out), then row v with payload u (in).
__global__ void k_add_keys(const dcg_edge *recs, uint64_t count, uint32_t n, uint32_t nb, uint32_t *keys,
                           uint32_t *vals, unsigned long long *counters) {
    const uint32_t dirbit = 1u << add_dir_bit(nb);
    const uint32_t in_top = (uint32_t)(((uint64_t)dirbit << 1) - 1);
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const dcg_edge e = recs[i];
        const bool ok = e.u < n && e.v < n;
        keys[i] = ok ? e.u : dirbit - 1;
        vals[i] = e.v;
        keys[count + i] = ok ? (dirbit | e.v) : in_top;
        vals[count + i] = e.u;
        if (!ok) atomicAdd(&counters[C_INVALID], 1ull);
    }
}

__global__ void k_split_add_keys(uint32_t *keys, uint64_t count, uint32_t nb) {
    const uint32_t dirbit = 1u << add_dir_bit(nb);
    const uint32_t in_top = (uint32_t)(((uint64_t)dirbit << 1) - 1);
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < 2 * count;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t x = keys[i];
        const uint32_t top = i < count ? dirbit - 1 : in_top;
        keys[i] = x == top ? ~0u : (x & (dirbit - 1));
    }
}

// Group heads of a sorted key array (high 32 bits for 64-bit keys).
template <typename K>
__device__ __forceinline__ uint32_t row_of(K k);
template <>
__device__ __forceinline__ uint32_t row_of<uint64_t>(uint64_t k) { return (uint32_t)(k >> 32); }
template <>
__device__ __forceinline__ uint32_t row_of<uint32_t>(uint32_t k) { return k; }

template <typename K>
__global__ void k_head_flags(const K *keys, uint64_t count, uint8_t *flags) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t r = row_of<K>(keys[i]);
        flags[i] = r != ~0u && (i == 0 || row_of<K>(keys[i - 1]) != r) ? 1 : 0;
    }
}

// Small batches: one 1,024-thread workgroup walks the keys in tiles and
// compacts the heads in order (one launch instead of a flag pass and a
// device-wide selection).
constexpr uint64_t kSmallHeads = 8192;

template <typename K>
__global__ __launch_bounds__(1024) void k_group_heads_small(const K *keys, uint64_t count, uint32_t *gstart,
                                                            unsigned long long *ngroups) {
    __shared__ uint32_t wsum[16];
    const uint32_t w = threadIdx.x / kWave;
    uint32_t base = 0;
    for (uint64_t t0 = 0; t0 < count; t0 += 1024) {
        const uint64_t i = t0 + threadIdx.x;
        bool head = false;
        if (i < count) {
            const uint32_t r = row_of<K>(keys[i]);
            head = r != ~0u && (i == 0 || row_of<K>(keys[i - 1]) != r);
        }
        const uint64_t m = __ballot(head);
        if (lane_id() == 0) wsum[w] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t before = 0, tot = 0;
        for (uint32_t q = 0; q < 16; ++q) {
            before += q < w ? wsum[q] : 0u;
            tot += wsum[q];
        }
        if (head) gstart[base + before + lanes_below(m)] = (uint32_t)i;
        base += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) *ngroups = base;
}

// Heads of the runs of equal rows in a sorted key array (records of invalid
// edges, row ~0, sort last and form no group), compacted in key order: group g
// spans [gstart[g], gstart[g + 1]).
template <typename K>
static void group_heads(Engine &e, const K *keys, uint64_t count, uint32_t *gstart, unsigned long long *ngroups) {
    if (count <= kSmallHeads) {
        hipLaunchKernelGGL(k_group_heads_small<K>, dim3(1), dim3(1024), 0, e.stream, keys, count, gstart, ngroups);
        return;
    }
    e.d_flags.ensure(count);
    hipLaunchKernelGGL(k_head_flags<K>, dim3(grid_for(count)), dim3(256), 0, e.stream, keys, count, e.d_flags.p);
    const size_t tb = select_temp_bytes(count);
    select_flagged(e.stream, e.d_flags.p, count, gstart, ngroups, e.temp(tb), tb);
}

template <typename K>
__device__ __forceinline__ uint64_t group_end(const K *keys, uint64_t s, uint64_t count, uint32_t row) {
    uint64_t lo = s + 1, hi = count;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (row_of<K>(keys[mid]) <= row) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// End of group g: the next group's head, or for the last group the end of its
// run of equal rows.
template <typename K>
__device__ __forceinline__ uint64_t group_end_of(const K *keys, const uint32_t *gstart, uint64_t g, uint64_t ngroups,
                                                 uint64_t gs, uint64_t count, uint32_t row) {
    return g + 1 < ngroups ? (uint64_t)gstart[g + 1] : group_end<K>(keys, gs, count, row);
}

__device__ __forceinline__ uint32_t load_slot(const StoreView &s, uint64_t slot) { return *row_base(s, slot); }
// Whether a slot is read from pinned host memory (its chunk is not resident).
__device__ __forceinline__ bool host_slot(const StoreView &s, uint64_t slot) {
    return s.host && s.chunk_frame[(uint32_t)(slot >> kChunkLog2)] < 0;
}
// 64-byte lines of slots [from, to) of a row read from host memory.
__device__ __forceinline__ uint32_t host_lines(const StoreView &s, uint64_t off, uint32_t from, uint32_t to) {
    return to > from && host_slot(s, off + from) ? (to - from + 15) / 16 : 0u;
}
// Per-chunk line reads for the cache policy (null when there is no cache).
__device__ __forceinline__ void note_lines(float *reads, uint64_t off, uint32_t from, uint32_t to) {
    if (reads && to > from) atomicAdd(&reads[(off + from) >> kChunkLog2], (float)((to - from + 15) / 16));
}

constexpr uint32_t kBigDelete = 1024;   // rows above this use one 1024-thread workgroup

// Index of the first record of x's run among the group's sorted targets
// [gs, ge) (lower bound on the low halves).
__device__ __forceinline__ uint64_t target_lb(const uint64_t *keys, uint64_t gs, uint64_t ge, uint32_t x) {
    uint64_t lo = gs, hi = ge;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if ((uint32_t)keys[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint64_t target_ub(const uint64_t *keys, uint64_t from, uint64_t ge, uint32_t x) {
    uint64_t a = from, b = ge;
    while (a < b) {
        const uint64_t mid = (a + b) >> 1;
        if ((uint32_t)keys[mid] <= x) a = mid + 1;
        else b = mid;
    }
    return a;
}

struct DelResult {
    uint32_t end, removed, moved, removed_below;
};

// A row's fences: f[j] is the id at slot 16j of the sorted prefix, for j < cap.
struct Fences {
    uint32_t *f;
    uint32_t cap;
};

__device__ __forceinline__ Fences fences_of(const uint32_t *fbase, uint32_t *fence, uint32_t row) {
    const uint32_t b = fbase[row];
    return Fences{fence + b, fbase[row + 1] - b};
}

// An entry that lands on a fence position of the prefix refreshes that fence
// (compaction moves prefix entries towards the row start).
__device__ __forceinline__ void fence_put(const Fences &F, uint32_t pos, uint32_t x) {
    if (pos % kFenceStride == 0 && pos / kFenceStride < F.cap) F.f[pos / kFenceStride] = x;
}

// One wave over slots [from, to) of a row: remove the first still-requested
// occurrences of the group's targets, survivors compacted in order.  Matches
// are resolved in lane order through per-run counters, so the occurrences
// removed are exactly the first ones (source_local_chunk_store.h:685-739).
// Tombstones never match a target and stay in place.  `removed_below` counts
// removals at positions below `mark`.
__device__ DelResult wave_delete_range(const StoreView &s, uint64_t off, uint32_t from, uint32_t to, uint32_t mark,
                                       const uint64_t *keys, uint64_t gs, uint64_t ge, uint32_t *used,
                                       const Fences &F) {
    const uint32_t lane = lane_id();
    uint32_t write = from, removed = 0, moved = 0, below = 0;
    for (uint32_t t0 = from; t0 < to; t0 += kWave) {
        const uint32_t i = t0 + lane;
        const bool valid = i < to;
        const uint32_t x = valid ? load_slot(s, off + i) : 0u;
        const uint64_t lo = target_lb(keys, gs, ge, x);
        const bool match = valid && lo < ge && (uint32_t)keys[lo] == x;
        bool rem = false;
        uint64_t mm = __ballot(match);
        while (mm) {
            const int l = __ffsll((unsigned long long)mm) - 1;
            if ((int)lane == l) {
                const uint32_t requested = (uint32_t)(target_ub(keys, lo + 1, ge, x) - lo);
                const uint32_t old = atomicAdd(&used[lo], 1u);
                rem = old < requested;
            }
            mm &= mm - 1;
        }
        const bool keep = valid && !rem;
        const uint64_t km = __ballot(keep);
        const uint32_t rtile = (uint32_t)__popcll(__ballot(rem));
        below += (uint32_t)__popcll(__ballot(rem && i < mark));
        if (removed + rtile) {
            const uint32_t pos = write + lanes_below(km);
            const bool mv = keep && pos != i;
            if (mv) store_slot(s, off + pos, x);
            if (mv && i < mark) fence_put(F, pos, x);
            moved += (uint32_t)__popcll(__ballot(mv));
        }
        write += (uint32_t)__popcll(km);
        removed += rtile;
    }
    return DelResult{write, removed, moved, below};
}

// One wave: drop the tombstones of slots [0, to), order kept.  Returns the new
// end in .end and the live entries below `mark` in .removed_below.
__device__ DelResult wave_compact_tombs(const StoreView &s, uint64_t off, uint32_t to, uint32_t mark,
                                        const Fences &F) {
    const uint32_t lane = lane_id();
    uint32_t write = 0, moved = 0, live_below = 0;
    for (uint32_t t0 = 0; t0 < to; t0 += kWave) {
        const uint32_t i = t0 + lane;
        const uint32_t x = i < to ? load_slot(s, off + i) : kTomb;
        const bool keep = !(x & kTomb);
        const uint64_t km = __ballot(keep);
        live_below += (uint32_t)__popcll(__ballot(keep && i < mark));
        const uint32_t pos = write + lanes_below(km);
        const bool mv = keep && pos != i;
        if (mv) store_slot(s, off + pos, x);
        if (mv && i < mark) fence_put(F, pos, x);
        moved += (uint32_t)__popcll(__ballot(mv));
        write += (uint32_t)__popcll(km);
    }
    return DelResult{write, 0, moved, live_below};
}

// Fenced search of a sorted prefix [0, S) for the first `c` live occurrences
// of v, by a subgroup of W lanes (W = 64: the whole wave; W = 16: a quarter at
// bit `shift` of the wave's ballots); q is the lane's index in the subgroup and
// every lane of the subgroup takes part with the same v and c.
//  1. W-ary search over the fences (HBM): k = fences below v, so the lower bound
//     of v lies in slots (16(k-1), 16k].
//  2. 16 lanes read the 16-slot window at fence k-1 (one 64-byte line of the
//     row) and tombstone the first live occurrences of v in it (bit 31 set, id
//     bits kept, so the prefix stays ordered and its fences valid); while more
//     are wanted and the window's last slot does not pass v, the next window.
// Returns the occurrences taken; `probes` and `lines` (kept by subgroup lane 0)
// count the row slots read and the windows read from host memory.
template <uint32_t W>
__device__ uint32_t fenced_take(const StoreView &s, uint64_t off, uint32_t S, const Fences &F, uint32_t v,
                                uint32_t c, uint32_t q, uint32_t shift, uint32_t &probes, uint32_t &lines,
                                float *reads) {
    static_assert(W == 16 || W == 64, "subgroup of 16 or 64 lanes");
    const uint64_t below = (1ull << q) - 1;
    const uint32_t nfc = (S + kFenceStride - 1) / kFenceStride;
    uint32_t lo = 0, hi = F.cap < nfc ? F.cap : nfc;
    while (hi - lo > W) {
        const uint32_t step = (hi - lo + W - 1) / W;
        const uint32_t idx = lo + q * step;
        const bool p = idx < hi && (F.f[idx] & ~kTomb) < v;
        const uint64_t m = __ballot(p);
        const uint32_t cnt = (uint32_t)__popcll(W == 64 ? m : (m >> shift) & 0xFFFFull);
        if (!cnt) {
            hi = lo;
            break;
        }
        const uint32_t nhi = lo + cnt * step;
        lo += (cnt - 1) * step + 1;
        hi = nhi < hi ? nhi : hi;
    }
    if (hi > lo) {
        const uint32_t idx = lo + q;
        const bool p = idx < hi && (F.f[idx] & ~kTomb) < v;
        const uint64_t m = __ballot(p);
        lo += (uint32_t)__popcll(W == 64 ? m : (m >> shift) & 0xFFFFull);
    }
    uint32_t pos = lo ? (lo - 1) * kFenceStride : 0u, taken = 0;
    while (true) {
        const uint32_t i = pos + q;
        const bool valid = q < kFenceStride && i < S;
        const uint32_t x = valid ? load_slot(s, off + i) : 0u;
        const uint32_t id = x & ~kTomb;
        const bool live = valid && id == v && !(x & kTomb);
        uint64_t lm = __ballot(live);
        if (W != 64) lm = (lm >> shift) & 0xFFFFull;
        if (live && (uint32_t)__popcll(lm & below) < c - taken) store_slot(s, off + i, x | kTomb);
        const uint32_t nlive = (uint32_t)__popcll(lm);
        taken += nlive < c - taken ? nlive : c - taken;
        if (q == 0) {
            probes += S - pos < kFenceStride ? S - pos : kFenceStride;
            lines += host_slot(s, off + pos) ? 1u : 0u;
            note_lines(reads, off, pos, pos + 1);
        }
        // the run may go on past the window when its last slot is <= v
        uint64_t em = __ballot(valid && q == kFenceStride - 1 && id <= v);
        if (W != 64) em = (em >> shift) & 0xFFFFull;
        if (taken >= c || !em || pos + kFenceStride >= S) break;
        pos += kFenceStride;
    }
    return taken;
}

// Lower bound of v in the sorted prefix [0, S) for one thread: binary search
// over the fences (HBM), then over the at most 16 slots they leave open.
__device__ __forceinline__ uint32_t fenced_lb(const StoreView &s, uint64_t off, uint32_t S, const Fences &F,
                                              uint32_t v, uint32_t &probes) {
    const uint32_t nfc = (S + kFenceStride - 1) / kFenceStride;
    const uint32_t nf = F.cap < nfc ? F.cap : nfc;
    uint32_t lo = 0, hi = nf;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((F.f[mid] & ~kTomb) < v) lo = mid + 1;
        else hi = mid;
    }
    if (!lo) return 0;
    uint32_t a = (lo - 1) * kFenceStride + 1, b = lo < nf ? lo * kFenceStride : S;
    while (a < b) {
        const uint32_t mid = (a + b) >> 1;
        probes++;
        if ((load_slot(s, off + mid) & ~kTomb) < v) a = mid + 1;
        else b = mid;
    }
    return a;
}

// One lane per run of equal targets (the run's head): fenced_lb, then the live
// occurrences from there are tombstoned in order.  used[head] records what the
// run took; need_tail tells whether a run wants more.
__device__ __forceinline__ uint32_t search_prefix(const StoreView &s, uint64_t off, uint32_t S, const Fences &F,
                                                  const uint64_t *keys, uint64_t q, uint64_t gs, uint64_t ge,
                                                  uint32_t *used, uint32_t &probes, uint32_t &need_tail,
                                                  uint32_t &lines, float *reads) {
    const uint32_t v = (uint32_t)keys[q];
    if (q > gs && (uint32_t)keys[q - 1] == v) return 0;
    const uint32_t c = (uint32_t)(target_ub(keys, q + 1, ge, v) - q);
    lines += host_slot(s, off) ? 1u : 0u;   // the window the fences leave open
    note_lines(reads, off, 0, 1);
    uint32_t taken = 0;
    for (uint32_t j = fenced_lb(s, off, S, F, v, probes); j < S && taken < c; ++j) {
        const uint32_t x = load_slot(s, off + j);
        probes++;
        if ((x & ~kTomb) != v) break;
        if (!(x & kTomb)) {
            store_slot(s, off + j, x | kTomb);
            taken++;
        }
    }
    used[q] = taken;
    need_tail |= taken < c ? 1u : 0u;
    return taken;
}

constexpr uint32_t kQuarterRuns = 16;   // groups of up to this many targets search a run per quarter-wave

constexpr uint32_t kWaveSearch = kFenceMin;   // sorted prefixes at least this long are searched through fences
constexpr uint32_t kWaveTail = 4096;   // longest tail a wave scans for a large row
constexpr uint32_t kWaveRuns = 256;    // most targets a wave takes for a large row

// A row above kBigDelete slots still takes the wave path when the wave only
// binary-searches its sorted prefix and scans a short tail: few targets, and
// tombstones below the compaction point (k_del_big compacts).
__device__ __forceinline__ bool wave_searchable(const RowDesc &d, uint32_t S, uint64_t targets) {
    const uint32_t tombs = row_tombs(d);
    return S >= kWaveSearch && d.ext - S <= kWaveTail && targets <= kWaveRuns && (uint64_t)tombs * 4 <= d.ext &&
           tombs <= kMaxTombs / 2;
}

// Touched rows of at most kBigDelete slots, and larger rows that
// wave_searchable admits; the rest are listed for k_del_big.  `sorted` holds
// each row's sorted prefix length S.  A wave takes four groups at a time, one
// per quarter-wave:
//  * phase A (quarter): a group of at most kQuarterRuns targets on a prefix of
//    S >= kWaveSearch is searched through the row's fences (fenced_take, run
//    by run).  When every run was satisfied and the tombstones stay below a
//    quarter of the row, the quarter finishes the group (row descriptor, stats).
//    A single target in a row of at most 16 slots without such a prefix is
//    removed by the quarter from one read of the row.
//  * phase B (whole wave, one remaining group at a time): for a searched
//    group that still wants occurrences, wave_delete_range over the tail
//    [S, ext), which holds appended entries and never tombstones (the prefix
//    precedes the tail in the row's sequence, so exactly the first
//    occurrences go); a group of more targets is searched here (the whole wave
//    for one run of equal targets, search_prefix with a lane per run
//    otherwise); a prefix below kWaveSearch takes one wave_delete_range pass
//    over the row (S shrinks by the entries removed below it); and
//    tombstones above a quarter of the row take wave_compact_tombs.
// Compaction refreshes the fences of the prefix entries it moves.
__global__ __launch_bounds__(256) void k_del_rows(StoreView s, RowDesc *desc, uint32_t *sorted, uint32_t *hot,
                                                  const uint64_t *keys, uint64_t count,
                                                  const uint32_t *gstart, const unsigned long long *ngroups_p,
                                                  uint32_t *used, unsigned long long *counters, int count_stats,
                                                  uint32_t *big_groups, unsigned long long *nbig,
                                                  const uint32_t *fbase, uint32_t *fence, float *reads) {
    __shared__ WgSum<6> tot;
    wg_sum_init(tot);
    __syncthreads();
    const uint32_t lane = lane_id(), qd = lane >> 4, ql = lane & 15;
    const uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / kWave;
    const uint64_t nwaves = (uint64_t)gridDim.x * blockDim.x / kWave;
    const uint64_t ngroups = *ngroups_p;
    unsigned long long missing = 0, changed = 0, rd = 0, wr = 0, rows = 0, hl = 0;
    for (uint64_t base = wave * 4; base < ngroups; base += nwaves * 4) {
        // ---- phase A: one group per quarter-wave (values uniform in the quarter)
        const uint64_t g = base + qd;
        const bool valid = g < ngroups;
        uint64_t gs = 0, ge = 0;
        uint32_t row = 0, S = 0, Sraw = 0;
        RowDesc d{0, 0, 0};
        if (valid) {
            gs = gstart[g];
            row = (uint32_t)(keys[gs] >> 32);
            d = desc[row];
            ge = group_end_of<uint64_t>(keys, gstart, g, ngroups, gs, count, row);
            Sraw = sorted[row];
            S = Sraw < d.ext ? Sraw : d.ext;
        }
        const bool big = valid && d.ext > kBigDelete && !wave_searchable(d, S, ge - gs);
        if (big && ql == 0) big_groups[atomicAdd(nbig, 1ull)] = (uint32_t)gs;
        const bool qsearch = valid && !big && S >= kWaveSearch && ge - gs <= kQuarterRuns;
        uint32_t new_tombs = 0, probes = 0, need = 0, qlines = 0;
        if (qsearch) {
            const Fences F = fences_of(fbase, fence, row);
            for (uint64_t t = gs; t < ge; ++t) {
                const uint32_t v = (uint32_t)keys[t];
                if (t > gs && (uint32_t)keys[t - 1] == v) continue;
                const uint32_t c = (uint32_t)(target_ub(keys, t + 1, ge, v) - t);
                const uint32_t got = fenced_take<16>(s, d.off, S, F, v, c, ql, qd * 16, probes, qlines, reads);
                if (ql == 0) used[t] = got;
                new_tombs += got;
                need |= got < c ? 1u : 0u;
            }
        }
        // one target in a row of at most 16 slots without a searchable prefix:
        // the quarter reads the row (one line) and removes the first live
        // occurrence, the entries after it moving down one slot (order kept)
        const bool qtiny = valid && !big && !qsearch && ge - gs == 1 && d.ext <= kFenceStride && d.ext &&
                           (uint64_t)row_tombs(d) * 4 <= (uint64_t)d.ext - 1;
        uint32_t t_removed = 0, t_moved = 0, t_S2 = S;
        if (qtiny) {
            const uint32_t v = (uint32_t)keys[gs];
            const Fences F = fences_of(fbase, fence, row);
            const bool in = ql < d.ext;
            const uint32_t x = in ? load_slot(s, d.off + ql) : 0u;
            const uint64_t mm = (__ballot(in && x == v) >> (qd * 16)) & 0xFFFFull;
            if (mm) {
                const uint32_t first = (uint32_t)__ffsll((unsigned long long)mm) - 1;
                if (in && ql > first) {
                    store_slot(s, d.off + ql - 1, x);
                    if (ql < S) fence_put(F, ql - 1, x);
                }
                t_removed = 1;
                t_moved = d.ext - 1 - first;
                t_S2 = first < S ? S - 1 : S;
            }
            qlines += host_lines(s, d.off, 0, d.ext);
        }
        const uint32_t tombs_a = row_tombs(d) + new_tombs;
        const bool compact_a = tombs_a && (uint64_t)tombs_a * 4 > d.ext && d.ext <= kBigDelete;
        const bool done = (qsearch && !(need && S < d.ext) && !compact_a) || qtiny;
        if (ql == 0) hl += qlines;   // lines read in phase A, whoever finishes the group
        if (done && ql == 0) {
            if (qtiny) {
                if (t_removed) desc[row].ext = d.ext - 1;
                if (t_S2 != S || S != Sraw) sorted[row] = t_S2;
                mark_read(s, hot, row);   // the row was read whole (cache-policy signal)
                missing += 1 - t_removed;
                changed += t_removed;
                rd += d.ext;
                wr += t_moved;
            } else {
                if (new_tombs) desc[row].meta = make_meta(d.meta & 63u, tombs_a);
                missing += (ge - gs) - new_tombs;
                changed += new_tombs ? 1 : 0;
                rd += probes;
                wr += new_tombs;
            }
            rows++;
        }
        // ---- phase B: the whole wave finishes the remaining groups one by one
        uint64_t rest = __ballot(valid && !big && !done && ql == 0);
        while (rest) {
            const int src = __ffsll((unsigned long long)rest) - 1;
            rest &= rest - 1;
            const uint64_t bgs = __shfl(gs, src, kWave), bge = __shfl(ge, src, kWave);
            const uint32_t brow = __shfl(row, src, kWave), bS = __shfl(S, src, kWave);
            const RowDesc bd{__shfl(d.off, src, kWave), __shfl(d.ext, src, kWave), __shfl(d.meta, src, kWave)};
            const bool searched = __shfl(qsearch ? 1 : 0, src, kWave) != 0;
            uint32_t ext = bd.ext, S2 = bS, removed = 0, moved = 0, read = 0;
            uint32_t bt = searched ? __shfl(new_tombs, src, kWave) : 0u;
            uint32_t bneed = searched ? __shfl(need, src, kWave) : 0u;
            uint32_t bprobes = searched ? __shfl(probes, src, kWave) : 0u;
            bool full = false;
            const Fences F = fences_of(fbase, fence, brow);
            if (bS >= kWaveSearch) {
                if (!searched) {
                    uint32_t p = 0, nd = 0, nt = 0, ln = 0;
                    const uint32_t v0 = (uint32_t)keys[bgs];
                    if ((uint32_t)keys[bge - 1] == v0) {
                        // one run (the group's targets are sorted): the whole wave searches it
                        const uint32_t c = (uint32_t)(bge - bgs);
                        const uint32_t got = fenced_take<64>(s, bd.off, bS, F, v0, c, lane, 0, p, ln, reads);
                        if (lane == 0) {
                            used[bgs] = got;
                            nt = got;
                            nd = got < c ? 1u : 0u;
                        }
                    } else {
                        for (uint64_t q0 = bgs; q0 < bge; q0 += kWave)
                            if (q0 + lane < bge)
                                nt += search_prefix(s, bd.off, bS, F, keys, q0 + lane, bgs, bge, used, p, nd, ln,
                                                    reads);
                    }
                    for (int o = 32; o >= 1; o >>= 1) {
                        nt += __shfl_xor(nt, o, kWave);
                        p += __shfl_xor(p, o, kWave);
                    }
                    hl += ln;
                    bt = nt;
                    bprobes = p;
                    bneed = __ballot(nd != 0) ? 1u : 0u;
                }
                read += bprobes;
                removed += bt;
                if (bneed && bS < ext) {
                    __threadfence_block();   // the quarter's tombstone stores complete before the pass reads
                    const DelResult r = wave_delete_range(s, bd.off, bS, ext, 0, keys, bgs, bge, used, F);
                    if (lane == 0) {
                        hl += host_lines(s, bd.off, bS, ext);
                        note_lines(reads, bd.off, bS, ext);
                    }
                    read += ext - bS;
                    ext = r.end;
                    removed += r.removed;
                    moved += r.moved;
                }
            } else {
                const DelResult r = wave_delete_range(s, bd.off, 0, ext, bS, keys, bgs, bge, used, F);
                if (lane == 0) hl += host_lines(s, bd.off, 0, ext);
                read += ext;
                ext = r.end;
                removed = r.removed;
                moved = r.moved;
                S2 = bS - r.removed_below;
                full = true;
            }
            uint32_t tombs = row_tombs(bd) + bt;
            if (tombs && (uint64_t)tombs * 4 > ext && ext <= kBigDelete) {   // larger rows: k_del_big compacts
                __threadfence_block();   // the tombstone stores complete before the pass reads them
                const DelResult c = wave_compact_tombs(s, bd.off, ext, S2, F);
                if (lane == 0) hl += host_lines(s, bd.off, 0, ext);
                read += ext;
                ext = c.end;
                S2 = c.removed_below;
                moved += c.moved;
                tombs = 0;
                full = true;
            }
            if (lane == 0) {
                if (ext != bd.ext || tombs != row_tombs(bd)) {
                    desc[brow].ext = ext;
                    desc[brow].meta = make_meta(bd.meta & 63u, tombs);
                }
                if (S2 != bS || bS != sorted[brow]) sorted[brow] = S2;
                if (full) mark_read(s, hot, brow);   // a pass read the whole row (cache-policy signal)
                missing += (bge - bgs) - removed;
                changed += removed ? 1 : 0;
                rd += read;
                wr += moved + bt;
                rows++;
            }
        }
    }
    for (int o = 32; o >= 1; o >>= 1) {
        missing += __shfl_xor(missing, o, kWave);
        changed += __shfl_xor(changed, o, kWave);
        rd += __shfl_xor(rd, o, kWave);
        wr += __shfl_xor(wr, o, kWave);
        rows += __shfl_xor(rows, o, kWave);
        hl += __shfl_xor(hl, o, kWave);
    }
    unsigned long long *const dst[6] = {&counters[C_W_HOST_LINES], count_stats ? &counters[C_MISSING] : nullptr,
                                        count_stats ? &counters[C_CHANGED] : nullptr, &counters[C_W_READ],
                                        &counters[C_W_WRITTEN], &counters[C_W_ROWS]};
    const unsigned long long x[6] = {hl, missing, changed, rd, wr, rows};
    wg_sum_flush(tot, x, dst);
}

constexpr uint32_t kTileE = 4096;   // entries per workgroup tile of the large-row passes

struct BigLds {
    uint32_t runs[kTileE];   // run index of a matched entry, ~0u otherwise
    uint8_t gone[kTileE];
    uint32_t wsum[16];
    uint32_t nmatch;
    uint32_t tombs, probes, need_tail;
};

// Compact slots [from, to) of a row in place, in order, by one 1024-thread
// workgroup in tiles of 4,096 entries.  TOMBS: drop tombstones.  Otherwise:
// drop the first still-requested occurrences of the group's targets; each
// thread matches its 4 entries against the sorted targets and flags matches in
// LDS, then one wave resolves the tile's matches in row order through the
// per-run counters (exact first-occurrence rule).  Survivors are placed with a
// workgroup prefix sum; all reads of a tile precede its writes, and writes only
// move entries towards the row start.  Returns the new end; `kept_below` counts
// survivors at positions below `mark`.
template <bool TOMBS>
__device__ uint32_t compact_range(const StoreView &s, uint64_t off, uint32_t from, uint32_t to, uint32_t mark,
                                  const uint64_t *keys, uint64_t gs, uint64_t ge, uint32_t *used, BigLds &L,
                                  uint32_t &moved, uint32_t &kept_below, const Fences &F) {
    const uint32_t tid = threadIdx.x, lane = lane_id(), w = tid / kWave;
    uint32_t write = from;
    for (uint32_t t0 = from; t0 < to; t0 += kTileE) {
        if (tid == 0) L.nmatch = 0;
        __syncthreads();
        uint32_t xv[4];
        bool keep[4];
        for (int j = 0; j < 4; ++j) {
            const uint32_t e = j * 1024 + tid, i = t0 + e;
            xv[j] = i < to ? load_slot(s, off + i) : 0u;
            if (TOMBS) {
                keep[j] = i < to && !(xv[j] & kTomb);
            } else {
                uint32_t run = ~0u;
                if (i < to) {
                    const uint64_t lo = target_lb(keys, gs, ge, xv[j]);
                    if (lo < ge && (uint32_t)keys[lo] == xv[j]) {
                        run = (uint32_t)(lo - gs);
                        atomicAdd(&L.nmatch, 1u);
                    }
                }
                L.runs[e] = run;
                L.gone[e] = 0;
            }
        }
        if (!TOMBS) {
            __syncthreads();
            if (L.nmatch && w == 0) {
                for (uint32_t c0 = 0; c0 < kTileE; c0 += kWave) {
                    const uint32_t run = L.runs[c0 + lane];
                    uint64_t mm = __ballot(run != ~0u);
                    bool rem = false;
                    while (mm) {
                        const int l = __ffsll((unsigned long long)mm) - 1;
                        if ((int)lane == l) {
                            const uint64_t lo = gs + run;
                            const uint32_t x = (uint32_t)keys[lo];
                            const uint32_t requested = (uint32_t)(target_ub(keys, lo + 1, ge, x) - lo);
                            rem = atomicAdd(&used[lo], 1u) < requested;
                        }
                        mm &= mm - 1;
                    }
                    if (rem) L.gone[c0 + lane] = 1;
                }
            }
            __syncthreads();
            for (int j = 0; j < 4; ++j) keep[j] = t0 + j * 1024 + tid < to && !L.gone[j * 1024 + tid];
        }
        uint32_t pos[4];
        uint32_t base = write;
        for (int j = 0; j < 4; ++j) {
            const uint64_t km = __ballot(keep[j]);
            if (lane == 0) L.wsum[w] = (uint32_t)__popcll(km);
            __syncthreads();
            uint32_t before = 0, tot = 0;
            for (uint32_t q = 0; q < 16; ++q) {
                before += q < w ? L.wsum[q] : 0u;
                tot += L.wsum[q];
            }
            pos[j] = base + before + lanes_below(km);
            base += tot;
            __syncthreads();
        }
        for (int j = 0; j < 4; ++j) {
            const uint32_t i = t0 + j * 1024 + tid;
            if (keep[j] && i < mark) kept_below++;
            if (keep[j] && pos[j] != i) {
                store_slot(s, off + pos[j], xv[j]);
                if (i < mark) fence_put(F, pos[j], xv[j]);
                moved++;
            }
        }
        write = base;
        __syncthreads();
    }
    return write;
}

constexpr uint32_t kSearchPrefix = 1024;   // sorted prefixes at least this long are binary-searched
constexpr uint32_t kDelBigGrid = 512;      // two 1024-thread workgroups per CU

// Sum of a per-thread count over the 1024-thread workgroup (via LDS).
__device__ __forceinline__ uint32_t block_sum(uint32_t x, uint32_t &slot) {
    if (threadIdx.x == 0) slot = 0;
    __syncthreads();
    if (x) atomicAdd(&slot, x);
    __syncthreads();
    const uint32_t r = slot;
    __syncthreads();
    return r;
}

// Large rows (more than kBigDelete slots, or holding tombstones): one
// 1024-thread workgroup per row.
//  * Sorted prefix [0, S) of at least kSearchPrefix entries: the head of every
//    run of equal targets finds the run's value through the row's fences
//    (fenced_lb: a binary search of the fences in HBM, then of one 16-slot
//    window) and tombstones its first live occurrences (bit 31 set, id bits
//    kept, so the prefix stays ordered and its fences valid).
//    Then, only when some run still wants occurrences, compact_range over the
//    tail [S, ext) (appended entries, never tombstoned) continues the per-run
//    counters.  The prefix precedes the tail in the row's sequence, so exactly
//    the first occurrences go.
//  * Otherwise one compact_range pass over the whole row (tombstones never
//    match a target and stay); S shrinks by the entries removed below it.
//  * When tombstones pass a quarter of the row, a last pass compacts the whole
//    row; S becomes the number of live prefix entries.
__global__ __launch_bounds__(1024) void k_del_big(StoreView s, RowDesc *desc, uint32_t *sorted, uint32_t *hot,
                                                  const uint64_t *keys, uint64_t count, const uint32_t *big_groups,
                                                  const unsigned long long *nbig_p, uint32_t *used,
                                                  unsigned long long *counters, int count_stats,
                                                  const uint32_t *fbase, uint32_t *fence) {
    __shared__ BigLds L;
    const uint32_t tid = threadIdx.x;
    const uint64_t nbig = *nbig_p;
    for (uint64_t b = blockIdx.x; b < nbig; b += gridDim.x) {
        const uint64_t gs = big_groups[b];
        const uint32_t row = (uint32_t)(keys[gs] >> 32);
        const uint64_t ge = group_end<uint64_t>(keys, gs, count, row);
        const RowDesc d = desc[row];
        const uint32_t S = sorted[row] < d.ext ? sorted[row] : d.ext;
        const Fences F = fences_of(fbase, fence, row);
        uint32_t moved = 0, kept = 0, ext = d.ext, S2 = S, new_tombs = 0;
        uint64_t scanned = 0;
        bool full_read = false;
        if (S >= kSearchPrefix) {
            uint32_t tombs = 0, probes = 0, need_tail = 0;
            for (uint64_t q = gs + tid; q < ge; q += blockDim.x) {
                const uint32_t v = (uint32_t)keys[q];
                if (q > gs && (uint32_t)keys[q - 1] == v) continue;
                const uint32_t c = (uint32_t)(target_ub(keys, q + 1, ge, v) - q);
                uint32_t taken = 0;
                for (uint32_t j = fenced_lb(s, d.off, S, F, v, probes); j < S && taken < c; ++j) {
                    const uint32_t x = load_slot(s, d.off + j);
                    probes++;
                    if ((x & ~kTomb) != v) break;
                    if (!(x & kTomb)) {
                        store_slot(s, d.off + j, x | kTomb);
                        taken++;
                    }
                }
                used[q] = taken;
                tombs += taken;
                need_tail |= taken < c ? 1u : 0u;
            }
            new_tombs = block_sum(tombs, L.tombs);
            scanned += block_sum(probes, L.probes);
            if (block_sum(need_tail, L.need_tail) && S < d.ext) {
                ext = compact_range<false>(s, d.off, S, d.ext, 0, keys, gs, ge, used, L, moved, kept, F);
                scanned += d.ext - S;
            }
        } else {
            ext = compact_range<false>(s, d.off, 0, d.ext, S, keys, gs, ge, used, L, moved, kept, F);
            S2 = block_sum(kept, L.tombs);
            scanned += d.ext;
            full_read = true;
        }
        const uint32_t removed = new_tombs + (d.ext - ext);
        uint32_t tombs_left = row_tombs(d) + new_tombs;
        if (tombs_left && ((uint64_t)tombs_left * 4 > ext || tombs_left > kMaxTombs / 2)) {
            kept = 0;
            const uint32_t ext2 = compact_range<true>(s, d.off, 0, ext, S2, keys, gs, ge, used, L, moved, kept, F);
            S2 = block_sum(kept, L.tombs);
            scanned += ext;
            ext = ext2;
            tombs_left = 0;
            full_read = true;
        }
        if (tid == 0) {
            desc[row].ext = ext;
            desc[row].meta = make_meta(d.meta & 63u, tombs_left);
            sorted[row] = S2;
            if (count_stats) {
                atomicAdd(&counters[C_MISSING], (unsigned long long)((ge - gs) - removed));
                if (removed) atomicAdd(&counters[C_CHANGED], 1ull);
            }
            atomicAdd(&counters[C_SLOTS_READ], scanned);
            atomicAdd(&counters[C_B_ROWS], 1ull);
            if (new_tombs) atomicAdd(&counters[C_SLOTS_WRITTEN], (unsigned long long)new_tombs);
            if (full_read) mark_read(s, hot, row);   // a pass read the whole row (cache-policy signal)
        }
        if (moved) atomicAdd(&counters[C_SLOTS_WRITTEN], (unsigned long long)moved);
        __syncthreads();
    }
}

// Relocation demand per capacity class: a workgroup histogram in LDS, then one
// global atomic per class per workgroup.
__global__ __launch_bounds__(256) void k_add_plan(const RowDesc *desc, const uint32_t *keys, uint64_t count,
                                                  const uint32_t *gstart, const unsigned long long *ngroups_p,
                                                  uint64_t *tail) {
    __shared__ uint32_t hist[kClasses];
    for (uint32_t i = threadIdx.x; i < kClasses; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    const uint64_t ngroups = *ngroups_p;
    for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < ngroups;
         g += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t gs = gstart[g];
        const uint32_t row = keys[gs];
        const uint64_t ge = group_end_of<uint32_t>(keys, gstart, g, ngroups, gs, count, row);
        const RowDesc d = desc[row];
        const uint64_t nd = (uint64_t)d.ext + (ge - gs);
        if (nd > row_cap(d)) {
            uint64_t c = nd < kMinRowCap ? kMinRowCap : nd;
            const uint32_t cls = 64u - (uint32_t)__clzll((long long)(c - 1));
            atomicAdd(&hist[cls], 1u);
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kClasses; i += blockDim.x)
        if (hist[i]) atomicAdd((unsigned long long *)&tail[T_COUNT + i], (unsigned long long)hist[i]);
}

// Per class: blocks retired by earlier batches are taken first, the rest come
// from the bump tail (largest classes first, each aligned to its size).
__device__ void alloc_classes(uint64_t *tail, uint64_t total) {
    uint64_t t = tail[T_TAIL];
    for (int c = kClasses - 1; c >= 0; --c) {
        const uint64_t cnt = tail[T_COUNT + c];
        const uint64_t take = cnt < tail[T_FREE + c] ? cnt : tail[T_FREE + c];
        tail[T_TAKE + c] = take;
        tail[T_RANK + c] = 0;
        tail[T_COUNT + c] = 0;
        if (cnt == take) continue;
        const uint64_t a = 1ull << c;
        t = (t + a - 1) & ~(a - 1);
        tail[T_CURSOR + c] = t;
        t += (cnt - take) << c;
    }
    if (t > total) tail[T_OVERFLOW] = 1;
    else tail[T_TAIL] = t;
}

__global__ void k_alloc_classes(uint64_t *tail, uint64_t total) {
    if (threadIdx.x || blockIdx.x) return;
    alloc_classes(tail, total);
}

// Free-pool segments of the class-sorted pool keys (taken entries, ~0, sort last).
__global__ void k_pool_segments(const uint64_t *pool, uint64_t len, uint64_t *tail) {
    // pool[0, len) sorted by class: each class segment's first entry records
    // its start, its last entry its end (k_pool_counts turns ends into counts)
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < len; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t c = (uint32_t)(pool[i] >> 40);
        if (i == 0 || (uint32_t)(pool[i - 1] >> 40) != c) tail[T_PSTART + c] = i;
        if (i + 1 == len || (uint32_t)(pool[i + 1] >> 40) != c) tail[T_FREE + c] = i + 1;
    }
}

__global__ void k_pool_counts(uint64_t *tail) {
    const uint32_t c = threadIdx.x;
    if (c < kClasses && tail[T_FREE + c]) tail[T_FREE + c] -= tail[T_PSTART + c];
}

// New blocks of the rows that relocate: one pass per 256 groups reserves ranks
// in each class with one global atomic per class (LDS ranks inside the
// workgroup), so relocations do not serialise on a cursor.  Rank i of class c
// takes the i-th free-pool block of c, or a bump block past them.  Same
// relocation test as k_add_plan.  Both directions in one launch (blockIdx.y);
// their ranks share the class counters.
struct AppendSide {
    const RowDesc *desc;
    const uint32_t *keys;
    const uint32_t *gstart;
    const unsigned long long *ngroups;
    uint64_t *reloc;
};

__global__ __launch_bounds__(256) void k_add_assign(AppendSide out, AppendSide in, uint64_t count, uint64_t *tail,
                                                    uint64_t *pool) {
    __shared__ uint32_t hist[kClasses];
    __shared__ unsigned long long base[kClasses];
    if (tail[T_OVERFLOW]) return;
    const AppendSide &sd = blockIdx.y ? in : out;
    const RowDesc *desc = sd.desc;
    const uint32_t *keys = sd.keys, *gstart = sd.gstart;
    uint64_t *reloc = sd.reloc;
    const uint64_t ngroups = *sd.ngroups;
    for (uint64_t g0 = (uint64_t)blockIdx.x * blockDim.x; g0 < ngroups; g0 += (uint64_t)gridDim.x * blockDim.x) {
        for (uint32_t i = threadIdx.x; i < kClasses; i += blockDim.x) hist[i] = 0;
        __syncthreads();
        const uint64_t g = g0 + threadIdx.x;
        uint32_t cls = 0, rank = 0;
        if (g < ngroups) {
            const uint64_t gs = gstart[g];
            const uint32_t row = keys[gs];
            const uint64_t ge = group_end_of<uint32_t>(keys, gstart, g, ngroups, gs, count, row);
            const RowDesc d = desc[row];
            const uint64_t nd = (uint64_t)d.ext + (ge - gs);
            if (nd > row_cap(d)) {
                uint64_t c = nd < kMinRowCap ? kMinRowCap : nd;
                cls = 64u - (uint32_t)__clzll((long long)(c - 1));
                rank = atomicAdd(&hist[cls], 1u);
            }
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < kClasses; i += blockDim.x)
            base[i] = hist[i] ? atomicAdd((unsigned long long *)&tail[T_RANK + i], (unsigned long long)hist[i]) : 0ull;
        __syncthreads();
        if (cls) {
            const uint64_t i = base[cls] + rank, take = tail[T_TAKE + cls];
            if (i < take) {
                const uint64_t slot = tail[T_PSTART + cls] + i;
                reloc[g] = pool[slot] & kPoolOff;
                pool[slot] = ~0ull;   // taken
            } else {
                reloc[g] = tail[T_CURSOR + cls] + ((i - take) << cls);
            }
        }
        __syncthreads();
    }
}

// A relocated row's old block joins the free pool (appended past the pool's
// live entries); the next batch may reuse it (source_local_chunk_store.h:410-433
// retires a block until the next batch in the same way).
__device__ __forceinline__ void retire_block(uint64_t *tail, uint64_t *pool, const RowDesc &old) {
    const uint64_t i = atomicAdd((unsigned long long *)&tail[T_RETIRED], 1ull);
    pool[i] = ((uint64_t)(old.meta & 63u) << 40) | old.off;
}

constexpr uint32_t kBigAppend = 4096;   // relocations of larger rows use one 1024-thread workgroup

__device__ __forceinline__ uint32_t class_of(uint64_t nd) {
    const uint64_t c = nd < kMinRowCap ? kMinRowCap : nd;
    return 64u - (uint32_t)__clzll((long long)(c - 1));
}

// ------------------------------------------------- small batches, one launch
// Batches of up to kSmallSort records (and free pools of up to kSmallSort
// blocks): one 1,024-thread workgroup per direction sorts the keys in LDS with
// a bitonic network and writes them with their group heads (and, for appends,
// the relocation demand per class), in place of the key pass, the device radix
// sorts, the head compaction and the plan pass.  Sorted keys, payload order
// (stable, through the record index in the low word) and heads are the same
// as those of the device-wide path.
constexpr uint32_t kSmallSort = 4096;
constexpr uint64_t kMaxBatch = 1ull << 30;   // both directions' keys sort as one array of int-counted items

__device__ __forceinline__ void cmp_swap(uint64_t &a, uint64_t &b, bool up) {
    if ((a > b) == up) {
        const uint64_t t = a;
        a = b;
        b = t;
    }
}

// Merge steps j and j/2 of one bitonic stage go together: the four entries
// o, o + j/2, o + j, o + 3j/2 of a 2j-block are closed under both, so a thread
// exchanges them in registers and the stage needs half the barriers.
__device__ void lds_bitonic(uint64_t *s, uint32_t N) {   // N a power of two, s[0, N) ascending
    for (uint32_t k = 2; k <= N; k <<= 1)
        for (uint32_t j = k >> 1; j > 0;) {
            if (j >= 2) {
                const uint32_t h = j >> 1;
                for (uint32_t i = threadIdx.x; i < N / 4; i += blockDim.x) {
                    const uint32_t b = 4 * i - 3 * (i & (h - 1));
                    const bool up = (b & k) == 0;
                    uint64_t x0 = s[b], x1 = s[b + h], x2 = s[b + j], x3 = s[b + j + h];
                    cmp_swap(x0, x2, up);
                    cmp_swap(x1, x3, up);
                    cmp_swap(x0, x1, up);
                    cmp_swap(x2, x3, up);
                    s[b] = x0;
                    s[b + h] = x1;
                    s[b + j] = x2;
                    s[b + j + h] = x3;
                }
                j >>= 2;
            } else {
                for (uint32_t i = threadIdx.x; i < N / 2; i += blockDim.x) {
                    uint64_t x0 = s[2 * i], x1 = s[2 * i + 1];
                    cmp_swap(x0, x1, ((2 * i) & k) == 0);
                    s[2 * i] = x0;
                    s[2 * i + 1] = x1;
                }
                j = 0;
            }
            __syncthreads();
        }
}

// Heads of the runs of equal rows (high word) of s[0, count), row ~0 forming
// no group, compacted in order into gstart (and gs when given); every thread
// returns the group count.
__device__ uint32_t lds_heads(const uint64_t *s, uint32_t count, uint32_t *gstart, uint32_t *gs, uint32_t *wsum) {
    const uint32_t w = threadIdx.x / kWave;
    uint32_t base = 0;
    for (uint32_t t0 = 0; t0 < count; t0 += 1024) {
        const uint32_t i = t0 + threadIdx.x;
        bool head = false;
        if (i < count) {
            const uint32_t r = (uint32_t)(s[i] >> 32);
            head = r != ~0u && (i == 0 || (uint32_t)(s[i - 1] >> 32) != r);
        }
        const uint64_t m = __ballot(head);
        if (lane_id() == 0) wsum[w] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t before = 0, tot = 0;
        for (uint32_t q = 0; q < 16; ++q) {
            before += q < w ? wsum[q] : 0u;
            tot += wsum[q];
        }
        if (head) {
            const uint32_t g = base + before + lanes_below(m);
            gstart[g] = i;
            if (gs) gs[g] = i;
        }
        base += tot;
        __syncthreads();
    }
    return base;
}

__device__ __forceinline__ uint32_t pow2_at_least(uint32_t x) {
    uint32_t N = 1;
    while (N < x) N <<= 1;
    return N;
}

// Delete keys (u << 32 | v) for the out direction (block 0) and (v << 32 | u)
// for the in direction (block 1); records with an id out of range key ~0 and
// count as missing.  Also clears the mutation counters and both directions'
// occurrence counters (used[0, 2 count)).
__global__ __launch_bounds__(1024) void k_del_prep_small(const dcg_edge *recs, uint32_t count, uint32_t n,
                                                         uint64_t *kout, uint64_t *kin, uint32_t *gout,
                                                         uint32_t *gin, uint32_t *used,
                                                         unsigned long long *counters) {
    __shared__ uint64_t s[kSmallSort];
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t bad;
    const bool out = blockIdx.x == 0;
    if (threadIdx.x == 0) bad = 0;
    if (out)
        for (uint32_t j = threadIdx.x; j < C_NUM; j += blockDim.x)
            if (j != C_MISSING && j != C_GROUPS_OUT && j != C_GROUPS_IN) counters[j] = 0;
    for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) used[blockIdx.x * count + i] = 0;
    const uint32_t N = pow2_at_least(count);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < N; i += blockDim.x) {
        uint64_t k = ~0ull;
        if (i < count) {
            const dcg_edge e = recs[i];
            if (e.u >= n || e.v >= n) {
                if (out) atomicAdd(&bad, 1u);
            } else {
                k = out ? ((uint64_t)e.u << 32) | e.v : ((uint64_t)e.v << 32) | e.u;
            }
        }
        s[i] = k;
    }
    __syncthreads();
    lds_bitonic(s, N);
    uint64_t *dst = out ? kout : kin;
    for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) dst[i] = s[i];
    const uint32_t groups = lds_heads(s, count, out ? gout : gin, nullptr, wsum);
    if (threadIdx.x == 0) {
        counters[out ? C_GROUPS_OUT : C_GROUPS_IN] = groups;
        if (out) counters[C_MISSING] = bad;
    }
}

// Append keys: row u with payload v (block 0, out direction), row v with
// payload u (block 1); invalid records key ~0 and sort last in record order.
// Sorting (row << 32 | record index) gives the stable order of the pairs.
// Then the relocation demand per capacity class (the k_add_plan test) from
// the heads kept in LDS; block 0 also clears the mutation counters.
__global__ __launch_bounds__(1024) void k_add_prep_small(const dcg_edge *recs, uint32_t count, uint32_t n,
                                                         uint32_t *kout, uint32_t *vout, uint32_t *kin,
                                                         uint32_t *vin, uint32_t *gout, uint32_t *gin,
                                                         const RowDesc *dout, const RowDesc *din, uint64_t *tail,
                                                         unsigned long long *counters) {
    __shared__ uint64_t s[kSmallSort];
    __shared__ uint32_t gs[kSmallSort];
    __shared__ uint32_t wsum[16], hist[kClasses];
    __shared__ uint32_t bad;
    const bool out = blockIdx.x == 0;
    if (threadIdx.x == 0) bad = 0;
    for (uint32_t c = threadIdx.x; c < kClasses; c += blockDim.x) hist[c] = 0;
    if (out)
        for (uint32_t j = threadIdx.x; j < C_NUM; j += blockDim.x)
            if (j != C_INVALID && j != C_GROUPS_OUT && j != C_GROUPS_IN) counters[j] = 0;
    const uint32_t N = pow2_at_least(count);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < N; i += blockDim.x) {
        uint64_t k = ~0ull;
        if (i < count) {
            const dcg_edge e = recs[i];
            const bool ok = e.u < n && e.v < n;
            if (!ok) atomicAdd(&bad, 1u);
            k = ((uint64_t)(ok ? (out ? e.u : e.v) : ~0u) << 32) | i;
        }
        s[i] = k;
    }
    __syncthreads();
    lds_bitonic(s, N);
    uint32_t *kd = out ? kout : kin, *vd = out ? vout : vin;
    for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) {
        const uint64_t x = s[i];
        const dcg_edge e = recs[(uint32_t)x];
        kd[i] = (uint32_t)(x >> 32);
        vd[i] = out ? e.v : e.u;
    }
    const uint32_t groups = lds_heads(s, count, out ? gout : gin, gs, wsum);
    const uint32_t valid = count - bad;   // invalid records sort last
    const RowDesc *desc = out ? dout : din;
    for (uint32_t g = threadIdx.x; g < groups; g += blockDim.x) {
        const uint32_t a = gs[g], b = g + 1 < groups ? gs[g + 1] : valid;
        const RowDesc d = desc[(uint32_t)(s[a] >> 32)];
        const uint64_t nd = (uint64_t)d.ext + (b - a);
        if (nd > row_cap(d)) atomicAdd(&hist[class_of(nd)], 1u);
    }
    __syncthreads();
    for (uint32_t c = threadIdx.x; c < kClasses; c += blockDim.x)
        if (hist[c]) atomicAdd((unsigned long long *)&tail[T_COUNT + c], (unsigned long long)hist[c]);
    if (threadIdx.x == 0) {
        counters[out ? C_GROUPS_OUT : C_GROUPS_IN] = groups;
        if (out) counters[C_INVALID] = bad;
    }
}

// Free pool of up to kSmallSort entries: sort by class in LDS, keep the live
// entries in place, record each class's segment, set the retired cursor past
// them, then assign blocks per class (k_alloc_classes).  One workgroup.
__global__ __launch_bounds__(1024) void k_pool_small(uint64_t *pool, uint32_t len, uint32_t live, uint64_t *tail,
                                                     uint64_t total) {
    __shared__ uint64_t s[kSmallSort];
    __shared__ uint32_t pst[kClasses], pen[kClasses];
    const uint32_t N = pow2_at_least(len);
    for (uint32_t c = threadIdx.x; c < kClasses; c += blockDim.x) pst[c] = pen[c] = 0;
    for (uint32_t i = threadIdx.x; i < N; i += blockDim.x) s[i] = i < len ? pool[i] : ~0ull;
    __syncthreads();
    lds_bitonic(s, N);
    for (uint32_t i = threadIdx.x; i < live; i += blockDim.x) {
        const uint64_t x = s[i];
        const uint32_t c = (uint32_t)(x >> 40);
        if (i == 0 || (uint32_t)(s[i - 1] >> 40) != c) pst[c] = i;
        if (i + 1 == live || (uint32_t)(s[i + 1] >> 40) != c) pen[c] = i + 1;
        pool[i] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t c = 0; c < (uint32_t)kClasses; ++c) {
            tail[T_PSTART + c] = pst[c];
            tail[T_FREE + c] = pen[c] ? pen[c] - pst[c] : 0;
        }
        tail[T_RETIRED] = live;
        alloc_classes(tail, total);
    }
}

// Append (and relocation) per touched row.  A wave takes four groups at a
// time, one per quarter-wave: a quarter appends up to kQuarterAppend entries
// and copies a relocated row of up to kQuarterCopy slots itself; the whole
// wave then takes the remaining groups one at a time (larger appends and
// copies).  Rows above kBigAppend slots that must move are listed for
// k_add_big.  Relocated entries (tombstones included) keep their positions,
// so the row's sorted prefix length and fences stay valid.
constexpr uint32_t kQuarterAppend = 64, kQuarterCopy = 256;

__global__ __launch_bounds__(256) void k_add_rows(StoreView s, RowDesc *desc, uint32_t *hot, const uint32_t *keys,
                                                  const uint32_t *vals, uint64_t count, const uint32_t *gstart,
                                                  const unsigned long long *ngroups_p, uint64_t *tail,
                                                  uint64_t *pool, const uint64_t *reloc, unsigned long long *counters,
                                                  int count_stats, uint32_t *big_groups, unsigned long long *nbig) {
    __shared__ WgSum<3> tot;
    wg_sum_init(tot);
    __syncthreads();
    const uint32_t lane = lane_id(), qd = lane >> 4, ql = lane & 15;
    const uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / kWave;
    const uint64_t nwaves = (uint64_t)gridDim.x * blockDim.x / kWave;
    const uint64_t ngroups = *ngroups_p;
    const bool overflow = tail[T_OVERFLOW] != 0;
    unsigned long long relocated = 0, rd = 0, wr = 0;
    // one group's append by `width` lanes starting at lane index `li`
    auto append = [&](uint64_t g, uint64_t gs, uint32_t row, uint32_t k, RowDesc d, uint32_t li, uint32_t width,
                      bool leader) {
        const uint64_t nd = (uint64_t)d.ext + k;
        if (nd > row_cap(d)) {
            const uint32_t cls = class_of(nd);
            const uint64_t noff = reloc[g];
            for (uint32_t i = li; i < d.ext; i += width) store_slot(s, noff + i, load_slot(s, d.off + i));
            if (leader) {
                if (d.ext) mark_read(s, hot, row);
                if (d.meta) retire_block(tail, pool, d);
                relocated++;
                rd += d.ext;
                wr += d.ext;
            }
            d.off = noff;
            d.meta = make_meta(cls, row_tombs(d));
        }
        for (uint32_t i = li; i < k; i += width) store_slot(s, d.off + d.ext + i, vals[gs + i]);
        d.ext = (uint32_t)nd;
        if (leader) {
            wr += k;
            desc[row] = d;
        }
    };
    for (uint64_t base = wave * 4; base < ngroups; base += nwaves * 4) {
        const uint64_t g = base + qd;
        const bool valid = g < ngroups;
        uint64_t gs = 0;
        uint32_t row = 0, k = 0;
        RowDesc d{0, 0, 0};
        bool move = false;
        if (valid) {
            gs = gstart[g];
            row = keys[gs];
            k = (uint32_t)(group_end_of<uint32_t>(keys, gstart, g, ngroups, gs, count, row) - gs);
            d = desc[row];
            move = (uint64_t)d.ext + k > row_cap(d);
        }
        const bool skip = valid && move && overflow;   // no room: the host reports overflow
        const bool big = valid && move && !overflow && d.ext > kBigAppend;
        if (big && ql == 0) big_groups[atomicAdd(nbig, 1ull)] = (uint32_t)g;
        const bool quarter = valid && !skip && !big && k <= kQuarterAppend && (!move || d.ext <= kQuarterCopy);
        if (quarter) append(g, gs, row, k, d, ql, 16, ql == 0);
        uint64_t rest = __ballot(valid && !skip && !big && !quarter && ql == 0);
        while (rest) {
            const int src = __ffsll((unsigned long long)rest) - 1;
            rest &= rest - 1;
            const RowDesc bd{__shfl(d.off, src, kWave), __shfl(d.ext, src, kWave), __shfl(d.meta, src, kWave)};
            append(base + (uint32_t)(src >> 4), __shfl(gs, src, kWave), __shfl(row, src, kWave),
                   __shfl(k, src, kWave), bd, lane, kWave, lane == 0);
        }
    }
    for (int o = 32; o >= 1; o >>= 1) {
        relocated += __shfl_xor(relocated, o, kWave);
        rd += __shfl_xor(rd, o, kWave);
        wr += __shfl_xor(wr, o, kWave);
    }
    unsigned long long *const dst[3] = {count_stats ? &counters[C_RELOCATED] : nullptr, &counters[C_SLOTS_READ],
                                        &counters[C_SLOTS_WRITTEN]};
    const unsigned long long x[3] = {relocated, rd, wr};
    wg_sum_flush(tot, x, dst);
}

// Relocation of a large row: one 1024-thread workgroup copies the row into its
// new block (16-byte accesses where the row is aligned) and appends the batch.
__global__ __launch_bounds__(1024) void k_add_big(StoreView s, RowDesc *desc, uint32_t *hot, const uint32_t *keys,
                                                  const uint32_t *vals, uint64_t count, const uint32_t *gstart,
                                                  const uint32_t *big_groups, const unsigned long long *nbig_p,
                                                  const uint64_t *reloc, uint64_t *tail, uint64_t *pool,
                                                  unsigned long long *counters, int count_stats) {
    const uint64_t nbig = *nbig_p;
    for (uint64_t b = blockIdx.x; b < nbig; b += gridDim.x) {
        const uint64_t g = big_groups[b];
        const uint64_t gs = gstart[g];
        const uint32_t row = keys[gs];
        const uint64_t ge = group_end<uint32_t>(keys, gs, count, row);
        const uint32_t k = (uint32_t)(ge - gs);
        RowDesc d = desc[row];
        const uint64_t nd = (uint64_t)d.ext + k;
        const uint32_t cls = class_of(nd);
        const uint64_t noff = reloc[g];
        // blocks above 65,536 slots are whole chunks; copy chunk segment by segment
        for (uint64_t seg = 0; seg < d.ext; seg += kChunkEdges) {
            const uint32_t len = (uint32_t)min((uint64_t)kChunkEdges, (uint64_t)d.ext - seg);
            const uint32_t *src = row_base(s, d.off + seg);
            uint32_t *dst = frame_w(s, noff + seg);
            uint32_t *hdst = s.host ? s.host + noff + seg : nullptr;
            for (uint32_t i = threadIdx.x; i < len; i += blockDim.x) {
                const uint32_t x = src[i];
                if (dst) dst[i] = x;
                if (hdst) hdst[i] = x;
            }
        }
        for (uint32_t i = threadIdx.x; i < k; i += blockDim.x) store_slot(s, noff + d.ext + i, vals[gs + i]);
        __syncthreads();
        if (threadIdx.x == 0) {
            desc[row] = RowDesc{noff, (uint32_t)nd, make_meta(cls, row_tombs(d))};
            retire_block(tail, pool, d);
            mark_read(s, hot, row);
            if (count_stats) atomicAdd(&counters[C_RELOCATED], 1ull);
            atomicAdd(&counters[C_SLOTS_READ], (unsigned long long)d.ext);
            atomicAdd(&counters[C_SLOTS_WRITTEN], (unsigned long long)nd);
        }
        __syncthreads();
    }
}

static uint32_t wave_grid(uint64_t groups) {
    uint64_t waves = groups ? groups : 1;
    uint64_t blocks = (waves * kWave + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    return (uint32_t)blocks;
}

void Engine::mutate_delete(const dcg_edge *recs, uint64_t count, dcg_phase_stats *st) {
    if (!count) {
        fill<unsigned long long>(stream, d_counters.p, C_NUM, 0ull);
        return;
    }
    if (count >= kMaxBatch) throw Error(DCG_ERR_ARG, "update batches hold fewer than 2^30 records");
    d_k64a.ensure(2 * count);   // sorted keys, out then in direction
    d_gstart.ensure(count);
    d_gstart2.ensure(count);
    d_used.ensure(2 * count);   // occurrence counters, out then in direction
    uint64_t *kout = d_k64a.p, *kin = d_k64a.p + count;
    uint32_t *used_out = d_used.p, *used_in = d_used.p + count;
    if (count <= kSmallSort && !large_batches) {
        hipLaunchKernelGGL(k_del_prep_small, dim3(2), dim3(1024), 0, stream, recs, (uint32_t)count, n, kout, kin,
                           d_gstart.p, d_gstart2.p, d_used.p, d_counters.p);
    } else {
        fill<unsigned long long>(stream, d_counters.p, C_NUM, 0ull);
        d_k64c.ensure(2 * count);
        // both directions' keys unsorted in d_k64c, one sort into kout | kin
        const uint32_t nb = (uint32_t)bits_for(n ? n - 1 : 0);
        hipLaunchKernelGGL(k_del_keys, dim3(grid_for(count)), dim3(256), 0, stream, recs, count, n, nb, d_k64c.p,
                           d_counters.p);
        const size_t tb = radix_sort_temp_bytes<uint64_t>(2 * count, false);
        radix_sort<uint64_t>(stream, d_k64c.p, kout, nullptr, nullptr, 2 * count, 0, (int)del_dir_bit(nb) + 1,
                             temp(tb), tb);
        hipLaunchKernelGGL(k_split_del_keys, dim3(grid_for(2 * count)), dim3(256), 0, stream, kout, count, nb);
        group_heads<uint64_t>(*this, kout, count, d_gstart.p, d_counters.p + C_GROUPS_OUT);
        group_heads<uint64_t>(*this, kin, count, d_gstart2.p, d_counters.p + C_GROUPS_IN);
        fill<uint32_t>(stream, d_used.p, 2 * count, 0u);
    }
    StoreView v = view();
    const uint32_t wg = wave_grid(count);
    d_big_groups.ensure(2 * count);   // large-row lists, out then in direction
    uint32_t *big_out = d_big_groups.p, *big_in = d_big_groups.p + count;
    // each launch timed on its own (HIP events); the bytes of both directions
    // go to the first launch of each kernel, so sums and averages hold.  The
    // in-row pass runs on stream2 beside the out-row pass: the two touch
    // disjoint rows and have their own occurrence counters and large-row lists
    // (shared counters and read tallies are atomics).  A batch that the
    // one-workgroup sort took runs both passes on the engine stream: their
    // kernels take microseconds, less than the fork's and join's cross-stream
    // waits.
    const bool fork = count > kSmallSort || large_batches;
    hipStream_t s_in = fork ? stream2 : stream;
    // such a batch also times its pass as one record: an event between two
    // dependent launches costs ~5 us of gap, more than those kernels take
    auto tb = [&](int kernel, hipStream_t on) { return fork ? time_begin(kernel, -1, 0, on) : -1; };
    auto te = [&](int slot) {
        if (slot >= 0) time_end(slot);
    };
    const int pass = time_begin(DCG_K_DELETE_PASS, -1);
    if (fork) {
        DCG_HIP(hipEventRecord(fork_ev, stream));
        DCG_HIP(hipStreamWaitEvent(stream2, fork_ev, 0));
    }
    const int w0 = tb(DCG_K_DELETE_ROWS, stream);
    hipLaunchKernelGGL(k_del_rows, dim3(wg), dim3(256), 0, stream, v, d_out.p, d_sorted_out.p, d_hot_out.p, kout,
                       count, d_gstart.p, d_counters.p + C_GROUPS_OUT, used_out, d_counters.p, 1, big_out,
                       d_counters.p + C_NBIG, d_fbase_out.p, d_fence_out.p, policy_reads());
    te(w0);
    const int b0 = tb(DCG_K_DELETE_BIG, stream);
    hipLaunchKernelGGL(k_del_big, dim3(kDelBigGrid), dim3(1024), 0, stream, v, d_out.p, d_sorted_out.p, d_hot_out.p,
                       kout, count, big_out, d_counters.p + C_NBIG, used_out, d_counters.p, 1, d_fbase_out.p,
                       d_fence_out.p);
    te(b0);
    const int w1 = tb(DCG_K_DELETE_ROWS, s_in);
    hipLaunchKernelGGL(k_del_rows, dim3(wg), dim3(256), 0, s_in, v, d_in.p, d_sorted_in.p, d_hot_in.p, kin,
                       count, d_gstart2.p, d_counters.p + C_GROUPS_IN, used_in, d_counters.p, 0, big_in,
                       d_counters.p + C_NBIG_IN, d_fbase_in.p, d_fence_in.p, policy_reads());
    te(w1);
    const int b1 = tb(DCG_K_DELETE_BIG, s_in);
    hipLaunchKernelGGL(k_del_big, dim3(kDelBigGrid), dim3(1024), 0, s_in, v, d_in.p, d_sorted_in.p, d_hot_in.p,
                       kin, count, big_in, d_counters.p + C_NBIG_IN, used_in, d_counters.p, 0, d_fbase_in.p,
                       d_fence_in.p);
    te(b1);
    if (fork) {
        DCG_HIP(hipEventRecord(join_ev, stream2));
        DCG_HIP(hipStreamWaitEvent(stream, join_ev, 0));
    }
    time_end(pass);
    DCG_HIP(hipGetLastError());
    DCG_HIP(hipMemcpyAsync(h_counters, d_counters.p, C_NUM * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
    mut_slot[0] = w0;
    mut_slot[1] = w1;
    mut_slot[2] = b0;
    mut_slot[3] = b1;
    mut_slot[4] = pass;
}

void Engine::finish_delete(uint64_t count, dcg_phase_stats *st) {
    if (!count) return;
    const double wb =
        4.0 * (double)(h_counters[C_W_READ] + h_counters[C_W_WRITTEN]) + 16.0 * (double)h_counters[C_W_ROWS];
    const double bb = 4.0 * (double)(h_counters[C_SLOTS_READ] + h_counters[C_SLOTS_WRITTEN]) +
                      16.0 * (double)h_counters[C_B_ROWS];
    if (mut_slot[0] >= 0) {   // per-kernel records (a batch past the one-workgroup sort)
        timed[mut_slot[0]].bytes = wb;
        timed[mut_slot[1]].bytes = 0.0;
        timed[mut_slot[0]].host_lines = h_counters[C_W_HOST_LINES];
        timed[mut_slot[2]].bytes = bb;
        timed[mut_slot[3]].bytes = 0.0;
    }
    timed[mut_slot[4]].bytes = wb + bb;
    timed[mut_slot[4]].host_lines = h_counters[C_W_HOST_LINES];
    if (st) {
        st->touched_sources = h_counters[C_GROUPS_OUT];
        st->changed_sources = h_counters[C_CHANGED];
        st->missing_deletes = h_counters[C_MISSING];
    }
    num_edges -= count - h_counters[C_MISSING];
}

void Engine::mutate_add(const dcg_edge *recs, uint64_t count, dcg_phase_stats *st) {
    if (!count) {
        fill<unsigned long long>(stream, d_counters.p, C_NUM, 0ull);
        return;
    }
    if (count >= kMaxBatch) throw Error(DCG_ERR_ARG, "update batches hold fewer than 2^30 records");
    d_k32a.ensure(2 * count);   // sorted keys and payloads, out then in direction
    d_v32a.ensure(2 * count);
    d_gstart.ensure(count);
    d_gstart2.ensure(count);
    uint32_t *kout = d_k32a.p, *kin = d_k32a.p + count, *vout = d_v32a.p, *vin = d_v32a.p + count;
    if (count <= kSmallSort && !large_batches) {
        hipLaunchKernelGGL(k_add_prep_small, dim3(2), dim3(1024), 0, stream, recs, (uint32_t)count, n, kout, vout,
                           kin, vin, d_gstart.p, d_gstart2.p, d_out.p, d_in.p, d_tail.p, d_counters.p);
    } else {
        fill<unsigned long long>(stream, d_counters.p, C_NUM, 0ull);
        d_k32c.ensure(2 * count);
        d_v32c.ensure(2 * count);
        // both directions' (key, payload) pairs unsorted in the c buffers, one
        // stable sort into kout | kin (payloads into vout | vin)
        const uint32_t nb = (uint32_t)bits_for(n ? n - 1 : 0);
        hipLaunchKernelGGL(k_add_keys, dim3(grid_for(count)), dim3(256), 0, stream, recs, count, n, nb, d_k32c.p,
                           d_v32c.p, d_counters.p);
        const size_t tb = radix_sort_temp_bytes<uint32_t>(2 * count, true);
        radix_sort<uint32_t>(stream, d_k32c.p, kout, d_v32c.p, vout, 2 * count, 0, (int)add_dir_bit(nb) + 1,
                             temp(tb), tb);
        hipLaunchKernelGGL(k_split_add_keys, dim3(grid_for(2 * count)), dim3(256), 0, stream, kout, count, nb);
        group_heads<uint32_t>(*this, kout, count, d_gstart.p, d_counters.p + C_GROUPS_OUT);
        group_heads<uint32_t>(*this, kin, count, d_gstart2.p, d_counters.p + C_GROUPS_IN);
        hipLaunchKernelGGL(k_add_plan, dim3(grid_for(count, 256, 1024)), dim3(256), 0, stream, d_out.p, kout,
                           count, d_gstart.p, d_counters.p + C_GROUPS_OUT, d_tail.p);
        hipLaunchKernelGGL(k_add_plan, dim3(grid_for(count, 256, 1024)), dim3(256), 0, stream, d_in.p, kin,
                           count, d_gstart2.p, d_counters.p + C_GROUPS_IN, d_tail.p);
    }
    // free pool: live entries and the blocks the previous batch retired, sorted
    // by class (taken entries, ~0, sort last); this batch retires past them
    const uint64_t pool_live = pool_len - pool_taken;
    if (d_pool.n < pool_live + 2 * count + 1) {
        DevBuf<uint64_t> grown;
        grown.alloc(pool_live + 2 * count + 1 + pool_live / 2);
        if (pool_len)
            DCG_HIP(hipMemcpyAsync(grown.p, d_pool.p, pool_len * 8, hipMemcpyDeviceToDevice, stream));
        std::swap(grown.p, d_pool.p);
        std::swap(grown.n, d_pool.n);
    }
    if (pool_len <= kSmallSort && !large_batches) {
        hipLaunchKernelGGL(k_pool_small, dim3(1), dim3(1024), 0, stream, d_pool.p, (uint32_t)pool_len,
                           (uint32_t)pool_live, d_tail.p, total_slots);
    } else {
        fill<uint64_t>(stream, d_tail.p + T_FREE, kClasses, 0ull);
        d_k64c.ensure(pool_len);
        const size_t pb = radix_sort_temp_bytes<uint64_t>(pool_len, false);
        radix_sort<uint64_t>(stream, d_pool.p, d_k64c.p, nullptr, nullptr, pool_len, 0, kPoolKeyBits, temp(pb), pb);
        hipLaunchKernelGGL(k_pool_segments, dim3(grid_for(pool_live, 256, 4096)), dim3(256), 0, stream, d_k64c.p,
                           pool_live, d_tail.p);
        hipLaunchKernelGGL(k_pool_counts, dim3(1), dim3(64), 0, stream, d_tail.p);
        if (pool_live)
            DCG_HIP(hipMemcpyAsync(d_pool.p, d_k64c.p, pool_live * 8, hipMemcpyDeviceToDevice, stream));
        fill<uint64_t>(stream, d_tail.p + T_RETIRED, 1, pool_live);
        hipLaunchKernelGGL(k_alloc_classes, dim3(1), dim3(64), 0, stream, d_tail.p, total_slots);
    }
    d_k64a.ensure(count);
    d_k64b.ensure(count);
    uint64_t *reloc_out = d_k64a.p, *reloc_in = d_k64b.p;
    const AppendSide side_out{d_out.p, kout, d_gstart.p, d_counters.p + C_GROUPS_OUT, reloc_out};
    const AppendSide side_in{d_in.p, kin, d_gstart2.p, d_counters.p + C_GROUPS_IN, reloc_in};
    hipLaunchKernelGGL(k_add_assign, dim3(grid_for(count, 256, 1024), 2), dim3(256), 0, stream, side_out, side_in,
                       count, d_tail.p, d_pool.p);
    StoreView v = view();
    const uint32_t wg = wave_grid(count);
    d_big_groups.ensure(2 * count);   // large-row lists, out then in direction
    uint32_t *big_out = d_big_groups.p, *big_in = d_big_groups.p + count;
    const int slot = time_begin(DCG_K_APPEND_ROWS, -1);
    hipLaunchKernelGGL(k_add_rows, dim3(wg), dim3(256), 0, stream, v, d_out.p, d_hot_out.p, kout, vout, count,
                       d_gstart.p, d_counters.p + C_GROUPS_OUT, d_tail.p, d_pool.p, reloc_out, d_counters.p, 1,
                       big_out, d_counters.p + C_NBIG);
    hipLaunchKernelGGL(k_add_big, dim3(64), dim3(1024), 0, stream, v, d_out.p, d_hot_out.p, kout, vout, count,
                       d_gstart.p, big_out, d_counters.p + C_NBIG, reloc_out, d_tail.p, d_pool.p, d_counters.p, 1);
    hipLaunchKernelGGL(k_add_rows, dim3(wg), dim3(256), 0, stream, v, d_in.p, d_hot_in.p, kin, vin, count,
                       d_gstart2.p, d_counters.p + C_GROUPS_IN, d_tail.p, d_pool.p, reloc_in, d_counters.p, 1,
                       big_in, d_counters.p + C_NBIG_IN);
    hipLaunchKernelGGL(k_add_big, dim3(64), dim3(1024), 0, stream, v, d_in.p, d_hot_in.p, kin, vin, count,
                       d_gstart2.p, big_in, d_counters.p + C_NBIG_IN, reloc_in, d_tail.p, d_pool.p, d_counters.p, 1);
    time_end(slot);
    DCG_HIP(hipGetLastError());
    // the add stage's counters have their own pinned mirror: a snapshot step
    // reads both stages' counters after one wait
    DCG_HIP(hipMemcpyAsync(h_counters + kCounterSlots, d_counters.p, C_NUM * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, stream));
    DCG_HIP(hipMemcpyAsync(h_tail, d_tail.p, T_NUM * sizeof(uint64_t), hipMemcpyDeviceToHost, stream));
    add_slot = slot;
}

// After the add mutation's readbacks landed.  A store that ran out of slots
// skipped the rows it could not relocate (rows stay consistent) and is
// reported here.
void Engine::finish_add(uint64_t count, dcg_phase_stats *st) {
    if (!count) return;
    const uint64_t *tl = h_tail;
    const unsigned long long *hc = h_counters + kCounterSlots;
    const int slot = add_slot;
    pool_len = tl[T_RETIRED];
    pool_taken = 0;
    for (int c = 0; c < kClasses; ++c) pool_taken += tl[T_TAKE + c];
    reused_blocks += pool_taken;
    timed[slot].bytes = 4.0 * (double)(hc[C_SLOTS_READ] + hc[C_SLOTS_WRITTEN]) +
                        16.0 * (double)(hc[C_GROUPS_OUT] + hc[C_GROUPS_IN]);
    if (tl[T_OVERFLOW]) throw Error(DCG_ERR_CAPACITY, "chunk store slot space exhausted (raise store_slack)");
    used_slots = tl[T_TAIL];
    if (st) {
        st->touched_sources = hc[C_GROUPS_OUT];
        st->changed_sources = hc[C_GROUPS_OUT];
        st->invalid_additions = hc[C_INVALID];
        st->relocated_rows = hc[C_RELOCATED];
    }
    num_edges += count - hc[C_INVALID];
}

// ---------------------------------------------------------------------- export
__global__ void k_export_deg(const RowDesc *out, uint32_t n, uint64_t *deg) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x)
        deg[v] = row_live(out[v]);
}

__global__ void k_export_rows(StoreView s, uint32_t n, const uint64_t *rp, uint32_t *col) {
    const uint32_t lane = lane_id();
    const uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / kWave;
    const uint64_t nwaves = (uint64_t)gridDim.x * blockDim.x / kWave;
    for (uint64_t v = wave; v < n; v += nwaves) {
        const RowDesc d = s.out[v];
        if (!row_tombs(d)) {
            for (uint32_t i = lane; i < d.ext; i += kWave) col[rp[v] + i] = load_slot(s, d.off + i);
            continue;
        }
        uint64_t w = rp[v];   // skip tombstones, order kept
        for (uint32_t t0 = 0; t0 < d.ext; t0 += kWave) {
            const uint32_t i = t0 + lane;
            const uint32_t x = i < d.ext ? load_slot(s, d.off + i) : kTomb;
            const uint64_t km = __ballot(!(x & kTomb));
            if (!(x & kTomb)) col[w + lanes_below(km)] = x;
            w += (uint64_t)__popcll(km);
        }
    }
}

void Engine::export_csr(uint64_t *row_ptr, uint32_t *col, uint64_t *m) {
    DevBuf<uint64_t> deg, rp;
    deg.alloc(n + 1);
    rp.alloc(n + 1);
    fill<uint64_t>(stream, deg.p, n + 1, 0ull);
    hipLaunchKernelGGL(k_export_deg, dim3(grid_for(n)), dim3(256), 0, stream, d_out.p, n, deg.p);
    const size_t tb = scan_temp_bytes<uint64_t>((uint64_t)n + 1);
    exclusive_sum<uint64_t>(stream, deg.p, rp.p, (uint64_t)n + 1, temp(tb), tb);
    uint64_t total = 0;
    DCG_HIP(hipMemcpyAsync(&total, rp.p + n, 8, hipMemcpyDeviceToHost, stream));
    sync();
    if (m) *m = total;
    if (!row_ptr) return;
    DCG_HIP(hipMemcpyAsync(row_ptr, rp.p, (n + 1) * 8, hipMemcpyDeviceToHost, stream));
    if (col && total) {
        DevBuf<uint32_t> dc;
        dc.alloc(total);
        hipLaunchKernelGGL(k_export_rows, dim3(wave_grid(n)), dim3(256), 0, stream, view(), n, rp.p, dc.p);
        DCG_HIP(hipMemcpyAsync(col, dc.p, total * 4, hipMemcpyDeviceToHost, stream));
        sync();
    }
    sync();
}

// --------------------------------------------------- per-phase read accounting
// Folds per-vertex expansion counters into per-chunk edge reads (the cache
// policy's signal) and splits reads into HBM-cache and host-link reads.
// Fold the listed rows' read passes into per-chunk reads for the cache policy
// (64-byte lines: a pass over a row segment reads ceil(len / 16) lines; rows)
// and into the cached / host edge-read counters; counters and the list are
// cleared for the next fold.
__global__ void k_fold_reads(StoreView s, const unsigned long long *nlist_p, float *reads, float *rows,
                             unsigned long long *counters) {
    __shared__ WgSum<2> tot;
    wg_sum_init(tot);
    __syncthreads();
    unsigned long long cached = 0, host = 0;
    const uint64_t nlist = *nlist_p;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nlist;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t e = s.hot_list[i], v = e & 0x7FFFFFFFu;
        const bool in = (e >> 31) != 0;
        uint32_t *hot = in ? s.hot_in : s.hot_out;
        const uint32_t h = hot[v];
        hot[v] = 0;
        if (!in) s.hot_window[v] = window_add(s.hot_window[v], h);   // a listed row is folded once
        const RowDesc d = in ? s.in[v] : s.out[v];
        uint64_t a = d.off, end = d.off + d.ext;
        while (a < end) {
            const uint64_t c = a >> kChunkLog2;
            const uint64_t ce = (c + 1) << kChunkLog2;
            const uint64_t b = end < ce ? end : ce;
            if (reads) {
                atomicAdd(&reads[c], (float)h * (float)((b - a + 15) / 16));   // 64-byte lines read
                atomicAdd(&rows[c], (float)h);
            }
            if (s.chunk_frame[c] >= 0) cached += (unsigned long long)h * (b - a);
            else host += (unsigned long long)h * (b - a);
            a = b;
        }
    }
    // wave-reduce the two counters before the global atomics
    for (int d = 32; d >= 1; d >>= 1) {
        cached += __shfl_xor(cached, d, kWave);
        host += __shfl_xor(host, d, kWave);
    }
    unsigned long long *const dst[2] = {&counters[C_CACHED_READS], &counters[C_HOST_READS]};
    const unsigned long long x[2] = {cached, host};
    wg_sum_flush(tot, x, dst);
}

void Engine::fold_enqueue(int stage) {
    if (cfg.placement == DCG_PLACE_DEVICE) return;
    fill<unsigned long long>(stream, d_counters.p + C_CACHED_READS, 2, 0ull);
    float *reads = cache.d_reads.n ? cache.d_reads.p : nullptr;
    float *rows = cache.d_reads.n ? cache.d_reads.p + cache.chunks : nullptr;
    // at most 2n rows are listed; the grid strides over what the list holds
    hipLaunchKernelGGL(k_fold_reads, dim3(grid_for(2ull * n, 256, 4096)), dim3(256), 0, stream, view(), d_hot_count.p,
                       reads, rows, d_counters.p);
    fill<unsigned long long>(stream, d_hot_count.p, 1, 0ull);
    DCG_HIP(hipGetLastError());
    DCG_HIP(hipMemcpyAsync(h_fold + 2 * stage, d_counters.p + C_CACHED_READS, 2 * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, stream));
}

// After the fold's readback landed (and, for a device store, after the
// phase's edges_scanned is known: every read is an HBM read).
void Engine::fold_finish(dcg_phase_stats *st, int stage) {
    if (cfg.placement == DCG_PLACE_DEVICE) {
        if (st) st->cached_edge_reads += st->edges_scanned;
        cum_cached_reads += st ? st->edges_scanned : 0;
        return;
    }
    const unsigned long long *f = h_fold + 2 * stage;
    cum_cached_reads += f[0];
    cum_host_reads += f[1];
    if (st) {
        st->cached_edge_reads += f[0];
        st->host_edge_reads += f[1];
    }
}

void Engine::fold_reads(dcg_phase_stats *st) {
    fold_enqueue();
    sync();
    fold_finish(st);
}

}  // namespace dcg
