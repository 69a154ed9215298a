// dcg_internal.hpp -- shared device structures and wave64 primitives of the
// MI355X incremental streaming-graph engine (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/dcg.h"

namespace dcg {

constexpr uint32_t kInf = 0xFFFFFFFFu;
constexpr uint64_t kInfState = 0xFFFFFFFFFFFFFFFFull;   // (value INF, parent INF)
constexpr uint32_t kChunkLog2 = 16;                       // 65,536-edge chunks
constexpr uint64_t kChunkEdges = 1ull << kChunkLog2;
constexpr uint32_t kMinRowCap = 4;                        // source_local_chunk_store.h:133-145
constexpr uint32_t kWave = 64;
constexpr uint32_t kBlock = 256;                          // 4 waves per workgroup
constexpr uint32_t kPBlock = 1024;                        // persistent closures: 16 waves per workgroup
constexpr uint32_t kBigRow = 256;                         // rows above this go to the piece list
constexpr uint32_t kPiece = 512;                          // edges per big-row piece
constexpr int kClasses = 40;                              // power-of-two capacity classes

// One adjacency row: a power-of-two block of slots (meta == 0 for an empty row).
// A block of <= 65,536 slots is aligned to its capacity, so it never crosses a
// chunk; larger blocks are whole, aligned runs of chunks.
//   ext   slots in use: live entries plus tombstones
//   meta  log2(capacity) in bits 0..5, tombstone count in bits 6..31
// Only rows mutated by the large-row delete path hold tombstones: a deleted
// entry of their sorted prefix keeps its id in bits 0..30 and gets bit 31
// (kTomb), so the prefix stays binary-searchable; readers skip such entries.
struct RowDesc {
    uint64_t off;
    uint32_t ext;
    uint32_t meta;
};

constexpr uint32_t kTomb = 0x80000000u;
constexpr uint32_t kTombShift = 6;
constexpr uint32_t kMaxTombs = (1u << (32 - kTombShift)) - 1;

__host__ __device__ __forceinline__ uint64_t row_cap(const RowDesc &d) { return d.meta ? 1ull << (d.meta & 63u) : 0ull; }
__host__ __device__ __forceinline__ uint32_t row_tombs(const RowDesc &d) { return d.meta >> kTombShift; }
__host__ __device__ __forceinline__ uint32_t row_live(const RowDesc &d) { return d.ext - row_tombs(d); }
__host__ __device__ __forceinline__ uint32_t make_meta(uint32_t log2cap, uint32_t tombs) {
    return log2cap | (tombs << kTombShift);
}

// Read/write view of the chunk store passed to kernels by value.
struct StoreView {
    RowDesc *out;              // n out-rows
    RowDesc *in;               // n in-rows
    uint32_t *host;            // pinned host slots (device-visible), null when device-placed
    uint32_t *frames;          // device chunk frames (frame f holds slots [f<<16, (f+1)<<16) of its chunk)
    const int32_t *chunk_frame;// chunk -> frame, -1 when not resident
    uint32_t *hot_out;         // per-vertex out-row expansions this snapshot
    uint32_t *hot_in;          // per-vertex in-row expansions this snapshot
    uint32_t *hot_list;        // rows whose counter left zero (v | direction << 31); null when device-placed
    unsigned long long *hot_count;
    uint32_t *hot_window;      // per vertex: out-row expansions of the last 4 snapshots (saturating bytes)
};

// Add h to the current snapshot's byte of a hotness window word (saturating at 255).
__host__ __device__ __forceinline__ uint32_t window_add(uint32_t w, uint32_t h) {
    const uint32_t b = (w & 0xFFu) + h;
    return (w & ~0xFFu) | (b > 255u ? 255u : b);
}

// Count one read pass over row v (hot is hot_out or hot_in).  With a host
// store the first count of a row since the last fold lists the row, so the
// fold visits only rows that were read.
__device__ __forceinline__ void mark_read(const StoreView &s, uint32_t *hot, uint32_t v) {
    if (!s.hot_list) {
        hot[v] += 1;
        return;
    }
    if (atomicAdd(&hot[v], 1u) == 0u)
        s.hot_list[atomicAdd(s.hot_count, 1ull)] = v | (hot == s.hot_in ? 0x80000000u : 0u);
}

// A device-placed store (no host copy) is its own frame array, chunk c in
// frame c: its slots resolve without the chunk table, one dependent load less
// on every row access.
__device__ __forceinline__ const uint32_t *row_base(const StoreView &s, uint64_t off) {
    if (!s.host) return s.frames + off;
    const uint32_t c = (uint32_t)(off >> kChunkLog2);
    const int32_t f = s.chunk_frame[c];
    return f >= 0 ? s.frames + (((uint64_t)f) << kChunkLog2) + (off & (kChunkEdges - 1))
                  : s.host + off;
}

// Frame copy of a slot, or null when its chunk is not resident.
__device__ __forceinline__ uint32_t *frame_w(const StoreView &s, uint64_t off) {
    if (!s.host) return s.frames + off;
    const int32_t f = s.chunk_frame[(uint32_t)(off >> kChunkLog2)];
    return f >= 0 ? s.frames + (((uint64_t)f) << kChunkLog2) + (off & (kChunkEdges - 1)) : nullptr;
}

// Write-through: a store reaches the frame of a resident chunk and the pinned
// host copy, so the host copy is always current and eviction copies nothing.
__device__ __forceinline__ void store_slot(const StoreView &s, uint64_t off, uint32_t v) {
    uint32_t *f = frame_w(s, off);
    if (f) *f = v;
    if (s.host) s.host[off] = v;
}

// Weight policy: REFERENCE (u+v)%128+1 (api.cuh:26-28), SEEDED splitmix64,
// UNIT (BFS), ZERO (CC label forwarding).
enum WeightKind : uint32_t { W_REFERENCE = 0, W_SEEDED = 1, W_UNIT = 2, W_ZERO = 3 };
struct WeightParams {
    uint32_t kind;
    uint32_t wmax;
    uint64_t seed;
};

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

__host__ __device__ __forceinline__ uint32_t edge_weight(const WeightParams &w, uint32_t u, uint32_t v) {
    switch (w.kind) {
    case W_REFERENCE: return (uint32_t)(((uint64_t)u + v) % 128u + 1u);
    case W_SEEDED: return 1u + (uint32_t)(mix64(w.seed ^ (((uint64_t)u << 32) | v)) % w.wmax);
    case W_UNIT: return 1u;
    default: return 0u;
    }
}

// ---- wave64 primitives -------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    // popcount of mask bits below this lane (v_mbcnt)
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Positions in a global list for the flagged lanes of a workgroup, with one
// atomic on the list counter per workgroup.  Every thread of the workgroup
// calls it (block-uniform control flow); wcnt (one word per wave) and wbase
// live in LDS.
__device__ __forceinline__ uint32_t wg_append(bool want, uint32_t *counter, uint32_t *wcnt, uint32_t &wbase) {
    const uint64_t mask = __ballot(want);
    const uint32_t w = threadIdx.x / kWave, nw = (blockDim.x + kWave - 1) / kWave;
    if (lane_id() == 0) wcnt[w] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t q = 0; q < nw; ++q) t += wcnt[q];
        wbase = t ? atomicAdd(counter, t) : 0u;
    }
    __syncthreads();
    uint32_t pos = wbase;
    for (uint32_t q = 0; q < w; ++q) pos += wcnt[q];
    pos += lanes_below(mask);
    __syncthreads();   // wcnt and wbase are reused by the next call
    return pos;
}

// Workgroup totals for a few global counters: every wave adds its totals
// (held by lane 0) into LDS and the last wave of the workgroup to arrive adds
// the sums to the global counters, so a workgroup costs one atomic per counter
// instead of one per wave (atomics on one word serialise at the L2: ~16 k of
// them take ~0.2 ms).  No barrier at the end, so waves may finish at different
// times.  wg_sum_init runs before a barrier at kernel start; every wave of the
// workgroup calls wg_sum_flush exactly once.
template <int K>
struct WgSum {
    unsigned long long v[K];
    uint32_t arrived;
};

template <int K>
__device__ __forceinline__ void wg_sum_init(WgSum<K> &w) {
    if (threadIdx.x < K) w.v[threadIdx.x] = 0;
    if (threadIdx.x == 0) w.arrived = 0;
}

template <int K>
__device__ __forceinline__ void wg_sum_flush(WgSum<K> &w, const unsigned long long (&x)[K],
                                             unsigned long long *const (&dst)[K]) {
    if (lane_id() != 0) return;
#pragma unroll
    for (int j = 0; j < K; ++j)
        if (x[j]) atomicAdd(&w.v[j], x[j]);
    __threadfence_block();
    if (atomicAdd(&w.arrived, 1u) + 1 == (blockDim.x + kWave - 1) / kWave) {
        __threadfence_block();
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const unsigned long long sum = atomicAdd(&w.v[j], 0ull);
            if (sum && dst[j]) atomicAdd(dst[j], sum);
        }
    }
}

template <typename T>
__device__ __forceinline__ T wave_shfl(T v, int src) { return __shfl(v, src, kWave); }

// Inclusive prefix sum over the wave (Hillis-Steele with DPP-backed shuffles).
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    const uint32_t l = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, kWave);
        if (l >= (uint32_t)d) x += y;
    }
    return x;
}

// Like wg_append for k entries per lane: the lane's first position in a global
// list, with one atomic on the list counter per workgroup.  Every thread of the
// workgroup calls it (block-uniform control flow); wcnt and wbase live in LDS.
__device__ __forceinline__ uint32_t wg_reserve(uint32_t k, uint32_t *counter, uint32_t *wcnt, uint32_t &wbase) {
    const uint32_t incl = wave_incl_scan(k);
    const uint32_t w = threadIdx.x / kWave, nw = (blockDim.x + kWave - 1) / kWave;
    if (lane_id() == kWave - 1) wcnt[w] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t q = 0; q < nw; ++q) t += wcnt[q];
        wbase = t ? atomicAdd(counter, t) : 0u;
    }
    __syncthreads();
    uint32_t pos = wbase;
    for (uint32_t q = 0; q < w; ++q) pos += wcnt[q];
    pos += incl - k;
    __syncthreads();   // wcnt and wbase are reused by the next call
    return pos;
}

// Wave-aggregated append: lanes with `want` get consecutive slots of `counter`.
// Returns the slot for this lane (undefined where !want).
__device__ __forceinline__ uint32_t wave_append(uint32_t *counter, bool want) {
    const uint64_t mask = __ballot(want);
    if (!mask) return 0;
    const int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if ((int)lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = wave_shfl(base, leader);
    return base + lanes_below(mask);
}

// Control block of one persistent launch (frontier counters ring of 3).
// One (query, vertex) slot: the packed state word (value << 32 | parent; rooted
// programs: value << 32 | depth), the round stamp that dedups frontier entries,
// and the rooted programs' parent.  Kept together (16 B, one 64-byte line holds
// four) so a relaxation touches one line for its word and stamp, not two.
struct alignas(16) VSlot {
    uint64_t st;
    uint32_t tick;
    uint32_t par;
};

// Grid barriers of kBarTreeMin or more workgroups arrive in two levels: the
// workgroups of each of kBarGroups groups (blockIdx % kBarGroups) at their
// group's counter, the last of each group at the top counter.
constexpr uint32_t kBarGroups = 8;
constexpr uint32_t kBarTreeMin = 256;
struct alignas(256) BarWord {
    uint32_t w;
};

struct alignas(256) Ctl {
    uint32_t count[3];        // entries in the queue consumed in round r (index r%3)
    uint32_t cursor[3];       // small-row work cursor of round r
    unsigned long long big[3];// (entries << 40) | pieces for round r
    uint32_t big_cursor[3];   // piece cursor of round r
    uint32_t rounds;          // rounds executed
    uint32_t overflow;        // a queue overflowed its capacity
    uint32_t aff_count;       // affected-list length (invalidation)
    uint32_t done;            // dense PageRank sweeps: converged or at the round cap
    unsigned long long edges; // adjacency entries scanned
    unsigned long long expansions; // rows expanded
    uint32_t solo_round;      // round reached by a workgroup-0 stretch
    uint32_t solo_resume_b;   // that stretch stopped before the round's pieces
    unsigned long long bad;   // PR: non-finite values seen
    uint32_t far_count[2];    // entries of the two far piles (bucketed push closure)
    uint32_t far_par;         // pile that collects this epoch's far entries
    uint32_t advances;        // bucket advances taken
    unsigned long long far_scanned;   // far-pile entries examined by the advances
    uint32_t rmax[2];         // dense PageRank sweeps: largest |residual| (float bits) by round parity
    uint32_t bar_timeout;     // a workgroup waited ~10 s at the barrier (the launch gives up)
    // device wall clock (s_memrealtime ticks) at the first workgroup's start and
    // the last workgroup's end of a persistent launch: its duration also when
    // the launch is a node of a replayed graph
    unsigned long long t_first, t_last;
    // grid-barrier words, each on a 256-byte line of its own: every workgroup
    // takes one atomic on an arrival counter while the waiters poll the
    // generation word, and a shared line queues the polls behind the atomics
    BarWord bar_arrive;               // arrivals at the top level in the current generation
    BarWord bar_gen;                  // generation, moved by the last arrival
    BarWord bar_group[kBarGroups];    // arrivals per group of workgroups (two-level barrier)
};

// Per-stage arguments read by the kernels from device memory, so one captured
// graph of a stage's launches serves every snapshot: the stage's records (the
// device batch slot they were copied to, and their count) and the round-stamp
// bases of its closures.  A graph never holds a record address itself.
struct StepArgs {
    const dcg_edge *recs;
    uint64_t count;
    uint32_t stamp[4];
};

struct BigEntry {
    uint64_t off;      // first slot of the row
    uint64_t val;      // owner state word (min-lattice) or float bits (PR delta/deg)
    uint32_t g;        // global entry id (q*n + v)
    uint32_t deg;
    uint32_t piece0;   // first piece index
    uint32_t pad;
};

template <typename T>
__global__ void k_fill(T *p, uint64_t n, T v) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Fill kernel (used instead of hipMemsetAsync so the sequence stays capturable
// in a hipGraph).
template <typename T>
inline void fill(hipStream_t s, T *p, uint64_t n, T v) {
    if (!n) return;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill<T>, dim3((uint32_t)blocks), dim3(256), 0, s, p, n, v);
}

inline uint32_t grid_for(uint64_t work, uint32_t block = 256, uint32_t cap = 65536) {
    uint64_t g = (work + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (uint32_t)g;
}

}  // namespace dcg
