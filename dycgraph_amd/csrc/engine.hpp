// engine.hpp -- host-side engine object behind the C ABI (include/dcg.h).
#pragma once
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "dcg_internal.hpp"

namespace dcg {

struct Error : std::runtime_error {
    int status;
    Error(int st, const std::string &m) : std::runtime_error(m), status(st) {}
};

#define DCG_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            throw ::dcg::Error(DCG_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_) + \
                                                " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")"); \
    } while (0)

// Device buffer owned by the engine.
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        release();
        if (count) DCG_HIP(hipMalloc(&p, count * sizeof(T)));
        n = count;
    }
    void ensure(size_t count) {
        if (count > n) alloc(count);
    }
    size_t bytes() const { return n * sizeof(T); }
};

struct MlpState;   // learned cache policy (mlp.hip)
struct MlpDeleter {
    void operator()(MlpState *m) const;
};

// Per-snapshot chunk access statistics and cache residency.
struct CacheState {
    uint64_t chunks = 0, frames = 0;
    std::vector<int32_t> chunk_frame;   // host mirror
    std::vector<int64_t> frame_chunk;   // frame -> chunk (-1 free)
    DevBuf<int32_t> d_chunk_frame;
    DevBuf<uint32_t> d_frames;          // frames * 65536 slots (device placement: all chunks)
    DevBuf<float> d_reads;              // edges read per chunk, this snapshot
    DevBuf<float> d_hist;               // 4-snapshot history of reads (chunk-major, 4 floats)
    DevBuf<float> d_touch;              // update records touching each chunk, this batch
    DevBuf<float> d_rows_last;          // rows expanded per chunk in the last snapshot
    DevBuf<float> d_score;
    DevBuf<int64_t> d_load;             // (frame, chunk) pairs of one admission decision
    std::vector<uint32_t> last_use;     // snapshot of last access (LRU tie-break)
    std::vector<float> hot_score;       // the last maintenance's 4-snapshot read sums (hotness ranking)
    uint64_t loaded = 0, evicted = 0, h2d = 0;
    std::unique_ptr<MlpState, MlpDeleter> mlp;
};

struct Engine {
    dcg_config cfg{};
    std::string err;
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;          // the in-row delete pass runs here, beside the out-row pass
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    int cu_count = 0;
    uint32_t coop_grid[10] = {};   // persistent grid per mode (5 + mode: static convergence)
    bool coop_plain = false;       // cooperative launches refused: plain launches of the same grid
    uint64_t bar_spin = 1000000000ull;   // grid-barrier patience in wall-clock ticks (10 s at 100 MHz)
    struct Tune {
        uint32_t blocks_per_cu = 1;   // persistent workgroups (1,024 threads) per CU
        uint32_t big_row = kBigRow;   // rows above this are split into pieces
        uint32_t piece = kPiece;
        uint32_t graphs = 1;            // replay each stage's repair launches as a captured hipGraph
        uint32_t bucket_quarters = ~0u;   // push-closure bucket width: SSSP w_max*q/4, BFS q; 0 off, ~0 auto
        uint32_t solo_max = 2048;         // frontiers this small run on one workgroup (no grid barriers)
        uint32_t grid_shift = 0;          // persistent grids span (workgroups per CU * CUs) >> grid_shift
    } tune;
    // A stage the stage tuner has tuned runs its repair closures with its own
    // grid, solo threshold and row split (bucket widths stay shared: they live
    // in device memory); everything else uses `tune`.
    Tune stage_tune[2];
    bool stage_tuned[2] = {false, false};
    dcg_tune_result stage_result[2]{};   // the last tuning of each stage
    void forget_stage_tunes() {
        stage_tuned[0] = stage_tuned[1] = false;
        stage_result[0] = stage_result[1] = dcg_tune_result{};
    }
    bool large_batches = false;           // every batch through the device-wide sort path (no one-workgroup sort)
    uint32_t tune_arm = 0;                // repair stages still to be tuned (tune_stage)
    dcg_tune_result last_tune{};          // outcome of the last tuned stage

    uint32_t n = 0;
    uint32_t Q = 1;
    uint32_t app = DCG_APP_SSSP;
    std::vector<uint32_t> sources;
    std::vector<uint32_t> qapps;        // per-query program index in the program table
    bool apps_given = false;            // dcg_config.query_apps was set
    // The program table whose kernels the engine launches (the built-in
    // SSSP/BFS/CC set unless dcg_set_program installed another).
    const dcg_program *prog = nullptr;
    bool rooted(uint32_t q) const { return prog->flags[qapps[q]] & DCG_PROGRAM_ROOTED; }
    bool vertex_seeds(uint32_t q) const { return prog->flags[qapps[q]] & DCG_PROGRAM_VERTEX_SEEDS; }
    DevBuf<uint32_t> d_sources, d_qapps;
    WeightParams wp{};
    bool loaded = false, started = false;
    uint64_t snapshots = 0;

    // ---- chunk store ----
    uint64_t total_slots = 0;
    uint64_t used_slots = 0;           // host copy of the bump tail
    uint64_t num_edges = 0;
    DevBuf<RowDesc> d_out, d_in;
    uint32_t *h_slab = nullptr;        // pinned host slots (DCG_PLACE_HOST)
    DevBuf<uint64_t> d_tail;           // [0] bump tail, [1] overflow flag, [2..] class cursors
    DevBuf<uint32_t> d_hot_out, d_hot_in;
    DevBuf<uint32_t> d_hot_window;             // reference hotness window (4 saturating bytes per vertex)
    DevBuf<uint32_t> d_hot_list;               // rows read since the last fold (host store only)
    DevBuf<unsigned long long> d_hot_count;
    DevBuf<uint32_t> d_sorted_out, d_sorted_in;   // per row: length of the sorted prefix
    // Fences of the sorted prefixes, in HBM: row r owns fence[fbase[r] .. fbase[r+1]),
    // fence j holding the id at slot 16j of the row (valid for 16j below its
    // current prefix length); rows whose prefix started below 16 have none.
    DevBuf<uint32_t> d_fbase_out, d_fbase_in, d_fence_out, d_fence_in;
    DevBuf<uint64_t> d_pool;           // free blocks (class << 40 | offset) of relocated rows
    uint64_t pool_len = 0, pool_taken = 0, reused_blocks = 0;
    CacheState cache;

    // ---- vertex state ----
    DevBuf<VSlot> d_vs;                // Q*n slots: packed state (value<<32 | parent; CC: label<<32 | depth),
                                       // dedup stamp, CC parent
    bool has_par = false;              // a query runs a rooted program (CC): parents in the slots
    DevBuf<uint8_t> d_cc_dirty;        // per 1,024 state slots: stamped by a CC improvement since the last parent check
    DevBuf<uint32_t> d_far0, d_far1, d_fstamp;   // bucketed push closure: far piles, pile epochs
    DevBuf<uint32_t> d_thr, d_minfar, d_fdelta, d_epoch;
    void update_buckets();
    uint32_t max_out_deg = 0;          // largest out-degree at load
    // Bucket width in use: the set one, or (auto) 64 quarters on graphs whose
    // largest out-degree is at most 64 (grids, road networks: high diameter,
    // where the ordered closure pays) and none otherwise.
    uint32_t bucket_width() const {
        if (tune.bucket_quarters != ~0u) return tune.bucket_quarters;
        return max_out_deg <= 64 ? 64u : 0u;
    }
    uint32_t stamp = 1;
    DevBuf<uint32_t> d_qa, d_qb, d_qc, d_aff;
    DevBuf<BigEntry> d_big;
    DevBuf<Ctl> d_ctl;
    DevBuf<dcg_round_rec> d_rtrace;    // closure round trace (empty: off)
    DevBuf<uint32_t> d_rtcur;          // its append cursor
    Ctl *h_ctl = nullptr;              // pinned readback
    StepArgs *h_step = nullptr;        // pinned: [0] delete stage, [1] add stage
    DevBuf<StepArgs> d_step;
    hipGraphExec_t graph_exec[2] = {nullptr, nullptr};   // captured repair launches of each stage
    bool capturing = false;
    void run_stage(int stage);
    bool build_stage_graph(int stage);
    // part 0: the whole stage; for the delete stage, part 1 ends after the
    // reset of the affected set and part 2 runs the rest (pull, push, CC check)
    void enqueue_repair(int stage, int part = 0);

    // CPU/GPU co-execution of the insertion closure (coexec.hip).  The CPU
    // owns a vertex set; during an add stage it keeps those vertices' state
    // and relaxes their out-rows on host threads, the GPU closure sends it
    // proposals for them and receives its proposals for the others, until
    // neither side has any.
    struct CoExec {
        bool on = false, active = false;
        DevBuf<uint8_t> d_owner;                 // n: 1 = CPU-owned
        DevBuf<unsigned long long> d_bprop;      // Q*n proposal minima (~0: none)
        DevBuf<uint32_t> d_blist, d_bcount;      // proposed slots, their count
        DevBuf<uint32_t> d_idx;                  // the CPU-owned vertices (ascending)
        DevBuf<uint64_t> d_vals;                 // gather / scatter staging (Q x |C|)
        DevBuf<uint64_t> d_props;                // CPU -> GPU proposals (g << 32 | unused, packed)
        std::vector<uint32_t> vertices;          // CPU-owned vertex ids, ascending
        std::vector<int32_t> index;              // n: position in `vertices`, or -1
        std::vector<uint64_t> rp;                // host CSR of the CPU-owned out-rows
        std::vector<uint32_t> col;
        std::vector<uint64_t> st;                // Q x |C| packed (value << 32 | parent)
        std::vector<uint32_t> stamp;             // Q x |C| frontier dedup
        uint32_t round_stamp = 0;
        int threads = 1;
        uint64_t relax = 0, to_cpu = 0, to_gpu = 0, rounds = 0;   // last stage's counters
        void reset() {
            on = active = false;
            for (DevBuf<uint32_t> *b : {&d_blist, &d_bcount, &d_idx}) b->release();
            d_owner.release();
            d_bprop.release();
            d_vals.release();
            d_props.release();
            vertices.clear();
            index.clear();
            rp.clear();
            col.clear();
            st.clear();
            stamp.clear();
        }
    } cx;
    void coexec_set(const uint8_t *cpu_owned);
    void coexec_add_stage();
    void enqueue_coexec_seeds();
    void enqueue_coexec_round(const std::vector<uint32_t> &gs, const std::vector<uint64_t> &packed);
    void coexec_apply(const std::vector<uint32_t> &gs, const std::vector<uint64_t> &packed, uint32_t stamp, Ctl *c4);

    // Diagnostic traces (traces.hip).  Repair snapshot: the first delete stage
    // after arming writes the affected set, its in-rows and the states of the
    // involved vertices before and after the repair.
    std::string i16_path;
    uint32_t i16_query = 0;
    bool i16_armed = false;
    void i16_capture();
    void gather_rows(const RowDesc *desc, const std::vector<uint32_t> &ids, std::vector<uint64_t> &offsets,
                     std::vector<uint32_t> &out);
    // Cache-candidate trace: the admitted prefix of the last hotness ranking.
    std::ofstream f1;
    bool f1_on = false, f1_init = false;
    std::vector<uint32_t> rank_ids, rank_deg, f1_prev_ids, f1_prev_deg;
    uint64_t rank_cap = 0, f1_prev_edges = 0;
    void f1_open(const char *path);
    void f1_record(uint32_t batch, dcg_trace_stats *out);
    void drop_graphs();
    // PageRank
    DevBuf<float> d_rank, d_res, d_delta;
    DevBuf<uint32_t> d_queued;
    uint32_t *pr_pending = nullptr;    // points at one of d_qa/d_qb/d_qc
    uint64_t pr_pending_count = 0;

    // ---- update stream ----
    // The whole stream stays in engine-owned pinned host memory (the role of
    // the reference's Loader arrays and mapped WEdge buffers); each stage copies
    // its slice into a device batch slot inside the stage (or finds it there,
    // copied ahead on copy_stream by prefetch_records).
    dcg_edge *h_adds = nullptr, *h_dels = nullptr;
    uint64_t n_adds = 0, n_dels = 0;
    struct BatchSlot {
        DevBuf<dcg_edge> buf;
        uint64_t begin = ~0ull, count = 0;   // the slice it holds (~0: none)
        bool pending = false;                // copied on copy_stream; `ready` not yet waited for
        hipEvent_t ready = nullptr;          // end of its copy on copy_stream
        hipEvent_t used = nullptr;           // end of the last work that read it
        bool read = false;                   // handed out since the last mark_batches_used
        uint64_t stamp = 0;                  // fill order (the older slot is refilled first)
    };
    enum { kDels = 0, kAdds = 1 };
    BatchSlot slots[2][2];                   // [deletions | additions][two slots]
    uint64_t slot_clock = 0;
    hipStream_t copy_stream = nullptr;
    // The device records of [begin, begin+count) of one direction, copied in
    // engine-stream order unless a prefetch already holds them; null for 0.
    const dcg_edge *batch_records(int kind, uint64_t begin, uint64_t count);
    // Copy a slice into the other slot on copy_stream (double buffering).
    void prefetch_records(int kind, uint64_t begin, uint64_t count);
    void fill_slot(int kind, BatchSlot &s, uint64_t begin, uint64_t count, hipStream_t on);
    void forget_batches();
    void mark_batches_used();
    void sync_all();   // engine, second and copy streams

    // ---- phase scratch ----
    DevBuf<uint8_t> d_temp;            // scratch of the sort, scan and selection primitives
    DevBuf<uint64_t> d_k64a, d_k64b;   // sorted delete keys (out then in) / relocation targets
    DevBuf<uint64_t> d_k64c;           // unsorted delete keys; sorted free pool
    DevBuf<uint32_t> d_k32a, d_v32a;   // add keys / payloads, sorted (out then in)
    DevBuf<uint32_t> d_k32c, d_v32c;   // add keys / payloads, unsorted
    DevBuf<uint32_t> d_gstart, d_gstart2, d_uniq;
    DevBuf<uint32_t> d_used;           // per-record occurrence counters
    DevBuf<uint32_t> d_big_groups;     // large-row mutation paths
    DevBuf<uint8_t> d_flags;           // group-head flags of a sorted key array
    DevBuf<unsigned long long> d_counters;  // phase counters (see store.hip)
    static constexpr int kCounterSlots = 32;   // pinned mirror of d_counters
    unsigned long long *h_counters = nullptr;
    static constexpr int kTailWords = 3 + 6 * kClasses;   // store tail block (u64 words)
    uint64_t *h_tail = nullptr;                            // pinned copy of it after an add mutation
    unsigned long long *h_fold = nullptr;                  // pinned: cached / host edge reads of a fold, per stage
    int mut_slot[5] = {-1, -1, -1, -1, -1};                // timing slots of the last delete mutation
    int add_slot = -1;                                     // timing slot of the last add mutation

    // cumulative counters
    uint64_t cum_cached_reads = 0, cum_host_reads = 0;

    // per-kernel measurement: HIP events around launches on `stream`
    struct Timed {
        int kernel;
        int ev;
        int ctl;       // Ctl block index whose counters price the launch (-1: mutation)
        double bytes;  // preset bytes (mutation kernels)
        uint64_t host_lines = 0;
        hipStream_t on = nullptr;   // stream the launch and its events are on
    };
    dcg_kernel_stats kstat[DCG_K_NUM] = {};
    std::vector<hipEvent_t> events;   // pairs (start, end)
    std::vector<Timed> timed;
    int time_begin(int kernel, int ctl, double bytes = 0, hipStream_t on = nullptr);   // on: default the engine stream
    void time_end(int slot);
    void collect_timing();
    void collect_timing_lazy();
    // Closures inside a captured stage graph: each replay stamps the device wall
    // clock at a closure's first workgroup start and last workgroup end in its
    // Ctl block; after the stage's wait the host copy (h_ctl) gives its time and
    // its rows and adjacency entries.
    struct GraphTimer {
        int kernel;
        int ctl;
    };
    double wall_khz = 100000.0;   // device wall-clock rate (hipDeviceAttributeWallClockRate)
    std::vector<GraphTimer> graph_timers[2];
    int capture_stage = -1;
    void collect_graph_timers(int stage);

    explicit Engine(const dcg_config &c);
    ~Engine();

    // store.hip
    void build_fences(const uint32_t *keys, const uint32_t *vals, uint64_t m, const uint64_t *rp,
                      const uint32_t *sorted, DevBuf<uint32_t> &fbase, DevBuf<uint32_t> &fence);
    void load_edges_device(uint32_t nn, const uint32_t *d_src, const uint32_t *d_dst, uint64_t m);
    void mutate_delete(const dcg_edge *recs, uint64_t count, dcg_phase_stats *st);
    void mutate_add(const dcg_edge *recs, uint64_t count, dcg_phase_stats *st);
    void export_csr(uint64_t *row_ptr, uint32_t *col, uint64_t *m);
    StoreView view();
    void fold_reads(dcg_phase_stats *st);   // enqueue + synchronise + finish
    void fold_enqueue(int stage = 0);   // the readback goes to the stage's pair in h_fold
    void fold_finish(dcg_phase_stats *st, int stage = 0);
    // the mutations enqueue their kernels and counter readbacks; finish_* read
    // them after the caller's synchronisation
    void finish_delete(uint64_t count, dcg_phase_stats *st);
    void finish_add(uint64_t count, dcg_phase_stats *st);

    // traverse.hip
    void alloc_state();
    void start_min_lattice();
    void delete_phase(uint64_t begin, uint64_t count, dcg_phase_stats *st);
    void add_phase(uint64_t begin, uint64_t count, dcg_phase_stats *st);
    // a stage's launches and readbacks, and its bookkeeping after the wait
    void delete_enqueue(uint64_t begin, uint64_t count);
    void delete_finish(uint64_t count, dcg_phase_stats *st);
    void add_enqueue(uint64_t begin, uint64_t count);
    void add_finish(uint64_t count, dcg_phase_stats *st);
    // both stages of a snapshot with one wait
    void snapshot_step(uint64_t db, uint64_t dc, uint64_t ab, uint64_t ac, dcg_phase_stats *dst,
                       dcg_phase_stats *ast);
    int del_repair_slot = -1, add_repair_slot = -1;   // timing slots of the stages' repairs
    uint32_t coop_blocks(const void *kernel);
    void autotune(dcg_tune_result *res);
    void repair_stage(int stage);   // run_stage, or tune_stage while tuning is armed
    void tune_stage(int stage);

    // pagerank.hip
    void start_pagerank();
    void pagerank_batch(uint64_t ab, uint64_t ac, uint64_t db, uint64_t dc, dcg_phase_stats *st);
    void pagerank_dense(Ctl *c, dcg_phase_stats *st);
    // dense sweeps pay once the frontier holds more than 1/32 of the vertices
    bool dense_pr(uint64_t frontier) const { return frontier * 32 > (uint64_t)n; }
    DevBuf<uint32_t> d_iota;
    DevBuf<uint32_t> d_prd_small;        // dense sweeps: in-rows of at most 64 slots
    DevBuf<uint32_t> d_prd_medium;       // dense sweeps: in-rows of 65..256 slots
    DevBuf<uint2> d_prd_big;             // dense sweeps: (row, first slot) pieces of larger in-rows

    // cache_policy.hip
    void cache_init();
    uint64_t hot_candidates(uint32_t query, uint64_t cap, uint32_t *ids, uint64_t ids_len);
    void fold_window();
    void cache_admit_measured();
    void cache_maintain(dcg_cache_stats *st);
    void cache_prefetch(uint64_t ab, uint64_t ac, uint64_t db, uint64_t dc, dcg_cache_stats *st);
    void cache_apply_plan(const std::vector<uint64_t> &desired, dcg_cache_stats *st);
    std::vector<uint64_t> plan_admission(const std::vector<float> &score, double admit_min, double margin);
    uint64_t move_budget() const;   // admissions per decision
    // per-chunk read counters of the cache policy (64-byte lines), null
    // without a cache
    float *policy_reads() { return cfg.placement == DCG_PLACE_DEVICE || !cache.d_reads.n ? nullptr : cache.d_reads.p; }

    void sync() { DCG_HIP(hipStreamSynchronize(stream)); }
    // Device bytes held: the store, the per-query state, batch scratch.
    void footprint(uint64_t &store, uint64_t &state, uint64_t &work) const {
        store = cache.d_chunk_frame.bytes() + cache.d_frames.bytes() + cache.d_reads.bytes() + cache.d_hist.bytes() +
                cache.d_touch.bytes() + cache.d_rows_last.bytes() + cache.d_score.bytes() + cache.d_load.bytes() +
                d_out.bytes() + d_in.bytes() + d_tail.bytes() + d_hot_out.bytes() + d_hot_in.bytes() +
                d_hot_window.bytes() + d_hot_list.bytes() + d_hot_count.bytes() + d_sorted_out.bytes() +
                d_sorted_in.bytes() + d_fbase_out.bytes() + d_fbase_in.bytes() + d_fence_out.bytes() +
                d_fence_in.bytes() + d_pool.bytes();
        state = d_sources.bytes() + d_qapps.bytes() + d_vs.bytes() + d_cc_dirty.bytes() +
                d_far0.bytes() + d_far1.bytes() + d_fstamp.bytes() + d_thr.bytes() + d_minfar.bytes() +
                d_fdelta.bytes() + d_epoch.bytes() + d_qa.bytes() + d_qb.bytes() + d_qc.bytes() +
                d_aff.bytes() + d_big.bytes() + d_ctl.bytes() + d_rtrace.bytes() + d_rtcur.bytes() + d_step.bytes() + d_rank.bytes() + d_res.bytes() +
                d_delta.bytes() + d_queued.bytes();
        work = d_temp.bytes() + d_k64a.bytes() + d_k64b.bytes() + d_k64c.bytes() + d_k32a.bytes() +
               d_v32a.bytes() + d_k32c.bytes() + d_v32c.bytes() + d_gstart.bytes() + d_gstart2.bytes() +
               d_uniq.bytes() + d_used.bytes() + d_big_groups.bytes() + d_flags.bytes() + d_counters.bytes() +
               d_iota.bytes() + d_prd_small.bytes() + d_prd_medium.bytes() + d_prd_big.bytes() + mlp_bytes();
        for (const auto &kind : slots)
            for (const BatchSlot &b : kind) work += b.buf.bytes();
    }
    uint64_t mlp_bytes() const;   // the learned policy's device buffers (mlp.hip)

    void *temp(size_t bytes) {
        d_temp.ensure(bytes);
        return d_temp.p;
    }
};

// The built-in program table (traverse.hip) and the TravArgs/Ctl layout tag a
// plugin's table must carry.
const dcg_program *builtin_program();
// Throws for a closure whose control block reports a barrier timeout or an overflow.
void check_launch(const Ctl &c, const char *stage);
uint32_t program_layout_tag();

// Random 64-byte line reads of pinned host memory that one chunk load (256 KiB
// streamed) costs in host-link time: ~4.6 us at 56 GB/s against ~320 M lines/s.
constexpr double kLoadLines = 1024.0;

inline double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

inline uint32_t ceil_pow2_cap(uint64_t deg) {
    if (deg == 0) return 0;
    uint64_t c = deg < kMinRowCap ? kMinRowCap : deg;
    c--;
    c |= c >> 1; c |= c >> 2; c |= c >> 4; c |= c >> 8; c |= c >> 16; c |= c >> 32;
    return (uint32_t)(c + 1);
}

}  // namespace dcg
