// traverse.hip -- host orchestration of one snapshot's delete and add stages and
// of the static convergence, over the repair kernels of vertex_kernels.hpp
// (instantiated here for the built-in SSSP/BFS/CC programs and launched through
// a program table, so a plugin's programs run through the same code), plus the
// PageRank residual push and dense sweeps.
#include <algorithm>
#include <cmath>

#include "engine.hpp"
#include "mlp.hpp"
#include "vertex_kernels.hpp"

namespace dcg {

using BuiltinSet = ProgramSet<apps::SSSP, apps::BFS, apps::CC>;
static_assert(DCG_APP_SSSP == 0 && DCG_APP_BFS == 1 && DCG_APP_CC == 2, "built-in program indices are dcg_app");

const dcg_program *builtin_program() {
    static const dcg_program t = ProgramTable<BuiltinSet>::make("sssp+bfs+cc");
    return &t;
}

uint32_t program_layout_tag() { return kTravArgsTag; }

// Launch kernel `k` of the engine's program table on the engine stream.
static void launch_pk(Engine &e, int k, uint32_t grid, uint32_t block, void **args) {
    DCG_HIP(hipLaunchKernel(e.prog->kernels[k], dim3(grid), dim3(block), args, 0, e.stream));
}

// ------------------------------------------------------------- engine-only kernels
__global__ void k_ctl_setup(Ctl *c, const uint32_t *src0, uint32_t add0, uint32_t init1, const uint32_t *aff_src) {
    if (threadIdx.x || blockIdx.x) return;
    c->count[0] = (src0 ? *src0 : 0u) + add0;
    c->count[1] = init1;
    c->count[2] = 0;
    for (int i = 0; i < 3; ++i) {
        c->cursor[i] = 0;
        c->big[i] = 0;
        c->big_cursor[i] = 0;
    }
    c->rounds = 0;
    c->overflow = 0;
    c->aff_count = aff_src ? *aff_src : 0u;
    c->edges = 0;
    c->expansions = 0;
    c->bad = 0;
    c->solo_round = 0;
    c->solo_resume_b = 0;
    c->far_count[0] = c->far_count[1] = 0;
    c->far_par = 0;
    c->advances = 0;
    c->far_scanned = 0;
    c->rmax[0] = c->rmax[1] = 0;
    c->bar_timeout = 0;
    c->bar_arrive.w = c->bar_gen.w = 0;
    for (uint32_t i = 0; i < kBarGroups; ++i) c->bar_group[i].w = 0;
    c->t_first = ~0ull;
    c->t_last = 0;
}

// Bucket thresholds of a push closure: 0 (everything relaxed waits in the far
// pile until the first advance) or INF for queries without buckets.
__global__ void k_far_setup(uint32_t *thr, uint32_t *minfar, const uint32_t *fdelta, uint32_t Q) {
    for (uint32_t q = threadIdx.x; q < Q; q += blockDim.x) {
        thr[q] = fdelta[q] == kInf ? kInf : 0u;
        minfar[q] = kInf;
        minfar[Q + q] = kInf;
    }
}

// ------------------------------------------------------------------- host side
void check_launch(const Ctl &c, const char *stage) {
    if (c.bar_timeout) throw Error(DCG_ERR_HIP, std::string("a grid barrier timed out in ") + stage);
    if (c.overflow) throw Error(DCG_ERR_CAPACITY, std::string("frontier overflow in ") + stage);
}

// `a`'s stamp base (stamp0, or its step slot) is the first stamp of the phase.
// With a list: the listed vertex-query slots.  Without: the slots of each CC
// query stamped in this stage, in the 1,024-slot blocks a CC improvement
// marked; the marks are cleared once every CC query was checked.
static void cc_fix_parents(Engine &e, const TravArgs &a, const uint32_t *list, const Ctl *lc) {
    if (!e.has_par) return;
    TravArgs ta = a;
    if (list) {
        const uint64_t work = (uint64_t)e.n * e.Q;
        uint32_t base = 0;
        void *args[] = {&ta, &list, &lc, &base};
        launch_pk(e, DCG_PK_FIX_PARENTS, grid_for(work, 256, 16384), 256, args);
        return;
    }
    const uint64_t blocks = (e.n >> kDirtyLog2) + 2;   // a query's range spans at most this many blocks
    for (uint32_t q = 0; q < e.Q; ++q)
        if (e.rooted(q)) {
            uint32_t base = q * e.n;
            void *args[] = {&ta, &list, &lc, &base};
            launch_pk(e, DCG_PK_FIX_PARENTS, grid_for(blocks, 1, 16384), 256, args);
        }
    fill<uint8_t>(e.stream, e.d_cc_dirty.p, e.d_cc_dirty.n, (uint8_t)0);
}

uint32_t Engine::coop_blocks(const void *kernel) {
    int per_cu = 0;
    DCG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kPBlock, 0));
    if (per_cu < 1) per_cu = 1;
    const int want = tune.blocks_per_cu ? (int)tune.blocks_per_cu : 1;
    if (per_cu > want) per_cu = want;
    // a smaller grid pays less per grid barrier (fewer arrivals) on graphs
    // whose frontiers never fill the chip
    const uint32_t grid = (uint32_t)(per_cu * cu_count) >> tune.grid_shift;
    return grid ? grid : 1u;
}

// The min-lattice closures come from the engine's program table; PageRank's
// from the built-in instantiation.
static const void *persistent_fn(const Engine &e, uint32_t mode, bool is_static) {
    switch (mode) {
    case M_PUSH: return e.prog->kernels[is_static ? DCG_PK_STATIC : DCG_PK_PUSH];
    case M_PULL: return e.prog->kernels[DCG_PK_PULL];
    case M_INVALIDATE: return e.prog->kernels[DCG_PK_INVALIDATE];
    case M_PR:
        return is_static ? (const void *)&k_persistent<BuiltinSet, M_PR, 1> : (const void *)&k_persistent<BuiltinSet, M_PR>;
    default: return (const void *)&k_persistent<BuiltinSet, M_PR_AMEND>;
    }
}

void launch_persistent(Engine &e, uint32_t mode, TravArgs &a, bool is_static = false) {
    const void *fn = persistent_fn(e, mode, is_static);
    const uint32_t gi = mode + (is_static ? 5u : 0u);   // grid sizes per instantiation
    if (!e.coop_grid[gi]) e.coop_grid[gi] = e.coop_blocks(fn);
    a.rtag = (uint32_t)(a.ctl - e.d_ctl.p) | mode << 8 | (is_static ? DCG_ROUND_STATIC : 0u);
    void *args[] = {&a};
    static const int kid[5] = {DCG_K_PUSH, DCG_K_PULL, DCG_K_INVALIDATE, DCG_K_PR_SOLVE, DCG_K_PR_AMEND};
    const int k = is_static ? DCG_K_STATIC : kid[mode];
    if (mode == M_PUSH && a.far0)
        hipLaunchKernelGGL(k_far_setup, dim3(1), dim3(64), 0, e.stream, a.thr, a.minfar, a.fdelta, a.Q);
    if (e.capturing) {
        // graph node: a plain launch of the same co-resident grid (the grid
        // barrier needs only co-residency; it gives up after ~10 s otherwise);
        // every replay times it by the wall-clock stamps in its control block
        e.graph_timers[e.capture_stage].push_back(Engine::GraphTimer{k, (int)(a.ctl - e.d_ctl.p)});
        DCG_HIP(hipLaunchKernel(fn, dim3(e.coop_grid[gi]), dim3(kPBlock), args, 0, e.stream));
        return;
    }
    mlp_join(e);   // no training kernel beside a closure
    const int slot = e.time_begin(k, (int)(a.ctl - e.d_ctl.p));
    const hipError_t err = e.coop_plain ? hipErrorIllegalState
                                        : hipLaunchCooperativeKernel(fn, dim3(e.coop_grid[gi]), dim3(kPBlock), args,
                                                                     0, e.stream);
    if (err == hipErrorIllegalState) {
        // the device's cooperative-launch resource is held by another process
        // (one that launched cooperatively before): launch the same
        // occupancy-sized grid plainly, as inside graphs, from now on
        (void)hipGetLastError();
        e.coop_plain = true;
        DCG_HIP(hipLaunchKernel(fn, dim3(e.coop_grid[gi]), dim3(kPBlock), args, 0, e.stream));
    } else {
        DCG_HIP(err);
    }
    e.time_end(slot);
}

// Stamps: each launch owns a window of 2^20 round stamps; the slots' stamps are
// cleared when the 32-bit stamp space would wrap.
__global__ void k_clear_ticks(VSlot *vs, uint64_t count) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
        vs[i].tick = 0;
}

static void clear_ticks(Engine &e) {
    const uint64_t count = (uint64_t)e.n * e.Q;
    if (count) hipLaunchKernelGGL(k_clear_ticks, dim3(grid_for(count)), dim3(256), 0, e.stream, e.d_vs.p, count);
}

// Before a phase: keep `windows` stamp windows ahead without a wrap, so every
// stamp the phase takes is above the first one (the CC parent check below
// selects the phase's vertices by stamp).
static void reserve_stamps(Engine &e, uint32_t windows) {
    constexpr uint64_t kWindow = 1ull << 20;
    if ((uint64_t)e.stamp + (uint64_t)(windows + 2) * kWindow > 0xFFFFFFFFull) {
        clear_ticks(e);
        e.stamp = 1;
    }
}

static uint32_t take_stamps(Engine &e) {
    constexpr uint32_t kWindow = 1u << 20;
    if (e.stamp > 0xFFFFFFFFu - 2 * kWindow) {
        clear_ticks(e);
        e.stamp = 1;
    }
    const uint32_t s = e.stamp;
    e.stamp += kWindow;
    return s;
}

static TravArgs base_args(Engine &e) {
    TravArgs a{};
    a.s = e.view();
    a.n = e.n;
    a.Q = e.Q;
    a.vs = e.d_vs.p;
    a.cc_dirty = e.d_cc_dirty.p;
    a.sources = e.d_sources.p;
    a.qapp = e.d_qapps.p;
    a.w = e.wp;
    a.app = e.app;
    a.max_rounds = 1u << 20;
    a.qa = e.d_qa.p;
    a.qb = e.d_qb.p;
    a.qcap = e.d_qa.n;
    a.aff = e.d_aff.p;
    a.affcap = e.d_aff.n;
    a.big = e.d_big.p;
    a.bigcap = e.d_big.n;
    a.rank = e.d_rank.p;
    a.res = e.d_res.p;
    a.delta = e.d_delta.p;
    a.queued = e.d_queued.p;
    a.eps = (float)e.cfg.pr_epsilon;
    a.sign = 1.f;
    a.big_row = e.tune.big_row;
    a.piece = e.tune.piece;
    a.solo_max = e.tune.solo_max;
    a.bar_spin = e.bar_spin;
    if (e.d_rtrace.p) {
        a.rtrace = e.d_rtrace.p;
        a.rtcur = e.d_rtcur.p;
        a.rtcap = (uint32_t)e.d_rtrace.n;
    }
    if (e.cx.active) {
        a.owner = e.cx.d_owner.p;
        a.bprop = e.cx.d_bprop.p;
        a.blist = e.cx.d_blist.p;
        a.bcount = e.cx.d_bcount.p;
        a.bcap = (uint32_t)e.cx.d_blist.n;
    }
    if (e.bucket_width() && e.d_far0.p) {
        a.far0 = e.d_far0.p;
        a.far1 = e.d_far1.p;
        a.fstamp = e.d_fstamp.p;
        a.thr = e.d_thr.p;
        a.minfar = e.d_minfar.p;
        a.fdelta = e.d_fdelta.p;
        a.epoch = e.d_epoch.p;
    }
    return a;
}

// Per-query bucket widths of the push closure: SSSP w_max * q / 4 (the
// reference's ordered repair uses 128 for its (u+v)%128+1 weights,
// ordered_gpu_repair.cuh:116-160), BFS q levels, CC none.
void Engine::update_buckets() {
    if (!d_fdelta.p) return;
    const uint32_t wmax = wp.kind == W_SEEDED ? wp.wmax : 128u;
    const uint32_t qt = bucket_width();
    std::vector<uint32_t> h(Q);
    for (uint32_t q = 0; q < Q; ++q)
        h[q] = rooted(q) || !qt ? kInf : prog->bucket_width(qapps[q], qt, wmax);
    DCG_HIP(hipMemcpyAsync(d_fdelta.p, h.data(), Q * 4, hipMemcpyHostToDevice, stream));
    sync();
}

void Engine::alloc_state() {
    drop_graphs();
    const uint64_t qn = (uint64_t)n * Q;
    if (qn >= 0xFFFFFFFFull) throw Error(DCG_ERR_UNSUPPORTED, "num_queries * num_nodes must be below 2^32");
    const uint64_t cap = qn ? qn : 1;
    d_qa.alloc(cap);
    d_qb.alloc(cap);
    d_qc.alloc(cap);
    d_big.alloc(cap < (1ull << 26) ? cap : (1ull << 26));
    d_ctl.alloc(8);
    d_vs.alloc(cap);
    fill<VSlot>(stream, d_vs.p, cap, VSlot{~0ull, 0u, kInf});
    has_par = false;
    if (app == DCG_APP_PR) {
        d_rank.alloc(cap);
        d_res.alloc(cap);
        d_delta.alloc(cap);
        d_queued.alloc(cap);
    } else {
        d_aff.alloc(cap);
        bool cc = false, ordered = false;
        for (uint32_t q = 0; q < Q; ++q) (rooted(q) ? cc : ordered) = true;
        if (cc) {
            has_par = true;
            const uint64_t blocks = (cap >> kDirtyLog2) + 1;
            d_cc_dirty.alloc(blocks);
            fill<uint8_t>(stream, d_cc_dirty.p, blocks, (uint8_t)0);
        }
        if (ordered) {
            d_far0.alloc(cap);
            d_far1.alloc(cap);
            d_fstamp.alloc(cap);
            fill<uint32_t>(stream, d_fstamp.p, cap, 0u);
            d_thr.alloc(Q);
            d_minfar.alloc(2ull * Q);
            d_fdelta.alloc(Q);
            d_epoch.alloc(1);
            fill<uint32_t>(stream, d_epoch.p, 1, 1u);
            update_buckets();
        }
    }
}

// Autotuner: coordinate descent over (workgroups per CU, large-row threshold,
// piece size), each candidate timed on a from-scratch closure of the loaded
// graph (HIP events around the launch).  Every candidate yields the same values.
void Engine::autotune(dcg_tune_result *res) {
    if (app == DCG_APP_PR) throw Error(DCG_ERR_UNSUPPORTED, "the autotuner tunes the min-lattice closure");
    auto measure = [&]() {
        for (auto &g : coop_grid) g = 0;
        drop_graphs();
        start_min_lattice();
        sync();
        collect_timing();
        const double ms = kstat[DCG_K_STATIC].ms;
        kstat[DCG_K_STATIC] = dcg_kernel_stats{};
        return ms;
    };
    const dcg_kernel_stats keep = kstat[DCG_K_STATIC];
    kstat[DCG_K_STATIC] = dcg_kernel_stats{};
    forget_stage_tunes();
    double best = measure();
    const double first = best;
    uint32_t evals = 1;
    const uint32_t bpc[] = {1, 2};
    const uint32_t rows[] = {64, 128, 256, 512, 1024, 4096};
    const uint32_t pieces[] = {256, 512, 1024, 2048, 4096};
    const uint32_t quarters[] = {0, 1, 4, 16, 64};
    tune.bucket_quarters = bucket_width();   // tune from the width in use
    for (int pass = 0; pass < 2; ++pass) {
        if (d_fdelta.p)
            for (uint32_t v : quarters) {
                if (v == tune.bucket_quarters) continue;
                const uint32_t old = tune.bucket_quarters;
                tune.bucket_quarters = v;
                update_buckets();
                const double t = measure();
                evals++;
                if (t < best) {
                    best = t;
                } else {
                    tune.bucket_quarters = old;
                    update_buckets();
                }
            }
        for (uint32_t v : bpc) {
            if (v == tune.blocks_per_cu) continue;
            const uint32_t old = tune.blocks_per_cu;
            tune.blocks_per_cu = v;
            const double t = measure();
            evals++;
            if (t < best) best = t;
            else tune.blocks_per_cu = old;
        }
        for (uint32_t v : rows) {
            if (v == tune.big_row) continue;
            const uint32_t old = tune.big_row;
            tune.big_row = v;
            const double t = measure();
            evals++;
            if (t < best) best = t;
            else tune.big_row = old;
        }
        for (uint32_t v : pieces) {
            if (v == tune.piece) continue;
            const uint32_t old = tune.piece;
            tune.piece = v;
            const double t = measure();
            evals++;
            if (t < best) best = t;
            else tune.piece = old;
        }
    }
    for (auto &g : coop_grid) g = 0;
    start_min_lattice();   // leave the converged state behind
    sync();
    collect_timing();
    kstat[DCG_K_STATIC] = keep;
    if (res) {
        res->blocks_per_cu = tune.blocks_per_cu;
        res->big_row = tune.big_row;
        res->piece = tune.piece;
        res->evaluations = evals;
        res->best_ms = best;
        res->bucket_quarters = tune.bucket_quarters;
        res->solo_max = tune.solo_max;
        res->first_ms = first;
        res->grid_shift = tune.grid_shift;
    }
}

void Engine::start_min_lattice() {
    const uint64_t qn = (uint64_t)n * Q;
    TravArgs ia = base_args(*this);
    {
        void *args[] = {&ia};
        launch_pk(*this, DCG_PK_INIT, grid_for(qn), 256, args);
    }
    stamp = 1;
    std::vector<uint64_t> slot0(Q);
    uint64_t count0 = 0;
    for (uint32_t q = 0; q < Q; ++q) {
        slot0[q] = count0;
        count0 += vertex_seeds(q) ? n : 1;
    }
    DevBuf<uint64_t> d_slot0;
    d_slot0.alloc(Q);
    DCG_HIP(hipMemcpyAsync(d_slot0.p, slot0.data(), Q * 8, hipMemcpyHostToDevice, stream));
    {
        uint32_t *q0 = d_qc.p;
        const uint64_t *s0 = d_slot0.p;
        void *args[] = {&ia, &q0, &s0};
        launch_pk(*this, DCG_PK_FRONTIER, grid_for(n), 256, args);
    }
    Ctl *c = d_ctl.p;
    hipLaunchKernelGGL(k_ctl_setup, dim3(1), dim3(64), 0, stream, c, (const uint32_t *)nullptr, (uint32_t)count0, 0u,
                       (const uint32_t *)nullptr);
    TravArgs a = base_args(*this);
    a.in0 = d_qc.p;
    a.ctl = c;
    a.stamp0 = take_stamps(*this);
    launch_persistent(*this, M_PUSH, a, true);
    cc_fix_parents(*this, a, nullptr, nullptr);
    DCG_HIP(hipGetLastError());
    DCG_HIP(hipMemcpyAsync(h_ctl, c, sizeof(Ctl), hipMemcpyDeviceToHost, stream));
    sync();
    check_launch(h_ctl[0], "static convergence");
    fold_reads(nullptr);
}

// Algorithmic bytes of a stage's closures (control blocks [lo, hi)): 24 B per
// expanded row, 12 B per adjacency entry.
static double stage_bytes(const Ctl *c, int lo, int hi) {
    double b = 0;
    for (int i = lo; i < hi; ++i) b += 24.0 * (double)c[i].expansions + 12.0 * (double)c[i].edges;
    return b;
}

// Delete stage of one snapshot (framework.cuh:4811-4957, re-ordered: the
// physical deletion runs first; the invalidation then scans post-delete rows,
// which only drops deleted edges that the seed step already examined).
// Enqueued in one call and finished after the stream's wait in another, so a
// snapshot step can enqueue both stages before it waits once.
void Engine::delete_enqueue(uint64_t begin, uint64_t count) {
    const dcg_edge *recs = batch_records(kDels, begin, count);
    mlp_release_training(*this);   // the learned policy trains beside the delete mutation
    mutate_delete(recs, count, nullptr);
    if (!count) return;
    reserve_stamps(*this, 2);
    h_step[0] = StepArgs{recs, count, {take_stamps(*this), take_stamps(*this), 0u, 0u}};
    DCG_HIP(hipMemcpyAsync(d_step.p, h_step, sizeof(StepArgs), hipMemcpyHostToDevice, stream));
    const int rs = time_begin(DCG_K_REPAIR_DELETE, -1);
    del_repair_slot = rs;
    if (i16_armed) i16_capture();   // the repair in two parts around the snapshot's captures
    else repair_stage(0);
    time_end(rs);
    mark_batches_used();
    fold_enqueue(0);
    DCG_HIP(hipGetLastError());
    DCG_HIP(hipMemcpyAsync(h_ctl, d_ctl.p, 4 * sizeof(Ctl), hipMemcpyDeviceToHost, stream));
}

void Engine::delete_finish(uint64_t count, dcg_phase_stats *st) {
    if (!count) return;
    finish_delete(count, st);
    for (int i = 0; i < 4; ++i)
        check_launch(h_ctl[i], "the delete stage");
    timed[del_repair_slot].bytes = stage_bytes(h_ctl, 1, 4);
    if (graph_exec[0]) collect_graph_timers(0);
    if (st) {
        st->affected = h_ctl[1].aff_count;
        st->rounds = h_ctl[1].rounds + h_ctl[3].rounds;
        st->bucket_advances = h_ctl[3].advances;
        st->far_entries = h_ctl[3].far_scanned;
        st->edges_scanned = h_ctl[1].edges + h_ctl[2].edges + h_ctl[3].edges;
    }
    fold_finish(st, 0);
}

void Engine::delete_phase(uint64_t begin, uint64_t count, dcg_phase_stats *st) {
    delete_enqueue(begin, count);
    if (!count) return;
    sync();   // one wait per stage: mutation counters, repair control blocks, read fold
    delete_finish(count, st);
}

// Add stage (framework.cuh:3525-3756): append, seed from the added edges, close.
void Engine::add_enqueue(uint64_t begin, uint64_t count) {
    const dcg_edge *recs = batch_records(kAdds, begin, count);
    mutate_add(recs, count, nullptr);
    if (!count) return;
    reserve_stamps(*this, 2);
    h_step[1] = StepArgs{recs, count, {take_stamps(*this), take_stamps(*this), 0u, 0u}};
    DCG_HIP(hipMemcpyAsync(d_step.p + 1, h_step + 1, sizeof(StepArgs), hipMemcpyHostToDevice, stream));
    const int rs = time_begin(DCG_K_REPAIR_ADD, -1);
    add_repair_slot = rs;
    if (cx.on) coexec_add_stage();   // the closure shared with host threads
    else repair_stage(1);
    time_end(rs);
    mark_batches_used();
    fold_enqueue(1);
    DCG_HIP(hipGetLastError());
    DCG_HIP(hipMemcpyAsync(h_ctl + 4, d_ctl.p + 4, 2 * sizeof(Ctl), hipMemcpyDeviceToHost, stream));
}

void Engine::add_finish(uint64_t count, dcg_phase_stats *st) {
    if (!count) return;
    finish_add(count, st);
    check_launch(h_ctl[4], "the add stage");
    check_launch(h_ctl[5], "the add stage");
    timed[add_repair_slot].bytes = stage_bytes(h_ctl, 5, 6);
    if (graph_exec[1]) collect_graph_timers(1);
    if (st && cx.on) {
        st->coexec_rounds = cx.rounds;
        st->cpu_relaxations = cx.relax;
        st->proposals_to_cpu = cx.to_cpu;
        st->proposals_to_gpu = cx.to_gpu;
    }
    if (st) {
        st->rounds = h_ctl[5].rounds;
        st->bucket_advances = h_ctl[5].advances;
        st->far_entries = h_ctl[5].far_scanned;
        st->edges_scanned = h_ctl[5].edges;
    }
    fold_finish(st, 1);
}

void Engine::add_phase(uint64_t begin, uint64_t count, dcg_phase_stats *st) {
    add_enqueue(begin, count);
    if (!count) return;
    sync();   // one wait per stage: mutation counters, repair control blocks, read fold
    add_finish(count, st);
}

// One snapshot step: the delete stage and the add stage enqueued back to back
// and finished after a single wait (the two stages use separate control
// blocks, step arguments and host mirrors).  The co-executed add stage and the
// repair-snapshot capture wait on the host inside a stage; with either on the
// stages run one after the other.
void Engine::snapshot_step(uint64_t db, uint64_t dc, uint64_t ab, uint64_t ac, dcg_phase_stats *dst,
                           dcg_phase_stats *ast) {
    if (cx.on || i16_armed) {
        delete_phase(db, dc, dst);
        add_phase(ab, ac, ast);
        return;
    }
    delete_enqueue(db, dc);
    add_enqueue(ab, ac);
    sync();
    delete_finish(dc, dst);
    add_finish(ac, ast);
}

// The repair launches of a stage, reading the batch and the round stamps from
// d_step[stage].  Delete (framework.cuh:4811-4957, re-ordered: the physical
// deletion has run; the invalidation scans post-delete rows, which only drops
// deleted edges that the seed step already examined): seeds -> invalidation
// closure -> reset -> one pull pass -> push closure.  Add: seeds -> push closure.
void Engine::enqueue_repair(int stage, int part) {
    Ctl *c = d_ctl.p;
    TravArgs a = base_args(*this);
    a.step = d_step.p + stage;
    constexpr uint32_t kSeedBlocks = 512;   // 1,024-thread workgroups, grid-stride over count * Q records
    if (stage == 0) {
        if (part != 2) {
            hipLaunchKernelGGL(k_ctl_setup, dim3(1), dim3(64), 0, stream, c + 0, (const uint32_t *)nullptr, 0u, 0u,
                               (const uint32_t *)nullptr);
            a.ctl = c + 0;
            {
                void *args[] = {&a};
                launch_pk(*this, DCG_PK_SEED_DELS, kSeedBlocks, 1024, args);
            }
            hipLaunchKernelGGL(k_ctl_setup, dim3(1), dim3(64), 0, stream, c + 1, &c[0].aff_count, 0u, 0u,
                               &c[0].aff_count);
            a.ctl = c + 1;
            a.in0 = d_aff.p;
            a.stamp_slot = 0;
            launch_persistent(*this, M_INVALIDATE, a);
            {
                void *args[] = {&a};   // a.ctl = c + 1: the affected list's length
                launch_pk(*this, DCG_PK_RESET, grid_for((uint64_t)n * Q, 256, 8192), 256, args);
            }
            if (part == 1) return;
        }
        a.in0 = d_aff.p;
        a.stamp_slot = 0;
        hipLaunchKernelGGL(k_ctl_setup, dim3(1), dim3(64), 0, stream, c + 2, &c[1].aff_count, 0u, 0u,
                           (const uint32_t *)nullptr);
        a.ctl = c + 2;
        a.max_rounds = 1;
        launch_persistent(*this, M_PULL, a);
        hipLaunchKernelGGL(k_ctl_setup, dim3(1), dim3(64), 0, stream, c + 3, &c[1].aff_count, 0u, 0u,
                           (const uint32_t *)nullptr);
        a.ctl = c + 3;
        a.max_rounds = 1u << 20;
        a.stamp_slot = 1;
        launch_persistent(*this, M_PUSH, a);
        cc_fix_parents(*this, a, d_aff.p, c + 1);
        cc_fix_parents(*this, a, nullptr, nullptr);
    } else {
        hipLaunchKernelGGL(k_ctl_setup, dim3(1), dim3(64), 0, stream, c + 4, (const uint32_t *)nullptr, 0u, 0u,
                           (const uint32_t *)nullptr);
        a.ctl = c + 4;
        a.stamp_slot = 0;
        {
            uint32_t *outq = d_qc.p;
            void *args[] = {&a, &outq};
            launch_pk(*this, DCG_PK_SEED_ADDS, kSeedBlocks, 1024, args);
        }
        hipLaunchKernelGGL(k_ctl_setup, dim3(1), dim3(64), 0, stream, c + 5, &c[4].aff_count, 0u, 0u,
                           (const uint32_t *)nullptr);
        a.ctl = c + 5;
        a.in0 = d_qc.p;
        a.stamp_slot = 1;
        launch_persistent(*this, M_PUSH, a);
        a.stamp_slot = 0;   // the parent check covers the seeds' round too
        cc_fix_parents(*this, a, nullptr, nullptr);
    }
}

// Capture the stage's repair launches as a graph unless one is ready; false
// when capture failed (graphs are then off and the launches go direct).
bool Engine::build_stage_graph(int stage) {
    if (!tune.graphs) return false;
    if (!graph_exec[stage]) {
        hipGraph_t g = nullptr;
        bool ok = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (ok) {
            capturing = true;
            capture_stage = stage;
            try {
                enqueue_repair(stage);
            } catch (...) {
                ok = false;
            }
            capturing = false;
            ok = hipStreamEndCapture(stream, &g) == hipSuccess && ok;
        }
        if (ok) ok = hipGraphInstantiate(&graph_exec[stage], g, nullptr, nullptr, 0) == hipSuccess;
        if (g) (void)hipGraphDestroy(g);
        if (!ok) {
            (void)hipGetLastError();
            graph_exec[stage] = nullptr;
            tune.graphs = 0;
            return false;
        }
    }
    return true;
}

// The stage's own closure parameters (when the stage tuner set them) for the
// scope's launches and captures; `tune` is restored at the end of the scope.
struct StageTuneScope {
    Engine &e;
    Engine::Tune saved;
    bool on, grid;
    StageTuneScope(Engine &en, int stage) : e(en), saved(en.tune), on(en.stage_tuned[stage]), grid(false) {
        if (!on) return;
        const Engine::Tune &s = e.stage_tune[stage];
        grid = s.blocks_per_cu != e.tune.blocks_per_cu || s.grid_shift != e.tune.grid_shift;
        e.tune.blocks_per_cu = s.blocks_per_cu;
        e.tune.grid_shift = s.grid_shift;
        e.tune.solo_max = s.solo_max;
        e.tune.big_row = s.big_row;
        e.tune.piece = s.piece;
        if (grid)
            for (auto &g : e.coop_grid) g = 0;
    }
    ~StageTuneScope() {
        if (!on) return;
        const uint32_t graphs = e.tune.graphs;   // a failed capture turned graphs off
        e.tune = saved;
        e.tune.graphs = graphs;
        if (grid)
            for (auto &g : e.coop_grid) g = 0;
    }
};

// Run a stage's repair launches: replay its captured graph (captured at the
// first call; a capture the runtime rejects turns graphs off for the engine and
// the launches are issued directly), or issue them directly.
void Engine::run_stage(int stage) {
    mlp_join(*this);   // before the stage's closures (and outside any capture)
    StageTuneScope scope(*this, stage);
    if (!build_stage_graph(stage)) {
        enqueue_repair(stage);
        return;
    }
    DCG_HIP(hipGraphLaunch(graph_exec[stage], stream));
}

// Co-executed add stage (coexec.hip drives it): the stage's seeds and first
// closure, issued directly (the owner arguments vary per stage) ...
void Engine::enqueue_coexec_seeds() { enqueue_repair(1); }

// ... and one round of the host's proposals: applied with fresh round stamps
// into the seed queue, then a push closure from the improved slots.
void Engine::enqueue_coexec_round(const std::vector<uint32_t> &gs, const std::vector<uint64_t> &packed) {
    reserve_stamps(*this, 2);
    h_step[1].stamp[0] = take_stamps(*this);
    h_step[1].stamp[1] = take_stamps(*this);
    DCG_HIP(hipMemcpyAsync(d_step.p + 1, h_step + 1, sizeof(StepArgs), hipMemcpyHostToDevice, stream));
    Ctl *c = d_ctl.p;
    hipLaunchKernelGGL(k_ctl_setup, dim3(1), dim3(64), 0, stream, c + 4, (const uint32_t *)nullptr, 0u, 0u,
                       (const uint32_t *)nullptr);
    coexec_apply(gs, packed, h_step[1].stamp[0], c + 4);
    hipLaunchKernelGGL(k_ctl_setup, dim3(1), dim3(64), 0, stream, c + 5, &c[4].aff_count, 0u, 0u,
                       (const uint32_t *)nullptr);
    TravArgs a = base_args(*this);
    a.step = d_step.p + 1;
    a.ctl = c + 5;
    a.in0 = d_qc.p;
    a.stamp_slot = 1;
    launch_persistent(*this, M_PUSH, a);
    DCG_HIP(hipGetLastError());
}

void Engine::repair_stage(int stage) {
    if (!tune_arm) {
        run_stage(stage);
        return;
    }
    --tune_arm;
    tune_stage(stage);
}

// Stage autotuner: the repair of this stage, run per candidate from the same
// post-mutation state (values, CC parents), each run with fresh round stamps
// and fresh round stamps, timed with HIP events around the replay of the stage's
// graph (captured for each parameter set before the clock starts).  A
// candidate's time is the minimum over its runs (as many as fit ~5 ms, at most
// 5); coordinate descent keeps a candidate only if it beats the best time so far
// by 3%, so timing noise does not move the parameters.  Every run reaches the
// same values (the fixed point does not depend on the parameters), so the state
// the last run leaves is the stage's result.
void Engine::tune_stage(int stage) {
    // candidates are set in `tune` itself: the stage's own parameters (if any)
    // are the starting point, and the choice becomes the stage's parameters
    const Tune global = tune;
    if (stage_tuned[stage]) {
        const Tune &s = stage_tune[stage];
        tune.blocks_per_cu = s.blocks_per_cu;
        tune.grid_shift = s.grid_shift;
        tune.solo_max = s.solo_max;
        tune.big_row = s.big_row;
        tune.piece = s.piece;
        for (auto &g : coop_grid) g = 0;
        drop_graphs();
    }
    stage_tuned[stage] = false;
    const uint64_t qn = (uint64_t)n * Q;
    DevBuf<VSlot> vs0;   // state words, stamps and parents after the mutation
    vs0.alloc(qn);
    DCG_HIP(hipMemcpyAsync(vs0.p, d_vs.p, qn * sizeof(VSlot), hipMemcpyDeviceToDevice, stream));
    hipEvent_t e0, e1;
    DCG_HIP(hipEventCreate(&e0));
    DCG_HIP(hipEventCreate(&e1));
    bool restore = false;
    auto run = [&]() {
        if (restore) {
            DCG_HIP(hipMemcpyAsync(d_vs.p, vs0.p, qn * sizeof(VSlot), hipMemcpyDeviceToDevice, stream));
        }
        restore = true;
        reserve_stamps(*this, 2);
        h_step[stage].stamp[0] = take_stamps(*this);
        h_step[stage].stamp[1] = take_stamps(*this);
        DCG_HIP(hipMemcpyAsync(d_step.p + stage, h_step + stage, sizeof(StepArgs), hipMemcpyHostToDevice, stream));
        // the graph of the current parameters is captured before the clock
        // starts: a candidate's time is its launches' device time only
        mlp_join(*this);
        build_stage_graph(stage);
        DCG_HIP(hipEventRecord(e0, stream));
        run_stage(stage);
        DCG_HIP(hipEventRecord(e1, stream));
        DCG_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        DCG_HIP(hipEventElapsedTime(&ms, e0, e1));
        DCG_HIP(hipMemcpy(h_ctl, d_ctl.p, 6 * sizeof(Ctl), hipMemcpyDeviceToHost));
        const int lo = stage == 0 ? 0 : 4, hi = stage == 0 ? 4 : 6;
        for (int i = lo; i < hi; ++i) check_launch(h_ctl[i], "a tuned stage");
        return (double)ms;
    };
    const bool ordered = d_fdelta.p != nullptr;
    if (ordered && tune.bucket_quarters == ~0u) tune.bucket_quarters = bucket_width();
    (void)run();   // the first run pays the cold start (caches, code objects); not compared
    double best = run();
    const int reps = best > 0 ? (int)std::min(5.0, std::max(1.0, std::ceil(5.0 / best))) : 5;
    for (int i = 1; i < reps; ++i) best = std::min(best, run());
    const double first = best;
    uint32_t evals = 1;
    auto timed_runs = [&]() {
        double t = run();
        for (int i = 1; i < reps; ++i) t = std::min(t, run());
        return t;
    };
    // a parameter change needs new grid sizes and a new graph
    auto params_changed = [&](bool buckets) {
        if (buckets) update_buckets();
        for (auto &g : coop_grid) g = 0;
        drop_graphs();
    };
    // one coordinate at a time: try its candidates, keep the best, move on
    auto axis = [&](uint32_t &param, std::initializer_list<uint32_t> values, bool buckets) {
        for (uint32_t v : values) {
            if (v == param) continue;
            const uint32_t old = param;
            param = v;
            params_changed(buckets);
            const double t = timed_runs();
            ++evals;
            if (t < 0.97 * best) {
                best = t;
            } else {
                param = old;
                params_changed(buckets);
            }
        }
    };
    const Tune start = tune;
    if (ordered) axis(tune.bucket_quarters, {4, 16, 64, 256}, true);
    axis(tune.blocks_per_cu, {1, 2}, false);
    axis(tune.grid_shift, {1, 2, 3, 5, 8}, false);
    axis(tune.solo_max, {0, 512, 2048, 8192}, false);
    axis(tune.big_row, {64, 128, 256, 1024, 4096}, false);
    axis(tune.piece, {64, 128, 512, 2048}, false);
    // Coordinate descent compares single runs on long stages, where one run
    // can differ from the next by more than the 3% margin: the choice is kept
    // only if it still beats the stage's starting parameters in three
    // interleaved runs each (best of three).
    const bool changed = start.bucket_quarters != tune.bucket_quarters || start.blocks_per_cu != tune.blocks_per_cu ||
                         start.grid_shift != tune.grid_shift || start.solo_max != tune.solo_max ||
                         start.big_row != tune.big_row || start.piece != tune.piece;
    if (changed) {
        const Tune chosen = tune;
        Tune from = start;
        from.graphs = tune.graphs;   // graphs stay off if a capture failed meanwhile
        double t_start = 1e300, t_chosen = 1e300;
        for (int i = 0; i < 3; ++i) {
            tune = from;
            params_changed(ordered);
            t_start = std::min(t_start, run());
            tune = chosen;
            params_changed(ordered);
            t_chosen = std::min(t_chosen, run());
        }
        ++evals;
        if (t_chosen < 0.97 * t_start) {
            best = std::min(best, t_chosen);
        } else {
            tune = from;
            params_changed(ordered);
            best = std::min(first, t_start);
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    stage_tune[stage] = tune;
    stage_tuned[stage] = true;
    // the other closures keep the engine-wide parameters; the bucket width
    // (device memory) and a graphs-off fallback stay as the tuner left them
    const uint32_t quarters = tune.bucket_quarters, graphs = tune.graphs;
    tune = global;
    tune.bucket_quarters = quarters;
    tune.graphs = graphs;
    for (auto &g : coop_grid) g = 0;
    drop_graphs();   // the next stage captures its graph with the chosen parameters
    const Tune &chosen = stage_tune[stage];
    last_tune = dcg_tune_result{};
    last_tune.blocks_per_cu = chosen.blocks_per_cu;
    last_tune.big_row = chosen.big_row;
    last_tune.piece = chosen.piece;
    last_tune.evaluations = evals;
    last_tune.best_ms = best;
    last_tune.bucket_quarters = ordered ? chosen.bucket_quarters : 0u;
    last_tune.solo_max = chosen.solo_max;
    last_tune.first_ms = first;
    last_tune.grid_shift = chosen.grid_shift;
    stage_result[stage] = last_tune;
}

void Engine::drop_graphs() {
    for (auto &g : graph_exec) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
    for (auto &list : graph_timers) list.clear();
}

// ------------------------------------------------------------------ PageRank
__global__ void k_pr_init(float *rank, float *res, uint32_t *queued, uint32_t *q0, uint32_t n, float base) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        rank[v] = 0.f;
        res[v] = base;
        queued[v] = 1;
        q0[v] = (uint32_t)v;
    }
}

// Unique sources of one PR batch (adds and deletes), deduplicated with stamps.
__global__ void k_pr_sources(const dcg_edge *adds, uint64_t na, const dcg_edge *dels, uint64_t nd, uint32_t n,
                             VSlot *vs, uint32_t stamp, uint32_t *out, Ctl *c) {
    const uint64_t total = na + nd;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t b = blockIdx.x * (uint64_t)blockDim.x; b < total; b += stride) {
        const uint64_t i = b + threadIdx.x;
        bool want = false;
        uint32_t u = 0;
        if (i < total) {
            const dcg_edge e = i < na ? adds[i] : dels[i - na];
            u = e.u;
            if (e.u < n && e.v < n) want = atomicExch(&vs[u].tick, stamp) != stamp;
        }
        const uint64_t mask = __ballot(want);
        if (mask) {
            const int leader = __ffsll((unsigned long long)mask) - 1;
            uint32_t base = 0;
            if ((int)lane_id() == leader) base = atomicAdd(&c->aff_count, (uint32_t)__popcll(mask));
            base = __shfl(base, leader, kWave);
            if (want) out[base + lanes_below(mask)] = u;
        }
    }
}

void Engine::start_pagerank() {
    const float base = (float)cfg.pr_base;
    hipLaunchKernelGGL(k_pr_init, dim3(grid_for(n)), dim3(256), 0, stream, d_rank.p, d_res.p, d_queued.p, d_qc.p, n,
                       base);
    pr_pending = d_qc.p;
    pr_pending_count = n;
    Ctl *c = d_ctl.p;
    if (dense_pr(pr_pending_count)) {
        pagerank_dense(c + 6, nullptr);
        fold_reads(nullptr);
        return;
    }
    hipLaunchKernelGGL(k_ctl_setup, dim3(1), dim3(64), 0, stream, c + 6, (const uint32_t *)nullptr,
                       (uint32_t)pr_pending_count, 0u, (const uint32_t *)nullptr);
    TravArgs a = base_args(*this);
    // converge: input = pending (qc), rounds alternate qa/qb
    a.ctl = c + 6;
    a.in0 = pr_pending;
    a.max_rounds = cfg.pr_max_rounds;
    launch_persistent(*this, M_PR, a, true);
    DCG_HIP(hipGetLastError());
    DCG_HIP(hipMemcpyAsync(h_ctl + 6, c + 6, sizeof(Ctl), hipMemcpyDeviceToHost, stream));
    sync();
    check_launch(h_ctl[6], "PageRank");
    if (h_ctl[6].bad) throw Error(DCG_ERR_NUMERIC, "PR non-finite rank/residual");
    const uint32_t r = h_ctl[6].rounds;
    // the unfinished frontier (if the round cap was hit) is the output of round r-1
    if (r == 0) {
        pr_pending_count = h_ctl[6].count[0];
    } else {
        pr_pending = ((r - 1) & 1) ? d_qb.p : d_qa.p;
        pr_pending_count = h_ctl[6].count[r % 3];
    }
    fold_reads(nullptr);
}

// After dense sweeps stopped at the round cap: the vertices still above eps
// become the pending frontier (queued, appended to `out`, count in ctl->aff_count).
__global__ void k_pr_pending(const float *res, uint32_t *queued, uint32_t n, float eps, uint32_t *out, Ctl *c) {
    for (uint32_t b = blockIdx.x * blockDim.x; b < n; b += gridDim.x * blockDim.x) {
        const uint32_t v = b + threadIdx.x;
        const bool want = v < n && fabsf(res[v]) > eps;
        if (want) queued[v] = 1;
        const uint64_t mask = __ballot(want);
        if (!mask) continue;
        const int leader = __ffsll((unsigned long long)mask) - 1;
        uint32_t base = 0;
        if ((int)lane_id() == leader) base = atomicAdd(&c->aff_count, (uint32_t)__popcll(mask));
        base = __shfl(base, leader, kWave);
        if (want) out[base + lanes_below(mask)] = v;
    }
}

__global__ void k_iota(uint32_t *q, uint32_t n) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) q[v] = v;
}

// Dense solve (no reference counterpart; the reference always pushes,
// pr_residual.cuh:53-188): when the frontier holds more than 1/32 of the
// vertices, every round consumes all residuals and each vertex pulls its new
// residual over its in-row (segmented wave sums, one atomic per row segment
// instead of one per edge), until no residual exceeds eps or the round cap.
// Same fixed point and the same invariant x + r as the push solve.
// ---- dense PageRank sweeps (a frontier above 1/32 of the vertices) ----------
// One round: consume every residual into the rank and the share each
// out-neighbour pulls (pr_residual.cuh:53-69 without the eps filter), then
// every vertex sums the shares over its in-row.  Rounds are plain launches; a
// round starts only while the previous round's largest |residual| is above
// eps and the round cap is not reached (Ctl::done), so launches past the end
// return at once.
constexpr uint32_t kDenseSmall = 64;     // in-rows up to this many slots: 4 lanes each
constexpr uint32_t kDenseMedium = 256;   // in-rows up to this many slots: 16 lanes each
constexpr uint32_t kDensePiece = 2048;   // larger in-rows: one wave per piece of this many slots

// In-rows by size: small and medium rows listed (count[0], count[2]), larger
// rows cut into pieces (count[1]); the total slot count prices a round (edges).
__global__ __launch_bounds__(1024) void k_prd_split(StoreView s, uint32_t n, uint32_t *small, uint32_t *medium,
                                                    uint2 *big, Ctl *c) {
    // list positions with one atomic per workgroup and list (one per wave on the
    // two list counters serialised into ~0.9 ms at n = 2^22)
    __shared__ uint32_t wcnt[16], wbase;
    __shared__ WgSum<1> tot;
    wg_sum_init(tot);
    __syncthreads();
    unsigned long long slots = 0;
    for (uint32_t b = blockIdx.x * blockDim.x; b < n; b += gridDim.x * blockDim.x) {
        const uint32_t v = b + threadIdx.x;
        const uint32_t ext = v < n ? s.in[v].ext : 0u;
        slots += ext;
        const bool sm = ext && ext <= kDenseSmall;
        const bool md = ext > kDenseSmall && ext <= kDenseMedium;
        const uint32_t np = ext > kDenseMedium ? (ext + kDensePiece - 1) / kDensePiece : 0u;
        const uint32_t ps = wg_reserve(sm ? 1u : 0u, &c->count[0], wcnt, wbase);
        if (sm) small[ps] = v;
        const uint32_t pm = wg_reserve(md ? 1u : 0u, &c->count[2], wcnt, wbase);
        if (md) medium[pm] = v;
        const uint32_t pb = wg_reserve(np, &c->count[1], wcnt, wbase);
        for (uint32_t p = 0; p < np; ++p) big[pb + p] = make_uint2(v, p * kDensePiece);
    }
    for (int d = 32; d >= 1; d >>= 1) slots += __shfl_xor(slots, d, kWave);
    unsigned long long *const dst[1] = {&c->far_scanned};   // slots per round
    const unsigned long long x[1] = {slots};
    wg_sum_flush(tot, x, dst);
}

__global__ void k_prd_consume(StoreView s, uint32_t n, uint32_t r, uint32_t max_rounds, float eps, float *rank,
                              float *res, float *delta, uint32_t *queued, Ctl *c) {
    if (*(volatile uint32_t *)&c->done) return;
    if (r >= max_rounds || (r > 0 && !(__uint_as_float(*(volatile uint32_t *)&c->rmax[(r - 1) & 1]) > eps))) {
        if (blockIdx.x == 0 && threadIdx.x == 0) c->done = 1;
        return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c->rmax[r & 1] = 0;
        c->rounds = r + 1;
        c->edges += c->far_scanned;
    }
    const apps::PageRank pr(eps, 0.f);
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        if (!isfinite(res[v]) || !isfinite(rank[v])) atomicExch((unsigned long long *)&c->bad, 1ull);
        queued[v] = 0;
        const float x = pr.CombineValueBuffer(v, &rank[v], &res[v]);   // every residual, no eps filter
        const uint32_t live = row_live(s.out[v]);
        delta[v] = live ? 0.85f * x / (float)live : 0.f;
    }
}

// A share of an in-row entry: a tombstone keeps its id under the tombstone bit,
// so the gather is issued for every entry and the tombstone's share dropped
// after it (the loads of a lane do not wait on each other).
__device__ __forceinline__ float share_of(const float *delta, uint32_t x) {
    const float d = delta[x & ~kTomb];
    return (x & kTomb) ? 0.f : d;
}

// Listed in-rows of at most LANES * PER slots: LANES lanes per row, each of
// which first loads its (at most PER) entries and then gathers their shares.
// Most in-rows of a power-law graph are short: rows of at most 64 slots take 4
// lanes, rows of 65..256 slots take 16 (a block of at most 256 slots lies in
// one chunk).
template <uint32_t LANES, uint32_t PER>
__global__ __launch_bounds__(256) void k_prd_pull_rows(StoreView s, const uint32_t *rows, const uint32_t *count,
                                                       const Ctl *c, const float *delta, float *res) {
    if (*(volatile const uint32_t *)&c->done) return;
    const uint32_t nrows = *count;
    const uint32_t lane = lane_id(), ql = lane & (LANES - 1);
    const uint64_t groups = (uint64_t)gridDim.x * blockDim.x / LANES;
    for (uint64_t i = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / LANES; i < nrows; i += groups) {
        const uint32_t v = rows[i];
        const RowDesc d = s.in[v];
        const uint32_t *base = row_base(s, d.off);
        uint32_t x[PER];
#pragma unroll
        for (uint32_t k = 0; k < PER; ++k) x[k] = ql + LANES * k < d.ext ? base[ql + LANES * k] : kTomb;
        float sum = 0.f;
#pragma unroll
        for (uint32_t k = 0; k < PER; ++k) sum += share_of(delta, x[k]);
        for (int o = LANES / 2; o >= 1; o >>= 1) sum += __shfl_xor(sum, o, LANES);
        if (ql == 0) {
            res[v] = sum;
            mark_read(s, s.hot_in, v);
        }
    }
}
static_assert(kDenseSmall == 4 * 16 && kDenseMedium == 16 * 16, "row classes match the lanes and loads");

// Pieces of larger in-rows: one wave per piece, one atomic per piece.
__global__ __launch_bounds__(256) void k_prd_pull_big(StoreView s, const uint2 *pieces, const Ctl *c,
                                                      const float *delta, float *res) {
    if (*(volatile const uint32_t *)&c->done) return;
    const uint32_t np = c->count[1];
    const uint32_t lane = lane_id();
    const uint64_t nwaves = (uint64_t)gridDim.x * blockDim.x / kWave;
    for (uint64_t i = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / kWave; i < np; i += nwaves) {
        const uint2 pc = pieces[i];
        const RowDesc d = s.in[pc.x];
        const uint32_t len = d.ext - pc.y < kDensePiece ? d.ext - pc.y : kDensePiece;
        // blocks are aligned to their power-of-two capacity and a piece starts at a
        // multiple of kDensePiece inside one, so a piece never crosses a chunk
        const uint32_t *base = row_base(s, d.off + pc.y);
        float sum = 0.f;
        // eight independent entry loads per lane, then their eight gathers
        constexpr uint32_t kUnroll = 8;
        for (uint32_t j0 = 0; j0 < len; j0 += kUnroll * kWave) {
            uint32_t x[kUnroll];
#pragma unroll
            for (uint32_t k = 0; k < kUnroll; ++k) {
                const uint32_t j = j0 + k * kWave + lane;
                x[k] = j < len ? base[j] : kTomb;
            }
#pragma unroll
            for (uint32_t k = 0; k < kUnroll; ++k) sum += share_of(delta, x[k]);
        }
        for (int o = 32; o >= 1; o >>= 1) sum += __shfl_xor(sum, o, kWave);
        if (lane == 0) {
            atomicAdd(&res[pc.x], sum);
            if (pc.y == 0) mark_read(s, s.hot_in, pc.x);
        }
    }
}

// Largest |residual| after round r's pull.
// Largest |residual| of the round: a wave and then a workgroup maximum, and one
// atomic per workgroup, skipped when the word already holds as much (a
// non-negative float orders like its bit pattern).  One atomic per wave on
// the single word serialised into ~0.2 ms per round on RMAT-22.
__global__ __launch_bounds__(256) void k_prd_rmax(const float *res, uint32_t n, uint32_t r, Ctl *c) {
    __shared__ float wmax[256 / kWave];
    if (*(volatile uint32_t *)&c->done) return;   // written by an earlier launch: the same for every thread
    float m = 0.f;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        m = fmaxf(m, fabsf(res[v]));
    for (int d = 32; d >= 1; d >>= 1) m = fmaxf(m, __shfl_xor(m, d, kWave));
    if (lane_id() == 0) wmax[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float b = wmax[0];
        for (uint32_t w = 1; w < blockDim.x / kWave; ++w) b = fmaxf(b, wmax[w]);
        const uint32_t bits = __float_as_uint(b);
        if (b > 0.f && bits > *(volatile uint32_t *)&c->rmax[r & 1]) atomicMax(&c->rmax[r & 1], bits);
    }
}

void Engine::pagerank_dense(Ctl *c, dcg_phase_stats *st) {
    const uint64_t slots_bound = total_slots / kDensePiece + n + 1;
    d_prd_small.ensure(n ? n : 1);
    d_prd_medium.ensure(n ? n : 1);
    d_prd_big.ensure(slots_bound);
    hipLaunchKernelGGL(k_ctl_setup, dim3(1), dim3(64), 0, stream, c, (const uint32_t *)nullptr, 0u, 0u,
                       (const uint32_t *)nullptr);
    fill<unsigned long long>(stream, &c->far_scanned, 1, 0ull);
    fill<uint32_t>(stream, &c->done, 1, 0u);
    fill<uint32_t>(stream, c->rmax, 2, 0u);
    const StoreView v = view();
    hipLaunchKernelGGL(k_prd_split, dim3(grid_for(n, 1024, 4096)), dim3(1024), 0, stream, v, n, d_prd_small.p,
                       d_prd_medium.p, d_prd_big.p, c);
    const uint32_t max_rounds = cfg.pr_max_rounds;   // 1..100, checked at creation
    const uint32_t cg = grid_for(n, 256, 4096), sg = grid_for((uint64_t)n * 4, 256, 65536);
    const uint32_t mg = grid_for(n, 256, 8192);   // medium rows: 16 lanes each, at most n / 65 of them
    const uint32_t rg = grid_for(n, 256, 1024);   // residual maximum: ~16 entries per thread at n = 2^22
    const float eps = (float)cfg.pr_epsilon;
    // rounds in chunks of 16 between host checks; launches past convergence return at once
    const int slot = time_begin(DCG_K_PR_SOLVE, -1);
    for (uint32_t r0 = 0;; r0 += 16) {
        for (uint32_t r = r0; r < r0 + 16; ++r) {
            hipLaunchKernelGGL(k_prd_consume, dim3(cg), dim3(256), 0, stream, v, n, r, max_rounds, eps, d_rank.p,
                               d_res.p, d_delta.p, d_queued.p, c);
            hipLaunchKernelGGL((k_prd_pull_rows<4, 16>), dim3(sg), dim3(256), 0, stream, v, d_prd_small.p,
                               &c->count[0], c, d_delta.p, d_res.p);
            hipLaunchKernelGGL((k_prd_pull_rows<16, 16>), dim3(mg), dim3(256), 0, stream, v, d_prd_medium.p,
                               &c->count[2], c, d_delta.p, d_res.p);
            hipLaunchKernelGGL(k_prd_pull_big, dim3(8192), dim3(256), 0, stream, v, d_prd_big.p, c, d_delta.p,
                               d_res.p);
            hipLaunchKernelGGL(k_prd_rmax, dim3(rg), dim3(256), 0, stream, d_res.p, n, r, c);
        }
        DCG_HIP(hipGetLastError());
        uint32_t done = 0;
        DCG_HIP(hipMemcpyAsync(&done, &c->done, 4, hipMemcpyDeviceToHost, stream));
        sync();
        if (done || (uint64_t)r0 + 16 > (uint64_t)max_rounds) break;
    }
    time_end(slot);
    hipLaunchKernelGGL(k_pr_pending, dim3(grid_for(n)), dim3(256), 0, stream, d_res.p, d_queued.p, n,
                       (float)cfg.pr_epsilon, d_qc.p, c);
    DCG_HIP(hipGetLastError());
    DCG_HIP(hipMemcpyAsync(h_ctl + (c - d_ctl.p), c, sizeof(Ctl), hipMemcpyDeviceToHost, stream));
    sync();
    const Ctl &h = h_ctl[c - d_ctl.p];
    // 24 B per row pulled (descriptor + share and residual) + 12 B per slot
    timed[slot].bytes = 24.0 * (double)n * (double)h.rounds + 12.0 * (double)h.edges;
    if (h.bad) throw Error(DCG_ERR_NUMERIC, "PR non-finite rank/residual");
    pr_pending = d_qc.p;
    pr_pending_count = h.aff_count;
    if (st) {
        st->rounds = h.rounds;
        st->edges_scanned = h.edges;
        st->affected = pr_pending_count;
    }
}

// One PR batch (framework.cuh:2692-2736): Amend(-1) over the old rows of every
// touched source, delete+add mutation, Amend(+1) over the new rows, then the
// residual solve capped at pr_max_rounds; an unfinished frontier carries over.
void Engine::pagerank_batch(uint64_t ab, uint64_t ac, uint64_t db, uint64_t dc, dcg_phase_stats *st) {
    Ctl *c = d_ctl.p;
    // Pending frontier must live in a buffer that neither amend nor converge
    // overwrites before it is consumed: move it to qc when needed.
    if (pr_pending != d_qc.p && pr_pending_count) {
        DCG_HIP(hipMemcpyAsync(d_qc.p, pr_pending, pr_pending_count * 4, hipMemcpyDeviceToDevice, stream));
    }
    pr_pending = d_qc.p;
    // unique touched sources (Runtime::BeginBatch sources) -> qa
    uint32_t *srcs = d_qa.p;
    hipLaunchKernelGGL(k_ctl_setup, dim3(1), dim3(64), 0, stream, c + 0, (const uint32_t *)nullptr, 0u, 0u,
                       (const uint32_t *)nullptr);
    const uint32_t stmp = take_stamps(*this);
    const dcg_edge *adds = batch_records(kAdds, ab, ac), *dels = batch_records(kDels, db, dc);
    hipLaunchKernelGGL(k_pr_sources, dim3(grid_for(ac + dc)), dim3(256), 0, stream, adds, ac, dels,
                       dc, n, d_vs.p, stmp, srcs, c + 0);
    // Amend(-1): appends after the carried entries in qc (count[1] starts there)
    hipLaunchKernelGGL(k_ctl_setup, dim3(1), dim3(64), 0, stream, c + 1, &c[0].aff_count, 0u,
                       (uint32_t)pr_pending_count, (const uint32_t *)nullptr);
    TravArgs a = base_args(*this);
    a.ctl = c + 1;
    a.in0 = srcs;
    a.qa = d_qc.p;   // round 0 output buffer
    a.max_rounds = 1;
    a.sign = -1.f;
    launch_persistent(*this, M_PR_AMEND, a);
    DCG_HIP(hipGetLastError());
    // topology mutation of the whole batch
    dcg_phase_stats dst{}, ast{};
    mutate_delete(dels, dc, &dst);
    sync();
    finish_delete(dc, &dst);
    mutate_add(adds, ac, &ast);
    mark_batches_used();
    sync();
    finish_add(ac, &ast);
    // Amend(+1) over the new rows; count[1] continues from ctl[1].count[1]
    hipLaunchKernelGGL(k_ctl_setup, dim3(1), dim3(64), 0, stream, c + 2, &c[0].aff_count, 0u, 0u,
                       (const uint32_t *)nullptr);
    // copy the running count into ctl[2].count[1]
    DCG_HIP(hipMemcpyAsync(&c[2].count[1], &c[1].count[1], 4, hipMemcpyDeviceToDevice, stream));
    a.ctl = c + 2;
    a.sign = 1.f;
    launch_persistent(*this, M_PR_AMEND, a);
    DCG_HIP(hipGetLastError());
    uint32_t frontier = 0;
    DCG_HIP(hipMemcpyAsync(&frontier, &c[2].count[1], 4, hipMemcpyDeviceToHost, stream));
    sync();
    if (dense_pr(frontier)) {
        DCG_HIP(hipMemcpyAsync(h_ctl, c, 3 * sizeof(Ctl), hipMemcpyDeviceToHost, stream));
        sync();
        for (int i = 0; i < 3; ++i)
            check_launch(h_ctl[i], "PageRank");
        const uint64_t amend_edges = h_ctl[1].edges + h_ctl[2].edges;
        const uint64_t touched = h_ctl[0].aff_count;
        pagerank_dense(c + 3, st);
        if (st) {
            st->requests = ac + dc;
            st->touched_sources = touched;
            st->missing_deletes = dst.missing_deletes;
            st->invalid_additions = ast.invalid_additions;
            st->relocated_rows = ast.relocated_rows;
            st->changed_sources = dst.changed_sources + ast.changed_sources;
            st->edges_scanned += amend_edges;
        }
        return;
    }
    // converge from qc (count = ctl[2].count[1])
    hipLaunchKernelGGL(k_ctl_setup, dim3(1), dim3(64), 0, stream, c + 3, &c[2].count[1], 0u, 0u,
                       (const uint32_t *)nullptr);
    TravArgs b = base_args(*this);
    b.ctl = c + 3;
    b.in0 = d_qc.p;
    b.max_rounds = cfg.pr_max_rounds;
    launch_persistent(*this, M_PR, b);
    DCG_HIP(hipGetLastError());
    DCG_HIP(hipMemcpyAsync(h_ctl, c, 4 * sizeof(Ctl), hipMemcpyDeviceToHost, stream));
    sync();
    for (int i = 0; i < 4; ++i) {
        check_launch(h_ctl[i], "PageRank");
        if (h_ctl[i].bad) throw Error(DCG_ERR_NUMERIC, "PR non-finite rank/residual");
    }
    const uint32_t r = h_ctl[3].rounds;
    if (r == 0) {
        pr_pending = d_qc.p;
        pr_pending_count = h_ctl[3].count[0];
    } else {
        pr_pending = ((r - 1) & 1) ? d_qb.p : d_qa.p;
        pr_pending_count = h_ctl[3].count[r % 3];
    }
    if (st) {
        st->requests = ac + dc;
        st->touched_sources = h_ctl[0].aff_count;
        st->missing_deletes = dst.missing_deletes;
        st->invalid_additions = ast.invalid_additions;
        st->relocated_rows = ast.relocated_rows;
        st->changed_sources = dst.changed_sources + ast.changed_sources;
        st->rounds = r;
        st->edges_scanned = h_ctl[1].edges + h_ctl[2].edges + h_ctl[3].edges;
        st->affected = pr_pending_count;
    }
}

}  // namespace dcg
