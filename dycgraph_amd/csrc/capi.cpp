// capi.cpp -- C ABI of the engine (include/dcg.h).  Every entry point catches
// exceptions and returns a status; nothing aborts across the ABI.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "engine.hpp"
#include "mlp.hpp"

using namespace dcg;

struct dcg_engine {
    Engine *e;
    std::string err;
};

namespace dcg {

Engine::Engine(const dcg_config &c) : cfg(c) {
    device = c.device;
    DCG_HIP(hipSetDevice(device));
    DCG_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    DCG_HIP(hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking));
    DCG_HIP(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
    DCG_HIP(hipEventCreateWithFlags(&join_ev, hipEventDisableTiming));
    DCG_HIP(hipStreamCreateWithFlags(&copy_stream, hipStreamNonBlocking));
    for (auto &kind : slots)
        for (BatchSlot &b : kind) {
            DCG_HIP(hipEventCreateWithFlags(&b.ready, hipEventDisableTiming));
            DCG_HIP(hipEventCreateWithFlags(&b.used, hipEventDisableTiming));
        }
    hipDeviceProp_t prop;
    DCG_HIP(hipGetDeviceProperties(&prop, device));
    cu_count = prop.multiProcessorCount;
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
        wall_khz = (double)khz;
    DCG_HIP(hipHostMalloc((void **)&h_ctl, 8 * sizeof(Ctl), hipHostMallocDefault));
    DCG_HIP(hipHostMalloc((void **)&h_step, 2 * sizeof(StepArgs), hipHostMallocDefault));
    d_step.alloc(2);
    DCG_HIP(hipHostMalloc((void **)&h_counters, 2 * Engine::kCounterSlots * sizeof(unsigned long long),
                          hipHostMallocDefault));
    DCG_HIP(hipHostMalloc((void **)&h_tail, Engine::kTailWords * sizeof(uint64_t), hipHostMallocDefault));
    DCG_HIP(hipHostMalloc((void **)&h_fold, 4 * sizeof(unsigned long long), hipHostMallocDefault));
    app = c.app;
    const bool mixed = c.query_apps && app != DCG_APP_PR;
    Q = app == DCG_APP_PR || (app == DCG_APP_CC && !mixed) ? 1u : std::max(1u, c.num_queries);
    if (c.sources && app != DCG_APP_PR) sources.assign(c.sources, c.sources + Q);
    else sources.assign(Q, 0u);
    prog = builtin_program();
    qapps.assign(Q, app == DCG_APP_PR ? 0u : app);
    if (mixed) {
        apps_given = true;
        for (uint32_t q = 0; q < Q; ++q) {
            // program indices: SSSP, BFS, CC in the built-in set (checked against
            // the installed set at graph load)
            if (c.query_apps[q] >= DCG_PROGRAM_MAX) throw Error(DCG_ERR_ARG, "query_apps entries out of range");
            qapps[q] = c.query_apps[q];
        }
    }
    wp.seed = c.weight_seed;
    wp.wmax = c.weight_max ? c.weight_max : 1;
    // SSSP weights; BFS (1) and CC (0) queries override per query in the kernels
    wp.kind = c.weight_mode == DCG_W_SEEDED ? W_SEEDED : W_REFERENCE;
}

int Engine::time_begin(int kernel, int ctl, double bytes, hipStream_t on) {
    const size_t slot = timed.size();
    if (2 * slot + 2 > events.size()) {
        for (int i = 0; i < 32; ++i) {
            hipEvent_t ev;
            DCG_HIP(hipEventCreate(&ev));
            events.push_back(ev);
        }
    }
    timed.push_back(Timed{kernel, (int)(2 * slot), ctl, bytes, 0, on ? on : stream});
    DCG_HIP(hipEventRecord(events[2 * slot], on ? on : stream));
    return (int)slot;
}

void Engine::time_end(int slot) { DCG_HIP(hipEventRecord(events[timed[slot].ev + 1], timed[slot].on)); }

// ---- update stream: device batch slots --------------------------------------
void Engine::forget_batches() {
    for (auto &kind : slots)
        for (BatchSlot &b : kind) {
            if (b.pending) DCG_HIP(hipEventSynchronize(b.ready));
            b.pending = false;
            b.begin = ~0ull;
            b.count = 0;
        }
}

// Copy [begin, begin+count) of one direction into slot `b` on stream `on`
// (after every earlier reader of the slot).  Host records are pinned, so the
// copy is an asynchronous DMA in stream order.
void Engine::fill_slot(int kind, BatchSlot &b, uint64_t begin, uint64_t count, hipStream_t on) {
    if (b.pending) DCG_HIP(hipEventSynchronize(b.ready));
    b.pending = false;
    if (b.buf.n < count) {
        DCG_HIP(hipEventSynchronize(b.used));   // the old buffer's readers have finished
        b.buf.alloc(count);
    }
    DCG_HIP(hipStreamWaitEvent(on, b.used, 0));
    const dcg_edge *h = (kind == kDels ? h_dels : h_adds) + begin;
    const int slot = time_begin(DCG_K_INGEST, -1, 8.0 * (double)count, on);
    DCG_HIP(hipMemcpyAsync(b.buf.p, h, count * sizeof(dcg_edge), hipMemcpyHostToDevice, on));
    time_end(slot);
    b.begin = begin;
    b.count = count;
    b.stamp = ++slot_clock;
}

const dcg_edge *Engine::batch_records(int kind, uint64_t begin, uint64_t count) {
    if (!count) return nullptr;
    BatchSlot *pick = nullptr;
    for (BatchSlot &b : slots[kind])
        if (b.begin == begin && b.count == count) pick = &b;
    if (pick) {
        if (pick->pending) {
            DCG_HIP(hipStreamWaitEvent(stream, pick->ready, 0));
            pick->pending = false;
        }
    } else {
        // refill the older slot that no prefetch is still filling
        BatchSlot &a = slots[kind][0], &b = slots[kind][1];
        if (a.pending != b.pending) pick = a.pending ? &b : &a;
        else pick = a.stamp <= b.stamp ? &a : &b;
        fill_slot(kind, *pick, begin, count, stream);
    }
    pick->read = true;
    return pick->buf.p;
}

// After a call enqueued the readers of its batch slots: their end, which a later
// copy into the slot waits for.  Only the slots this call read get the event (an
// event between launches costs a few microseconds of queue time); the others
// keep the one of their last readers.
void Engine::mark_batches_used() {
    for (auto &kind : slots)
        for (BatchSlot &b : kind)
            if (b.read) {
                DCG_HIP(hipEventRecord(b.used, stream));
                b.read = false;
            }
}

void Engine::prefetch_records(int kind, uint64_t begin, uint64_t count) {
    if (!count) return;
    for (BatchSlot &b : slots[kind])
        if (b.begin == begin && b.count == count) return;   // already there or on its way
    BatchSlot &a = slots[kind][0], &b = slots[kind][1];
    BatchSlot &t = a.stamp <= b.stamp ? a : b;
    fill_slot(kind, t, begin, count, copy_stream);
    DCG_HIP(hipEventRecord(t.ready, copy_stream));
    t.pending = true;
}

void Engine::sync_all() {
    DCG_HIP(hipStreamSynchronize(stream));
    DCG_HIP(hipStreamSynchronize(stream2));
    DCG_HIP(hipStreamSynchronize(copy_stream));
    mlp_sync(*this);
}

// Closures of a replayed stage graph: each one's span between its wall-clock
// stamps, priced from its Ctl block (copied to h_ctl before the stage's wait).
void Engine::collect_graph_timers(int stage) {
    for (const GraphTimer &g : graph_timers[stage]) {
        const Ctl &c = h_ctl[g.ctl];
        const double ms = c.t_last > c.t_first ? (double)(c.t_last - c.t_first) / wall_khz : 0.0;
        dcg_kernel_stats &k = kstat[g.kernel];
        k.launches++;
        k.ms += ms;
        k.edges += c.edges;
        k.expansions += c.expansions;
        k.algo_bytes += 24.0 * (double)c.expansions + 12.0 * (double)c.edges;
    }
}

// Fold recorded launches into kstat (call after the stream has synchronised).
// Stage calls fold their timing records lazily: an elapsed-time query per
// record is host time on the snapshot's critical path.  Records priced from
// their control blocks (closures launched outside graphs) are folded at once,
// before the next stage reuses the blocks; the statistics calls fold the rest.
void Engine::collect_timing_lazy() {
    bool priced = false;
    for (const Timed &t : timed) priced |= t.ctl >= 0;
    if (priced || timed.size() >= 256) collect_timing();
}

void Engine::collect_timing() {
    if (timed.empty()) return;
    // a copy still running on copy_stream (a prefetch) is folded by a later call
    std::vector<Timed> later;
    for (size_t i = 0; i < timed.size();) {
        if (timed[i].on == copy_stream && hipEventQuery(events[timed[i].ev + 1]) == hipErrorNotReady) {
            later.push_back(timed[i]);
            timed.erase(timed.begin() + (long)i);
        } else {
            ++i;
        }
    }
    Ctl ctl[8];
    bool priced = false;   // closure launches (graphs off) are priced from their Ctl blocks
    for (const Timed &t : timed) priced |= t.ctl >= 0;
    if (priced && d_ctl.p) {
        DCG_HIP(hipMemcpyAsync(ctl, d_ctl.p, sizeof(ctl), hipMemcpyDeviceToHost, stream));
        sync();
    }
    for (const Timed &t : timed) {
        float ms = 0.f;
        DCG_HIP(hipEventElapsedTime(&ms, events[t.ev], events[t.ev + 1]));
        dcg_kernel_stats &k = kstat[t.kernel];
        k.launches++;
        k.ms += ms;
        if (t.ctl >= 0) {
            const Ctl &c = ctl[t.ctl];
            k.edges += c.edges;
            k.expansions += c.expansions;
            k.algo_bytes += 24.0 * (double)c.expansions + 12.0 * (double)c.edges;
        } else {
            k.algo_bytes += t.bytes;
            k.host_lines += t.host_lines;
        }
    }
    timed.clear();
    // keep the pending copies' event pairs: move them to the front of the pool
    for (Timed &t : later) {
        const size_t slot = timed.size();
        std::swap(events[2 * slot], events[t.ev]);
        std::swap(events[2 * slot + 1], events[t.ev + 1]);
        t.ev = (int)(2 * slot);
        timed.push_back(t);
    }
}

Engine::~Engine() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (stream2) (void)hipStreamSynchronize(stream2);
    if (copy_stream) (void)hipStreamSynchronize(copy_stream);
    cache.mlp.reset();   // its events and pinned judge sums
    for (hipEvent_t ev : events) (void)hipEventDestroy(ev);
    for (auto &kind : slots)
        for (BatchSlot &b : kind) {
            b.buf.release();
            if (b.ready) (void)hipEventDestroy(b.ready);
            if (b.used) (void)hipEventDestroy(b.used);
        }
    if (h_adds) (void)hipHostFree(h_adds);
    if (h_dels) (void)hipHostFree(h_dels);
    if (h_slab) (void)hipHostFree(h_slab);
    if (h_ctl) (void)hipHostFree(h_ctl);
    if (h_step) (void)hipHostFree(h_step);
    drop_graphs();
    if (h_counters) (void)hipHostFree(h_counters);
    if (h_tail) (void)hipHostFree(h_tail);
    if (h_fold) (void)hipHostFree(h_fold);
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    if (join_ev) (void)hipEventDestroy(join_ev);
    if (copy_stream) (void)hipStreamDestroy(copy_stream);
    if (stream2) (void)hipStreamDestroy(stream2);
    if (stream) (void)hipStreamDestroy(stream);
}

}  // namespace dcg

static void check_slice(uint64_t begin, uint64_t count, uint64_t total) {
    if (begin > total || count > total - begin) throw Error(DCG_ERR_ARG, "update slice out of range");
}

template <typename F>
static int guarded(dcg_engine *h, F &&f) {
    if (!h || !h->e) return DCG_ERR_ARG;
    try {
        f(*h->e);
        h->err.clear();
        return DCG_OK;
    } catch (const Error &x) {
        h->err = x.what();
        return x.status;
    } catch (const std::bad_alloc &) {
        h->err = "host allocation failed";
        return DCG_ERR_CAPACITY;
    } catch (const std::exception &x) {
        h->err = x.what();
        return DCG_ERR_STATE;
    }
}

extern "C" {

int dcg_config_init(dcg_config *cfg) {
    if (!cfg) return DCG_ERR_ARG;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->app = DCG_APP_SSSP;
    cfg->num_queries = 1;
    cfg->weight_mode = DCG_W_REFERENCE;
    cfg->weight_max = 100;
    cfg->weight_seed = 1;
    cfg->placement = DCG_PLACE_DEVICE;
    cfg->cache_policy = DCG_CACHE_HOTNESS;
    cfg->cache_bytes = 4000000000ull;
    cfg->store_slack = 0.125;
    cfg->pr_base = 0.15;
    cfg->pr_epsilon = 1e-6;
    cfg->pr_max_rounds = 100;
    cfg->block_size = 256;
    cfg->device = 0;
    return DCG_OK;
}

const char *dcg_status_string(int status) {
    switch (status) {
    case DCG_OK: return "ok";
    case DCG_ERR_ARG: return "invalid argument";
    case DCG_ERR_HIP: return "HIP runtime error";
    case DCG_ERR_STATE: return "invalid engine state";
    case DCG_ERR_CAPACITY: return "capacity exhausted";
    case DCG_ERR_NUMERIC: return "non-finite PageRank state";
    case DCG_ERR_UNSUPPORTED: return "unsupported";
    case DCG_ERR_IO: return "file I/O failed";
    default: return "unknown status";
    }
}

int dcg_create(const dcg_config *cfg, dcg_engine **out) {
    if (!cfg || !out) return DCG_ERR_ARG;
    *out = nullptr;
    if (cfg->app > DCG_APP_PR || cfg->placement > DCG_PLACE_HOST || cfg->cache_policy > DCG_CACHE_LEARNED)
        return DCG_ERR_ARG;
    if (cfg->app == DCG_APP_PR) {
        // samples/hybrid_pr/hybrid_pr.cu:133-137
        if (!std::isfinite(cfg->pr_epsilon) || cfg->pr_epsilon <= 0 || (float)cfg->pr_epsilon == 0.f ||
            cfg->pr_max_rounds == 0 || cfg->pr_max_rounds > 100 || !(cfg->pr_base > 0))
            return DCG_ERR_ARG;
    }
    if (cfg->weight_mode == DCG_W_SEEDED && cfg->weight_max == 0) return DCG_ERR_ARG;
    if (cfg->app != DCG_APP_PR && cfg->num_queries > 1 && !cfg->sources) return DCG_ERR_ARG;
    dcg_engine *h = new (std::nothrow) dcg_engine();
    if (!h) return DCG_ERR_CAPACITY;
    try {
        h->e = new Engine(*cfg);
    } catch (const Error &x) {
        delete h;
        return x.status;
    } catch (const std::exception &) {
        delete h;
        return DCG_ERR_STATE;
    }
    *out = h;
    return DCG_OK;
}

void dcg_destroy(dcg_engine *h) {
    if (!h) return;
    delete h->e;
    delete h;
}

const char *dcg_last_error(const dcg_engine *h) { return h ? h->err.c_str() : "null engine"; }

int dcg_set_program(dcg_engine *h, const dcg_program *p) {
    return guarded(h, [&](Engine &e) {
        if (e.loaded) throw Error(DCG_ERR_STATE, "dcg_set_program after a graph load");
        if (e.app == DCG_APP_PR) throw Error(DCG_ERR_UNSUPPORTED, "PageRank runs the built-in residual push");
        if (p) {
            if (p->abi != DCG_PROGRAM_ABI || p->tag != program_layout_tag())
                throw Error(DCG_ERR_ARG, "program table of another ABI or kernel-argument layout");
            if (p->num_programs < 1 || p->num_programs > DCG_PROGRAM_MAX || !p->bucket_width)
                throw Error(DCG_ERR_ARG, "bad program table");
            for (const void *k : p->kernels)
                if (!k) throw Error(DCG_ERR_ARG, "program table without a kernel");
        }
        e.prog = p ? p : builtin_program();
        if (!e.apps_given) e.qapps.assign(e.Q, p ? 0u : e.app);
        for (auto &g : e.coop_grid) g = 0;
        e.drop_graphs();
    });
}

// Query program indices must exist in the installed program set.
static void check_programs(const Engine &e) {
    for (uint32_t q : e.qapps)
        if (q >= e.prog->num_programs) throw Error(DCG_ERR_ARG, "query_apps entry beyond the program set");
}

int dcg_load_edges_device(dcg_engine *h, uint32_t num_nodes, const uint32_t *d_src, const uint32_t *d_dst,
                          uint64_t num_edges) {
    return guarded(h, [&](Engine &e) {
        if (e.loaded) throw Error(DCG_ERR_STATE, "graph already loaded");
        if (num_nodes == 0 || (num_edges && (!d_src || !d_dst))) throw Error(DCG_ERR_ARG, "bad graph arguments");
        if (e.app != DCG_APP_PR) check_programs(e);
        e.load_edges_device(num_nodes, d_src, d_dst, num_edges);
        e.alloc_state();
    });
}

int dcg_load_edges(dcg_engine *h, uint32_t num_nodes, const uint32_t *src, const uint32_t *dst, uint64_t num_edges) {
    return guarded(h, [&](Engine &e) {
        if (e.loaded) throw Error(DCG_ERR_STATE, "graph already loaded");
        if (num_nodes == 0 || (num_edges && (!src || !dst))) throw Error(DCG_ERR_ARG, "bad graph arguments");
        if (e.app != DCG_APP_PR) check_programs(e);
        DevBuf<uint32_t> s, d;
        s.alloc(num_edges ? num_edges : 1);
        d.alloc(num_edges ? num_edges : 1);
        if (num_edges) {
            DCG_HIP(hipMemcpyAsync(s.p, src, num_edges * 4, hipMemcpyHostToDevice, e.stream));
            DCG_HIP(hipMemcpyAsync(d.p, dst, num_edges * 4, hipMemcpyHostToDevice, e.stream));
        }
        e.load_edges_device(num_nodes, s.p, d.p, num_edges);
        e.alloc_state();
    });
}

int dcg_load_csr(dcg_engine *h, uint32_t num_nodes, const uint64_t *row_ptr, const uint32_t *col) {
    if (!row_ptr) return DCG_ERR_ARG;
    const uint64_t m = row_ptr[num_nodes];
    std::vector<uint32_t> src(m);
    for (uint32_t u = 0; u < num_nodes; ++u)
        for (uint64_t k = row_ptr[u]; k < row_ptr[u + 1]; ++k) src[k] = u;
    return dcg_load_edges(h, num_nodes, src.data(), col, m);
}

int dcg_start(dcg_engine *h) {
    return guarded(h, [&](Engine &e) {
        if (!e.loaded) throw Error(DCG_ERR_STATE, "dcg_start before a graph load");
        // source clamp of hybrid_sssp.cu:391
        for (auto &s : e.sources) s = std::min(s, e.n - 1);
        e.d_sources.alloc(e.sources.size());
        DCG_HIP(hipMemcpyAsync(e.d_sources.p, e.sources.data(), e.sources.size() * 4, hipMemcpyHostToDevice, e.stream));
        e.d_qapps.alloc(e.qapps.size());
        DCG_HIP(hipMemcpyAsync(e.d_qapps.p, e.qapps.data(), e.qapps.size() * 4, hipMemcpyHostToDevice, e.stream));
        if (e.app == DCG_APP_PR) e.start_pagerank();
        else e.start_min_lattice();
        e.cache_admit_measured();
        e.collect_timing();
        e.started = true;
    });
}

int dcg_load_updates(dcg_engine *h, const dcg_edge *adds, uint64_t num_adds, const dcg_edge *dels, uint64_t num_dels) {
    return guarded(h, [&](Engine &e) {
        if ((num_adds && !adds) || (num_dels && !dels)) throw Error(DCG_ERR_ARG, "null update array");
        // no copy of the old stream is still read: every stage waits for its work,
        // and a pending prefetch is waited for here
        e.sync_all();
        e.forget_batches();
        if (e.h_adds) (void)hipHostFree(e.h_adds);
        if (e.h_dels) (void)hipHostFree(e.h_dels);
        e.h_adds = e.h_dels = nullptr;
        e.n_adds = e.n_dels = 0;
        DCG_HIP(hipHostMalloc((void **)&e.h_adds, (num_adds ? num_adds : 1) * sizeof(dcg_edge), hipHostMallocDefault));
        DCG_HIP(hipHostMalloc((void **)&e.h_dels, (num_dels ? num_dels : 1) * sizeof(dcg_edge), hipHostMallocDefault));
        if (num_adds) std::memcpy(e.h_adds, adds, num_adds * sizeof(dcg_edge));
        if (num_dels) std::memcpy(e.h_dels, dels, num_dels * sizeof(dcg_edge));
        e.n_adds = num_adds;
        e.n_dels = num_dels;
    });
}

int dcg_prefetch_updates(dcg_engine *h, uint64_t add_begin, uint64_t add_count, uint64_t del_begin,
                         uint64_t del_count) {
    return guarded(h, [&](Engine &e) {
        check_slice(add_begin, add_count, e.n_adds);
        check_slice(del_begin, del_count, e.n_dels);
        e.prefetch_records(Engine::kDels, del_begin, del_count);
        e.prefetch_records(Engine::kAdds, add_begin, add_count);
    });
}

int dcg_del_edge(dcg_engine *h, uint64_t begin, uint64_t count, dcg_phase_stats *st) {
    return guarded(h, [&](Engine &e) {
        if (!e.started) throw Error(DCG_ERR_STATE, "dcg_del_edge before dcg_start");
        if (e.app == DCG_APP_PR) throw Error(DCG_ERR_STATE, "PageRank batches use dcg_update_pagerank");
        check_slice(begin, count, e.n_dels);
        if (st) *st = dcg_phase_stats{};
        const double t0 = now_ms();
        e.delete_phase(begin, count, st);
        e.collect_timing_lazy();
        if (st) {
            st->requests = count;
            st->ms = now_ms() - t0;
        }
    });
}

int dcg_add_edge(dcg_engine *h, uint64_t begin, uint64_t count, dcg_phase_stats *st) {
    return guarded(h, [&](Engine &e) {
        if (!e.started) throw Error(DCG_ERR_STATE, "dcg_add_edge before dcg_start");
        if (e.app == DCG_APP_PR) throw Error(DCG_ERR_STATE, "PageRank batches use dcg_update_pagerank");
        check_slice(begin, count, e.n_adds);
        if (st) *st = dcg_phase_stats{};
        const double t0 = now_ms();
        e.add_phase(begin, count, st);
        e.collect_timing_lazy();
        if (st) {
            st->requests = count;
            st->ms = now_ms() - t0;
        }
    });
}

int dcg_snapshot(dcg_engine *h, uint64_t del_begin, uint64_t del_count, uint64_t add_begin, uint64_t add_count,
                 dcg_phase_stats *dst, dcg_phase_stats *ast) {
    return guarded(h, [&](Engine &e) {
        if (!e.started) throw Error(DCG_ERR_STATE, "dcg_snapshot before dcg_start");
        if (e.app == DCG_APP_PR) throw Error(DCG_ERR_STATE, "PageRank batches use dcg_update_pagerank");
        check_slice(del_begin, del_count, e.n_dels);
        check_slice(add_begin, add_count, e.n_adds);
        if (dst) *dst = dcg_phase_stats{};
        if (ast) *ast = dcg_phase_stats{};
        const double t0 = now_ms();
        e.snapshot_step(del_begin, del_count, add_begin, add_count, dst, ast);
        e.collect_timing_lazy();
        const double ms = now_ms() - t0;
        if (dst) {
            dst->requests = del_count;
            dst->ms = ms;
        }
        if (ast) {
            ast->requests = add_count;
            ast->ms = 0.0;   // the step's time is on the delete record
        }
    });
}

int dcg_update_pagerank(dcg_engine *h, uint64_t add_begin, uint64_t add_count, uint64_t del_begin, uint64_t del_count,
                        dcg_phase_stats *st) {
    return guarded(h, [&](Engine &e) {
        if (!e.started) throw Error(DCG_ERR_STATE, "dcg_update_pagerank before dcg_start");
        if (e.app != DCG_APP_PR) throw Error(DCG_ERR_STATE, "dcg_update_pagerank needs the PR app");
        check_slice(add_begin, add_count, e.n_adds);
        check_slice(del_begin, del_count, e.n_dels);
        if (st) *st = dcg_phase_stats{};
        const double t0 = now_ms();
        e.pagerank_batch(add_begin, add_count, del_begin, del_count, st);
        e.collect_timing_lazy();
        e.fold_reads(st);
        if (st) st->ms = now_ms() - t0;
    });
}

int dcg_cache_maintain(dcg_engine *h, dcg_cache_stats *st) {
    return guarded(h, [&](Engine &e) {
        if (!e.started) throw Error(DCG_ERR_STATE, "cache maintenance before dcg_start");
        e.cache_maintain(st);
    });
}

int dcg_cache_prefetch(dcg_engine *h, uint64_t ab, uint64_t ac, uint64_t db, uint64_t dc, dcg_cache_stats *st) {
    return guarded(h, [&](Engine &e) {
        if (!e.started) throw Error(DCG_ERR_STATE, "cache prefetch before dcg_start");
        check_slice(ab, ac, e.n_adds);
        check_slice(db, dc, e.n_dels);
        e.cache_prefetch(ab, ac, db, dc, st);
    });
}

int dcg_cache_policy_set(dcg_engine *h, uint32_t policy) {
    return guarded(h, [&](Engine &e) {
        if (policy > DCG_CACHE_LEARNED) throw Error(DCG_ERR_ARG, "unknown cache policy");
        e.cfg.cache_policy = policy;
    });
}

int dcg_hot_candidates(dcg_engine *h, uint32_t query, uint64_t cap_slots, uint32_t *ids, uint64_t ids_len,
                       uint64_t *count) {
    return guarded(h, [&](Engine &e) {
        if (!e.started) throw Error(DCG_ERR_STATE, "hot candidates before dcg_start");
        if (query >= e.Q || !count || (ids_len && !ids)) throw Error(DCG_ERR_ARG, "bad hot-candidate arguments");
        *count = e.hot_candidates(query, cap_slots, ids, ids_len);
    });
}

int dcg_set_repair_snapshot(dcg_engine *h, const char *path, uint32_t query) {
    return guarded(h, [&](Engine &e) {
        if (!path || !*path) {
            e.i16_armed = false;
            return;
        }
        if (e.app == DCG_APP_PR) throw Error(DCG_ERR_UNSUPPORTED, "repair snapshots are for min-lattice programs");
        if (query >= e.Q) throw Error(DCG_ERR_ARG, "query out of range");
        e.i16_path = path;
        e.i16_query = query;
        e.i16_armed = true;
    });
}

int dcg_set_cache_trace(dcg_engine *h, const char *path) {
    return guarded(h, [&](Engine &e) { e.f1_open(path); });
}

int dcg_trace_cache_candidates(dcg_engine *h, uint32_t batch, dcg_trace_stats *stats) {
    return guarded(h, [&](Engine &e) { e.f1_record(batch, stats); });
}

int dcg_hot_window(dcg_engine *h, uint32_t *out, size_t n) {
    return guarded(h, [&](Engine &e) {
        if (!e.loaded) throw Error(DCG_ERR_STATE, "no graph loaded");
        if (!out || n != e.n) throw Error(DCG_ERR_ARG, "window size mismatch");
        e.fold_window();   // the counters the ranking would fold first
        DCG_HIP(hipMemcpyAsync(out, e.d_hot_window.p, n * 4, hipMemcpyDeviceToHost, e.stream));
        e.sync();
    });
}

int dcg_gather_values(dcg_engine *h, uint32_t query, void *out, size_t n) {
    return guarded(h, [&](Engine &e) {
        if (!e.started) throw Error(DCG_ERR_STATE, "gather before dcg_start");
        if (!out || n != e.n || query >= e.Q) throw Error(DCG_ERR_ARG, "gather size/query mismatch");
        if (e.app == DCG_APP_PR) {
            DCG_HIP(hipMemcpyAsync(out, e.d_rank.p, n * 4, hipMemcpyDeviceToHost, e.stream));
            e.sync();
            return;
        }
        std::vector<VSlot> tmp(n);
        DCG_HIP(hipMemcpyAsync(tmp.data(), e.d_vs.p + (uint64_t)query * e.n, n * sizeof(VSlot),
                               hipMemcpyDeviceToHost, e.stream));
        e.sync();
        uint32_t *o = static_cast<uint32_t *>(out);
        for (size_t i = 0; i < n; ++i) o[i] = (uint32_t)(tmp[i].st >> 32);
    });
}

int dcg_gather_parents(dcg_engine *h, uint32_t query, uint32_t *out, size_t n) {
    return guarded(h, [&](Engine &e) {
        if (!e.started) throw Error(DCG_ERR_STATE, "gather before dcg_start");
        if (!out || n != e.n || query >= e.Q) throw Error(DCG_ERR_ARG, "gather size/query mismatch");
        if (e.app == DCG_APP_PR) throw Error(DCG_ERR_UNSUPPORTED, "PageRank has no parents");
        // rooted programs (CC) keep (value, depth) in the state word and the parent apart
        std::vector<VSlot> tmp(n);
        DCG_HIP(hipMemcpyAsync(tmp.data(), e.d_vs.p + (uint64_t)query * e.n, n * sizeof(VSlot),
                               hipMemcpyDeviceToHost, e.stream));
        e.sync();
        const bool rooted = e.rooted(query);
        for (size_t i = 0; i < n; ++i) out[i] = rooted ? tmp[i].par : (uint32_t)tmp[i].st;
    });
}

int dcg_gather_buffers(dcg_engine *h, uint32_t query, void *out, size_t n) {
    if (h && h->e && h->e->app != DCG_APP_PR) return dcg_gather_values(h, query, out, n);
    return guarded(h, [&](Engine &e) {
        if (!e.started) throw Error(DCG_ERR_STATE, "gather before dcg_start");
        if (!out || n != e.n || query >= e.Q) throw Error(DCG_ERR_ARG, "gather size/query mismatch");
        DCG_HIP(hipMemcpyAsync(out, e.d_res.p, n * 4, hipMemcpyDeviceToHost, e.stream));
        e.sync();
    });
}

int dcg_get_csr(dcg_engine *h, uint64_t *row_ptr, uint32_t *col, uint64_t *num_edges) {
    return guarded(h, [&](Engine &e) {
        if (!e.loaded) throw Error(DCG_ERR_STATE, "no graph loaded");
        e.export_csr(row_ptr, col, num_edges);
    });
}

int dcg_num_nodes(const dcg_engine *h, uint32_t *n) {
    if (!h || !h->e || !n) return DCG_ERR_ARG;
    *n = h->e->n;
    return DCG_OK;
}

int dcg_get_info(dcg_engine *h, dcg_engine_info *info) {
    return guarded(h, [&](Engine &e) {
        if (!info) throw Error(DCG_ERR_ARG, "null info");
        std::memset(info, 0, sizeof(*info));
        info->snapshots = e.snapshots;
        info->store_slots = e.total_slots;
        info->store_used_slots = e.used_slots;
        info->chunks = e.cache.chunks;
        info->frames = e.cache.frames;
        info->host_bytes = e.h_slab ? e.total_slots * 4 : 0;
        uint64_t store = 0, state = 0, work = 0;
        e.footprint(store, state, work);
        info->device_bytes = store + state + work;
        info->store_device_bytes = store;
        info->state_device_bytes = state;
        info->frame_bytes = e.cache.d_frames.bytes();
        info->cached_edge_reads = e.cum_cached_reads;
        info->host_edge_reads = e.cum_host_reads;
        info->free_blocks = e.pool_len - e.pool_taken;
        info->reused_blocks = e.reused_blocks;
    });
}

int dcg_get_kernel_stats(dcg_engine *h, uint32_t kernel, dcg_kernel_stats *out) {
    return guarded(h, [&](Engine &e) {
        if (!out || kernel >= DCG_K_NUM) throw Error(DCG_ERR_ARG, "bad kernel id");
        e.sync();
        e.collect_timing();
        *out = e.kstat[kernel];
    });
}

int dcg_reset_kernel_stats(dcg_engine *h) {
    return guarded(h, [&](Engine &e) {
        e.sync();
        e.collect_timing();
        for (auto &k : e.kstat) k = dcg_kernel_stats{};
    });
}

int dcg_autotune(dcg_engine *h, dcg_tune_result *out) {
    return guarded(h, [&](Engine &e) {
        if (!e.started) throw Error(DCG_ERR_STATE, "dcg_autotune before dcg_start");
        e.autotune(out);
    });
}

int dcg_autotune_stages(dcg_engine *h, uint32_t stages) {
    return guarded(h, [&](Engine &e) {
        if (!e.started) throw Error(DCG_ERR_STATE, "dcg_autotune_stages before dcg_start");
        if (e.app == DCG_APP_PR) throw Error(DCG_ERR_UNSUPPORTED, "the stage tuner tunes the min-lattice repair");
        e.tune_arm = stages;
    });
}

int dcg_autotune_stage_last(dcg_engine *h, uint32_t stage, dcg_tune_result *out) {
    return guarded(h, [&](Engine &e) {
        if (stage > 1) throw Error(DCG_ERR_ARG, "stage is 0 (delete) or 1 (add)");
        if (!out) throw Error(DCG_ERR_ARG, "null result");
        *out = e.stage_result[stage];
    });
}

int dcg_autotune_last(dcg_engine *h, dcg_tune_result *out) {
    return guarded(h, [&](Engine &e) {
        if (!out) throw Error(DCG_ERR_ARG, "null result");
        *out = e.last_tune;
    });
}

int dcg_set_tuning(dcg_engine *h, uint32_t blocks_per_cu, uint32_t big_row, uint32_t piece) {
    return guarded(h, [&](Engine &e) {
        auto pow2 = [](uint32_t x) { return x && !(x & (x - 1)); };
        if (blocks_per_cu > 16 || (big_row && big_row < 1) || (piece && (!pow2(piece) || piece > 65536)))
            throw Error(DCG_ERR_ARG, "bad tuning values");
        if (blocks_per_cu) e.tune.blocks_per_cu = blocks_per_cu;
        if (big_row) e.tune.big_row = big_row;
        if (piece) e.tune.piece = piece;
        e.forget_stage_tunes();
        for (auto &g : e.coop_grid) g = 0;
        e.drop_graphs();
    });
}

int dcg_set_closure_grid(dcg_engine *h, uint32_t grid_shift, uint32_t solo_max) {
    return guarded(h, [&](Engine &e) {
        if (grid_shift > 16) throw Error(DCG_ERR_ARG, "grid_shift above 16");
        e.tune.grid_shift = grid_shift;
        e.tune.solo_max = solo_max;
        e.forget_stage_tunes();
        for (auto &g : e.coop_grid) g = 0;
        e.drop_graphs();
    });
}

int dcg_set_cpu_domain(dcg_engine *h, const uint8_t *cpu_owned, size_t n) {
    return guarded(h, [&](Engine &e) {
        if (!e.loaded) throw Error(DCG_ERR_STATE, "set the CPU domain after the graph load");
        if (e.app == DCG_APP_PR) throw Error(DCG_ERR_UNSUPPORTED, "co-execution covers the min-lattice closure");
        if (cpu_owned && n != e.n) throw Error(DCG_ERR_ARG, "the CPU domain mask needs one byte per vertex");
        if (cpu_owned && e.prog != builtin_program())
            throw Error(DCG_ERR_UNSUPPORTED, "co-execution runs the built-in programs on the host");
        e.sync_all();
        e.coexec_set(cpu_owned);
    });
}

int dcg_set_barrier_timeout(dcg_engine *h, uint64_t microseconds) {
    return guarded(h, [&](Engine &e) {
        const double ticks = (double)microseconds * (e.wall_khz ? e.wall_khz : 100000.0) / 1000.0;
        e.bar_spin = ticks > 1.8e19 ? ~0ull : (uint64_t)ticks;
        e.drop_graphs();   // captured launches hold the old value
    });
}

int dcg_set_batch_maintenance(dcg_engine *h, int large) {
    return guarded(h, [&](Engine &e) { e.large_batches = large != 0; });
}

int dcg_set_graphs(dcg_engine *h, int on) {
    return guarded(h, [&](Engine &e) {
        e.tune.graphs = on ? 1u : 0u;
        e.drop_graphs();
    });
}

int dcg_set_buckets(dcg_engine *h, uint32_t quarters) {
    return guarded(h, [&](Engine &e) {
        if (quarters > 4096 && quarters != ~0u) throw Error(DCG_ERR_ARG, "bucket width out of range");
        e.tune.bucket_quarters = quarters;
        e.update_buckets();
        e.drop_graphs();
    });
}

int dcg_set_round_trace(dcg_engine *h, uint32_t capacity) {
    return guarded(h, [&](Engine &e) {
        if (capacity > (1u << 26)) throw Error(DCG_ERR_ARG, "round-trace capacity above 2^26 records");
        e.sync_all();
        e.drop_graphs();   // captured launches hold the trace's pointers
        e.d_rtrace.alloc(capacity);
        if (capacity) {
            e.d_rtcur.alloc(1);
            fill<uint32_t>(e.stream, e.d_rtcur.p, 1, 0u);
        } else {
            e.d_rtcur.release();
        }
        e.sync();
    });
}

int dcg_round_trace(dcg_engine *h, dcg_round_rec *out, size_t cap, size_t *written) {
    return guarded(h, [&](Engine &e) {
        if (!written || (cap && !out)) throw Error(DCG_ERR_ARG, "null output");
        *written = 0;
        if (!e.d_rtrace.p) throw Error(DCG_ERR_STATE, "the round trace is off");
        e.sync_all();
        uint32_t cur = 0;
        DCG_HIP(hipMemcpy(&cur, e.d_rtcur.p, 4, hipMemcpyDeviceToHost));
        const size_t held = std::min<size_t>(cur, e.d_rtrace.n);
        const size_t take = std::min(held, cap);
        if (take) DCG_HIP(hipMemcpy(out, e.d_rtrace.p, take * sizeof(dcg_round_rec), hipMemcpyDeviceToHost));
        *written = cur;
        fill<uint32_t>(e.stream, e.d_rtcur.p, 1, 0u);
        e.sync();
    });
}

int dcg_sync(dcg_engine *h) {
    return guarded(h, [&](Engine &e) { e.sync_all(); });
}

}  // extern "C"
