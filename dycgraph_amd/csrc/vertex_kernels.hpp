// vertex_kernels.hpp -- the engine's repair kernels as templates over a set of
// vertex programs (include/dcg_program.hpp), shared by libdcg.so (the built-in
// SSSP/BFS/CC set) and program plugins (include/dcg_program_plugin.hpp).
//
// Vertex state is one 64-bit word per (query, vertex).  Programs without the
// rooted witness: value in the high half, parent in the low half; a relaxation is
// one 64-bit atomicMin, so the value lattice (the reference's atomicMin + parent
// CAS loop, hybrid_sssp.cu:126-145) and the smallest-source tie-break of the pull
// repair (framework.cuh:229-244) come from the same instruction.  Rooted-witness
// programs (CC): value in the high half, hop depth from the value's root in the
// low half, parent in a separate array.  The value lattice is unchanged (CC:
// label = smallest id that reaches the vertex), and the (value, depth) order makes
// parents a shortest-hop tree inside each value class, so a deletion invalidates
// only the part of that tree that lost its path to the root (the exact
// counterpart of directed_label_witness.h) instead of every vertex with an equal
// label.
//
// One cooperative launch runs rounds until the frontier is empty:
//   phase A: waves take 64 frontier entries at a time; rows of <= 256 edges are
//            expanded edge-balanced inside the wave (wave64 prefix sum + binary
//            search of the owner); larger rows are queued as 512-edge pieces.
//   phase B: all waves share the pieces of the large rows.
// New frontier entries go to a workgroup LDS buffer published with one global
// atomic per phase.
//
// A query's program is picked by its qapp entry inside the kernels through
// Visit<ProgramSet<...>> (compile-time dispatch; hooks are inlined).
#pragma once
#include <type_traits>

#include "dcg_internal.hpp"
#include "../../include/dcg_program.hpp"

namespace dcg {

enum Mode : uint32_t { M_PUSH = 0, M_PULL = 1, M_INVALIDATE = 2, M_PR = 3, M_PR_AMEND = 4 };
constexpr uint32_t kDirtyLog2 = 10;   // CC parent-check block: 1,024 state slots
constexpr uint32_t kLdsQueries = 1024;   // queries whose far-pile minima a workgroup gathers in LDS

struct TravArgs {
    StoreView s;
    uint32_t n, Q;
    VSlot *vs;                 // per (query, vertex): state word, dedup stamp, CC parent
    uint8_t *cc_dirty;         // per 1,024 slots: a CC improvement stamped one of them (null without CC)
    const uint32_t *sources;
    const uint32_t *qapp;      // per-query program
    WeightParams w;            // SSSP weights (BFS: 1, CC: 0)
    uint32_t app;
    uint32_t max_rounds;
    const uint32_t *in0;
    uint32_t *qa, *qb;
    uint64_t qcap;
    uint32_t *aff;
    uint64_t affcap;
    BigEntry *big;
    uint64_t bigcap;
    Ctl *ctl;
    uint32_t stamp0;
    const StepArgs *step;      // when set, the round-stamp base is step->stamp[stamp_slot]
    uint32_t stamp_slot;
    uint32_t big_row;          // rows above this many entries are split into pieces
    uint32_t piece;            // piece size (power of two dividing 65,536)
    uint32_t solo_max;         // frontiers of at most this many entries run on workgroup 0 alone
    uint64_t bar_spin;         // wall-clock ticks a workgroup waits at the grid barrier before giving up
    // CPU/GPU co-execution of an insertion closure (null owner: off): a
    // relaxation into a CPU-owned vertex becomes a proposal, the minimum per
    // (query, vertex) in bprop, its slot listed once in blist
    const uint8_t *owner;
    unsigned long long *bprop;
    uint32_t *blist, *bcount;
    uint32_t bcap;
    // bucketed push closure: values at or above thr[q] wait in a far pile until
    // the near frontier drains; null far0 turns it off
    uint32_t *far0, *far1;
    uint32_t *fstamp;          // q*n: epoch in which the entry entered a far pile
    uint32_t *thr;             // per query: values below this are near
    uint32_t *minfar;          // 2*Q per-query minima of the two far piles
    const uint32_t *fdelta;    // per query bucket width (kInf: never far)
    uint32_t *epoch;           // epoch counter (one word, monotone across launches)
    // PageRank
    float *rank, *res, *delta;
    uint32_t *queued;
    float eps;
    float sign;
    // round trace (null: off): workgroup 0 appends one record per round
    dcg_round_rec *rtrace;
    uint32_t *rtcur;
    uint32_t rtcap;
    uint32_t rtag;             // control-block index | mode << 8 | DCG_ROUND_STATIC
};

__device__ __forceinline__ void trace_round(const TravArgs &a, uint32_t r, uint32_t frontier, uint32_t flags) {
    if (!a.rtrace) return;
    const uint32_t i = atomicAdd(a.rtcur, 1u);
    if (i < a.rtcap) a.rtrace[i] = dcg_round_rec{a.rtag | flags, r, frontier, 0u, (uint64_t)wall_clock64()};
}

constexpr uint64_t kBigShift = 36;
constexpr uint64_t kBigMask = (1ull << kBigShift) - 1;

__device__ __forceinline__ uint32_t stamp_base(const TravArgs &a) {
    return a.step ? a.step->stamp[a.stamp_slot] : a.stamp0;
}

__device__ __forceinline__ void decode_g(const TravArgs &a, uint32_t g, uint32_t &q, uint32_t &v) {
    if (a.Q == 1) {
        q = 0;
        v = g;
    } else {
        q = g / a.n;
        v = g - q * a.n;
    }
}

// ---- program sets: compile-time dispatch on a query's program index ------------
// A set of min-lattice programs (include/dcg_program.hpp); a query's qapp entry
// is its program's index in the set.  The built-in set is
// ProgramSet<apps::SSSP, apps::BFS, apps::CC> (indices = dcg_app).
template <class... P>
struct ProgramSet {
    static constexpr uint32_t size = sizeof...(P);
};

template <class P>
__device__ __forceinline__ P make_program(const TravArgs &a, uint32_t q) {
    const api::Env env{a.sources[q], a.w.kind == W_SEEDED ? (uint32_t)DCG_W_SEEDED : (uint32_t)DCG_W_REFERENCE,
                       a.w.wmax, a.w.seed};
    return P(env);
}

template <class F, class P0, class... Ps>
__device__ __forceinline__ auto visit_program(uint32_t idx, const TravArgs &a, uint32_t q, const F &f) {
    if constexpr (sizeof...(Ps) == 0) {
        return f(make_program<P0>(a, q));
    } else {
        if (idx == 0) return f(make_program<P0>(a, q));
        return visit_program<F, Ps...>(idx - 1, a, q, f);
    }
}

// Visit<PS>::on(a, q, f) calls f(program of query q).
template <class PS>
struct Visit;
template <class... P>
struct Visit<ProgramSet<P...>> {
    template <class F>
    __device__ __forceinline__ static auto on(const TravArgs &a, uint32_t q, const F &f) {
        return visit_program<F, P...>(sizeof...(P) == 1 ? 0u : a.qapp[q], a, q, f);
    }
};

template <class App>
using app_t = typename std::decay<App>::type;

template <class PS>
__device__ __forceinline__ bool rooted_query(const TravArgs &a, uint32_t q) {
    return Visit<PS>::on(a, q, [](const auto &app) -> bool { return app_t<decltype(app)>::kRootedDeletionWitness; });
}

// The (value, depth) word an edge s -> d offers under a rooted-witness program:
// the program's candidate value, and a depth that counts hops inside one value
// class: the source's depth + 1 when the edge keeps the value (CC: (label,
// depth + 1), the source's word + 1), 1 when it raises it.  So the offer never
// grows when the source's word shrinks (a source that gets a smaller value over
// a longer path still offers depth 1 over an edge that raises the value), the
// chaotic relaxation converges to a fixpoint every word of which some
// in-neighbour offers, and (value, depth) strictly grows along parents: the
// parents form a tree rooted at the seeds (depth 0).
template <class App>
__device__ __forceinline__ uint64_t rooted_offer(const App &app, uint32_t s, uint64_t sv, uint32_t d) {
    const uint32_t vs = (uint32_t)(sv >> 32);
    if (vs == kInf) return kInfState;
    const uint32_t c = app.AccumulateBuffer(s, d, app.TraversalEdgeWeight(s, d), vs);
    if (c == kInf) return kInfState;
    const uint32_t depth = c == vs ? (uint32_t)sv + 1u : 1u;
    return ((uint64_t)c << 32) | depth;
}

// Dependency test of the delete stage, claiming a dependent vertex once.
// A destination slot's state word and parent, loaded ahead of the edge
// operation so a wave issues the loads of several 64-edge steps together.
// States only decrease (and claims only set parents to INF), so a copy that is
// stale by the time the edge is examined only weakens the early-out tests; the
// atomics decide.
struct DstPre {
    uint64_t st;
    uint32_t par;
};
__device__ __forceinline__ DstPre load_pre(const TravArgs &a, uint32_t q, uint32_t x) {
    const VSlot s = a.vs[(uint64_t)q * a.n + x];
    return DstPre{s.st, s.par};
}

// Rooted-witness programs: the parent edge, or an offer equal to the vertex's
// (value, depth) word while the parent's offer is not (the reference's CC marks
// every equal-label successor instead, hybrid_cc.cu:60-64); the claim resets
// the parent entry.  Other programs: the program's IsDeletionDependency
// (api.cuh:43-49) with candidate = AccumulateBuffer over DeletionEdgeWeight;
// parent_tight is evaluated only when the candidate equals the destination's
// value (the one case the reference's rule reads it); the claim resets the
// parent half of the state word.  `sv` is the state word of the edge's source.
template <class PS>
__device__ __forceinline__ bool invalidate_op(const TravArgs &a, uint32_t q, uint32_t s, uint64_t sv, uint32_t d,
                                              uint32_t &out_g, bool has_pre = false, DstPre pre = {}) {
    const uint64_t qoff = (uint64_t)q * a.n;
    const uint64_t gd = qoff + d;
    return Visit<PS>::on(a, q, [&](const auto &app) -> bool {
        using App = app_t<decltype(app)>;
        if constexpr (App::kRootedDeletionWitness) {
            const uint32_t pc = has_pre ? pre.par : a.vs[gd].par;
            if (pc == kInf) return false;   // a root or already claimed
            const uint64_t sd = has_pre ? pre.st : a.vs[gd].st;
            bool dep;
            if (pc == s) dep = true;
            else if (rooted_offer(app, s, sv, d) == sd) dep = rooted_offer(app, pc, a.vs[qoff + pc].st, d) != sd;
            else dep = false;
            if (!dep) return false;
            if (atomicExch(&a.vs[gd].par, kInf) == kInf) return false;
            out_g = (uint32_t)gd;
            return true;
        } else {
            const uint64_t sd = has_pre ? pre.st : a.vs[gd].st;
            const uint32_t vd = (uint32_t)(sd >> 32), pd = (uint32_t)sd;
            if (d == a.sources[q] || pd == kInf) return false;
            const uint32_t vs = (uint32_t)(sv >> 32);
            const uint32_t cand = vs == kInf ? kInf : app.AccumulateBuffer(s, d, app.DeletionEdgeWeight(s, d), vs);
            bool ptight = false;
            if (pd != s && vs != kInf && cand == vd) {
                const uint32_t vp = (uint32_t)(a.vs[qoff + pd].st >> 32);
                ptight = vp != kInf && app.AccumulateBuffer(pd, d, app.DeletionEdgeWeight(pd, d), vp) == vd;
            }
            if (!App::IsDeletionDependency(s, d, vs, vd, (uint64_t)cand, pd, ptight)) return false;
            if (atomicExch(reinterpret_cast<uint32_t *>(&a.vs[gd].st), kInf) == kInf) return false;
            out_g = (uint32_t)gd;
            return true;
        }
    });
}

// Relax s -> d.  Returns 0 (no change to publish), 1 (d joins the next
// frontier) or 2 (d joins the far pile: bucketed closures only).  Rooted
// programs: (value, depth + 1) with one 64-bit atomicMin, the parent entry
// follows the winning word; others: (value, parent) with one 64-bit atomicMin,
// so the smallest source wins a tie (framework.cuh:229-244).
template <class PS>
__device__ __forceinline__ uint32_t push_op(const TravArgs &a, uint32_t q, uint32_t s, uint64_t sv, uint32_t d,
                                            uint32_t stamp, uint32_t &out_g, bool bucket = true,
                                            bool has_pre = false, DstPre pre = {}) {
    const uint64_t gd = (uint64_t)q * a.n + d;
    return Visit<PS>::on(a, q, [&](const auto &app) -> uint32_t {
        using App = app_t<decltype(app)>;
        if constexpr (App::kRootedDeletionWitness) {
            const uint64_t cand = rooted_offer(app, s, sv, d);
            if (cand >= (has_pre ? pre.st : a.vs[gd].st)) return 0;
            const uint64_t old = atomicMin((unsigned long long *)&a.vs[gd].st, (unsigned long long)cand);
            if (cand >= old) return 0;
            a.vs[gd].par = s;
            if (atomicExch(&a.vs[gd].tick, stamp) == stamp) return 0;
            a.cc_dirty[gd >> kDirtyLog2] = 1;
            out_g = (uint32_t)gd;
            return 1;
        } else {
            const uint32_t vs = (uint32_t)(sv >> 32);
            const uint32_t c = app.AccumulateBuffer(s, d, app.TraversalEdgeWeight(s, d), vs);
            if (c == kInf) return 0;
            const uint64_t cur = has_pre ? pre.st : a.vs[gd].st;
            if (c >= (uint32_t)(cur >> 32)) return 0;   // (a CPU-owned vertex's copy here only gets lower on the CPU)
            if (a.owner && a.owner[d]) {
                const unsigned long long p = ((unsigned long long)c << 32) | s;
                if (atomicMin(&a.bprop[gd], p) == ~0ull) {
                    const uint32_t i = atomicAdd(a.bcount, 1u);
                    if (i < a.bcap) a.blist[i] = (uint32_t)gd;
                    else a.ctl->overflow = 1;
                }
                return 0;
            }
            const uint64_t old =
                atomicMin((unsigned long long *)&a.vs[gd].st, (unsigned long long)(((uint64_t)c << 32) | s));
            if (c >= (uint32_t)(old >> 32)) return 0;
            if (bucket && a.far0 && c >= a.thr[q]) {
                const uint32_t ep = *(volatile uint32_t *)a.epoch;
                if (atomicExch(&a.fstamp[gd], ep) == ep) return 0;   // already in this epoch's pile
                out_g = (uint32_t)gd;
                return 2;
            }
            if (atomicExch(&a.vs[gd].tick, stamp) == stamp) return 0;
            out_g = (uint32_t)gd;
            return 1;
        }
    });
}

// Pull (non-rooted programs): the row owner `d` is the destination, `x` an
// in-neighbour; the packed (value, parent) candidate or ~0.
template <class PS>
__device__ __forceinline__ uint64_t pull_cand(const TravArgs &a, uint32_t q, uint32_t d, uint32_t x,
                                             bool has_pre = false, DstPre pre = {}) {
    const uint64_t qoff = (uint64_t)q * a.n;
    return Visit<PS>::on(a, q, [&](const auto &app) -> uint64_t {
        const uint32_t vx = (uint32_t)((has_pre ? pre.st : a.vs[qoff + x].st) >> 32);
        if (vx == kInf) return ~0ull;
        const uint32_t c = app.AccumulateBuffer(x, d, app.TraversalEdgeWeight(x, d), vx);
        if (c == kInf) return ~0ull;
        return ((uint64_t)c << 32) | x;
    });
}

__device__ __forceinline__ void pull_apply(const TravArgs &a, uint32_t q, uint32_t d, uint64_t packed) {
    if (packed == ~0ull) return;
    const uint64_t gd = (uint64_t)q * a.n + d;
    if (packed >= a.vs[gd].st) return;
    atomicMin((unsigned long long *)&a.vs[gd].st, (unsigned long long)packed);
}

// Pull of one in-edge x -> d: rooted programs take x's offer and its parent
// entry; others min-combine the packed candidate.
template <class PS>
__device__ __forceinline__ void pull_edge(const TravArgs &a, uint32_t q, uint32_t d, uint32_t x,
                                          bool has_pre = false, DstPre pre = {}) {
    const uint64_t qoff = (uint64_t)q * a.n;
    Visit<PS>::on(a, q, [&](const auto &app) -> int {
        using App = app_t<decltype(app)>;
        if constexpr (App::kRootedDeletionWitness) {
            const uint64_t cand = rooted_offer(app, x, has_pre ? pre.st : a.vs[qoff + x].st, d);
            const uint64_t gd = qoff + d;
            if (cand >= a.vs[gd].st) return 0;
            const uint64_t old = atomicMin((unsigned long long *)&a.vs[gd].st, (unsigned long long)cand);
            if (cand < old) a.vs[gd].par = x;
        } else {
            pull_apply(a, q, d, pull_cand<PS>(a, q, d, x, has_pre, pre));
        }
        return 0;
    });
}

// PageRank's Frontier::Add (include/framework/pr_residual.cuh:21-30) through the
// program's AccumulateBuffer / IsActiveNode hooks.
__device__ __forceinline__ bool pr_add(const TravArgs &a, uint32_t d, float c, uint32_t &out_g) {
    const apps::PageRank pr(a.eps, 0.f);
    const float old = pr.AccumulateBuffer(&a.res[d], c);
    const float val = old + c;
    if (!isfinite(val)) atomicExch((unsigned int *)&a.ctl->bad, 1u);
    if (!pr.IsActiveNode(d, old, 0.f) && pr.IsActiveNode(d, val, 0.f) && atomicCAS(&a.queued[d], 0u, 1u) == 0u) {
        out_g = d;
        return true;
    }
    return false;
}

template <class PS, uint32_t MODE>
// 0: nothing to publish, 1: next frontier, 2: far pile (push only).
__device__ __forceinline__ uint32_t edge_op(const TravArgs &a, uint32_t q, uint32_t owner, uint64_t val, uint32_t x,
                                            uint32_t stamp, uint32_t &out_g, bool has_pre = false, DstPre pre = {}) {
    if (x & kTomb) return 0;   // deleted entry of a large row
    if (MODE == M_PUSH) return push_op<PS>(a, q, owner, val, x, stamp, out_g, true, has_pre, pre);
    if (MODE == M_INVALIDATE) return invalidate_op<PS>(a, q, owner, val, x, out_g, has_pre, pre) ? 1u : 0u;
    if (MODE == M_PR || MODE == M_PR_AMEND) return pr_add(a, x, __uint_as_float((uint32_t)val), out_g) ? 1u : 0u;
    pull_edge<PS>(a, q, owner, x, has_pre, pre);
    return 0;
}

// The edge loops of the min-lattice modes are software-pipelined: a wave loads
// the row entries of the step after next and the destination slots of the next
// step before it runs the current step's operations, so the two dependent loads
// of a step overlap the work of the steps before it.  PageRank's residual adds
// read no slot.
template <uint32_t MODE>
constexpr bool preloads() {
    return MODE == M_PUSH || MODE == M_INVALIDATE || MODE == M_PULL;
}

// Append the lanes' results to the next frontier (and, while invalidating, to
// the affected list) with one atomic per wave.
template <uint32_t MODE>
__device__ __forceinline__ void emit(const TravArgs &a, uint32_t *outq, uint32_t *count, bool want, uint32_t g) {
    const uint64_t mask = __ballot(want);
    if (!mask) return;
    const int leader = __ffsll((unsigned long long)mask) - 1;
    const uint32_t lane = lane_id();
    uint32_t base = 0, abase = 0;
    if ((int)lane == leader) {
        base = atomicAdd(count, (uint32_t)__popcll(mask));
        if (MODE == M_INVALIDATE) abase = atomicAdd(&a.ctl->aff_count, (uint32_t)__popcll(mask));
    }
    base = __shfl(base, leader, kWave);
    const uint32_t r = lanes_below(mask);
    if (want) {
        if (base + r < a.qcap) outq[base + r] = g;
        else a.ctl->overflow = 1;
    }
    if (MODE == M_INVALIDATE) {
        abase = __shfl(abase, leader, kWave);
        if (want) {
            if (abase + r < a.affcap) a.aff[abase + r] = g;
            else a.ctl->overflow = 1;
        }
    }
}

// Append to one of the far piles with one atomic per wave.
__device__ __forceinline__ void emit_to(const TravArgs &a, uint32_t *pile, uint32_t *count, bool want, uint32_t g) {
    const uint64_t mask = __ballot(want);
    if (!mask) return;
    const int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if ((int)lane_id() == leader) base = atomicAdd(count, (uint32_t)__popcll(mask));
    base = __shfl(base, leader, kWave);
    if (want) {
        const uint32_t pos = base + lanes_below(mask);
        if (pos < a.qcap) pile[pos] = g;
        else a.ctl->overflow = 1;
    }
}

__device__ __forceinline__ void emit_far(const TravArgs &a, bool want, uint32_t g) {
    if (!__ballot(want)) return;
    const uint32_t fp = *(volatile uint32_t *)&a.ctl->far_par;
    emit_to(a, fp ? a.far1 : a.far0, &a.ctl->far_count[fp], want, g);
}

// A wave's private staging run in LDS: entries wait there and move to a global
// list in one atomic on the list's counter per flush (a flush when the run
// could not take another 64 entries, and at the end of the phase), so the
// counter that every wave of the grid shares sees one atomic per ~200 entries.
constexpr uint32_t kStage = 256;
struct Stage {
    uint32_t *buf;   // kStage entries of this wave's LDS
    uint32_t n;      // entries waiting (wave-uniform)
};

// Move the staged entries to list[count...] (and, while invalidating, to the
// affected list as well).
template <uint32_t MODE>
__device__ __forceinline__ void stage_flush(const TravArgs &a, Stage &s, uint32_t *list, uint32_t *count) {
    if (!s.n) return;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const uint32_t lane = lane_id();
    uint32_t base = 0, abase = 0;
    if (lane == 0) {
        base = atomicAdd(count, s.n);
        if (MODE == M_INVALIDATE) abase = atomicAdd(&a.ctl->aff_count, s.n);
    }
    base = __shfl(base, 0, kWave);
    if (MODE == M_INVALIDATE) abase = __shfl(abase, 0, kWave);
    for (uint32_t i = lane; i < s.n; i += kWave) {
        const uint32_t g = s.buf[i];
        if (base + i < a.qcap) list[base + i] = g;
        else a.ctl->overflow = 1;
        if (MODE == M_INVALIDATE) {
            if (abase + i < a.affcap) a.aff[abase + i] = g;
            else a.ctl->overflow = 1;
        }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    s.n = 0;
}

// Stage the ballot's entries (mask = ballot(want)); flush first when the run
// could not take them.
template <uint32_t MODE>
__device__ __forceinline__ void stage_put(const TravArgs &a, Stage &s, uint64_t mask, bool want, uint32_t g,
                                          uint32_t *list, uint32_t *count) {
    if (s.n + kWave > kStage) stage_flush<MODE>(a, s, list, count);
    if (want) s.buf[s.n + lanes_below(mask)] = g;
    s.n += (uint32_t)__popcll(mask);
}

// The far pile that appends go to in this round.
__device__ __forceinline__ uint32_t *far_pile(const TravArgs &a, uint32_t *&count) {
    const uint32_t fp = *(volatile uint32_t *)&a.ctl->far_par;
    count = &a.ctl->far_count[fp];
    return fp ? a.far1 : a.far0;
}

// End of a phase: the wave's staged frontier and far-pile entries go out.
template <uint32_t MODE>
__device__ __forceinline__ void stage_close(const TravArgs &a, Stage &sn, Stage &sf, uint32_t *out, uint32_t *count) {
    stage_flush<MODE>(a, sn, out, count);
    if (sf.n) {
        uint32_t *fc;
        uint32_t *pile = far_pile(a, fc);
        stage_flush<M_PUSH>(a, sf, pile, fc);
    }
}

// Workgroup output buffer of a persistent launch: waves append new frontier
// (and far-pile) entries with LDS atomics and the workgroup publishes them
// with one global atomic per phase (flush_out), instead of one global atomic
// per wave and 64-edge chunk on a single counter.  A ballot that finds the
// buffer full leaves sentinels behind, which the flush compacts away, and goes
// to the wave's staging run (flushed by the wave itself).
constexpr uint32_t kOutCap = 8192, kFarCap = 4096;
struct OutLds {
    uint32_t n, nf, cn, cf, base, fbase, abase;
    uint32_t near[kOutCap];
    uint32_t far[kFarCap];
};

template <uint32_t MODE>
__device__ __forceinline__ void emit_b(OutLds *Op, Stage &st, const TravArgs &a, uint32_t *outq, uint32_t *count,
                                       bool want, uint32_t g) {
    if (!Op) {   // workgroup 0 alone: direct appends
        emit<MODE>(a, outq, count, want, g);
        return;
    }
    OutLds &O = *Op;
    const uint64_t mask = __ballot(want);
    if (!mask) return;
    const int leader = __ffsll((unsigned long long)mask) - 1;
    const uint32_t pc = (uint32_t)__popcll(mask), r = lanes_below(mask);
    uint32_t base = 0;
    if ((int)lane_id() == leader) base = atomicAdd(&O.n, pc);
    base = __shfl(base, leader, kWave);
    if (base + pc <= kOutCap) {
        if (want) O.near[base + r] = g;
        return;
    }
    if (want && base + r < kOutCap) O.near[base + r] = kInf;
    stage_put<MODE>(a, st, mask, want, g, outq, count);
}

__device__ __forceinline__ void emit_far_b(OutLds *Op, Stage &st, const TravArgs &a, bool want, uint32_t g) {
    if (!Op) {
        emit_far(a, want, g);
        return;
    }
    OutLds &O = *Op;
    const uint64_t mask = __ballot(want);
    if (!mask) return;
    const int leader = __ffsll((unsigned long long)mask) - 1;
    const uint32_t pc = (uint32_t)__popcll(mask), r = lanes_below(mask);
    uint32_t base = 0;
    if ((int)lane_id() == leader) base = atomicAdd(&O.nf, pc);
    base = __shfl(base, leader, kWave);
    if (base + pc <= kFarCap) {
        if (want) O.far[base + r] = g;
        return;
    }
    if (want && base + r < kFarCap) O.far[base + r] = kInf;
    uint32_t *fc;
    uint32_t *pile = far_pile(a, fc);
    stage_put<M_PUSH>(a, st, mask, want, g, pile, fc);
}

// Wave 0 drops the sentinels of an overflowed buffer in place; returns the count.
__device__ __forceinline__ uint32_t compact_lds(uint32_t *buf, uint32_t n) {
    uint32_t w = 0;
    for (uint32_t i0 = 0; i0 < n; i0 += kWave) {
        const uint32_t i = i0 + lane_id();
        const uint32_t x = i < n ? buf[i] : kInf;
        const uint64_t m = __ballot(x != kInf);
        __builtin_amdgcn_wave_barrier();
        if (x != kInf) buf[w + lanes_below(m)] = x;
        __builtin_amdgcn_wave_barrier();
        w += (uint32_t)__popcll(m);
    }
    return w;
}

// Publish the workgroup's buffered entries (all threads of the workgroup).
template <uint32_t MODE>
__device__ void flush_out(OutLds &O, const TravArgs &a, uint32_t *outq, uint32_t *count) {
    __syncthreads();
    if (threadIdx.x < kWave) {
        const uint32_t cn = O.n <= kOutCap ? O.n : compact_lds(O.near, kOutCap);
        const uint32_t cf = O.nf <= kFarCap ? O.nf : compact_lds(O.far, kFarCap);
        if (threadIdx.x == 0) {
            O.cn = cn;
            O.cf = cf;
            O.base = cn ? atomicAdd(count, cn) : 0u;
            if (MODE == M_INVALIDATE) O.abase = cn ? atomicAdd(&a.ctl->aff_count, cn) : 0u;
            if (cf) {
                const uint32_t fp = *(volatile uint32_t *)&a.ctl->far_par;
                O.fbase = atomicAdd(&a.ctl->far_count[fp], cf) | (fp << 31);
            }
        }
    }
    __syncthreads();
    const uint32_t cn = O.cn, cf = O.cf;
    for (uint32_t i = threadIdx.x; i < cn; i += blockDim.x) {
        const uint32_t g = O.near[i];
        if (O.base + i < a.qcap) outq[O.base + i] = g;
        else a.ctl->overflow = 1;
        if (MODE == M_INVALIDATE) {
            if (O.abase + i < a.affcap) a.aff[O.abase + i] = g;
            else a.ctl->overflow = 1;
        }
    }
    if (cf) {
        uint32_t *pile = (O.fbase >> 31) ? a.far1 : a.far0;
        const uint32_t fb = O.fbase & 0x7FFFFFFFu;
        for (uint32_t i = threadIdx.x; i < cf; i += blockDim.x) {
            if (fb + i < a.qcap) pile[fb + i] = O.far[i];
            else a.ctl->overflow = 1;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) O.n = O.nf = 0;
    __syncthreads();
}

struct WaveLds {
    uint32_t pref[65];
    uint64_t val[64];
    uint32_t owner[64];
    uint32_t q[64];
    const uint32_t *base[64];
    uint32_t stage_near[kStage];   // overflow staging runs (Stage)
    uint32_t stage_far[kStage];
};

// Work split of one group of waves (the whole grid, or workgroup 0 alone).
struct Team {
    uint32_t wave, nwaves;   // this wave's index and the team's wave count
    uint32_t tid, nthreads;  // thread index and count (for plain loops)
};

// Phase A of round r: frontier entries in batches of 64 (the first batch of
// every wave is static; further batches come from an atomic cursor), rows up to
// `big_row` expanded edge-balanced in the wave, larger rows listed as pieces.
constexpr uint32_t kClaim = 4;   // batches a wave claims per cursor atomic

// Entries queued for the round in slot `slot`: the counter may pass the queue
// capacity when an append overflowed (the write was dropped and `overflow`
// set), so readers never index past qcap.
__device__ __forceinline__ uint32_t front_count(const TravArgs &a, uint32_t slot) {
    const uint32_t c = *(volatile uint32_t *)&a.ctl->count[slot];
    return c < a.qcap ? c : (uint32_t)a.qcap;
}

// One lane's frontier entry of a batch and its row: the entry, the row's
// extent (0: nothing to expand) and offset, and the value the row's edges carry.
struct BatchRow {
    uint32_t g, deg;
    uint64_t off, val;
};

// Load entry i's row (i >= cnt: none).  Issued one batch ahead of its use, so a
// state word may be older than the round's latest improvement of the vertex:
// such a vertex was improved in this round and is queued for the next one,
// which expands it with its new value (the same holds for a word read late).
template <uint32_t MODE>
__device__ __forceinline__ BatchRow batch_row(const TravArgs &a, uint32_t i, uint32_t cnt, uint32_t g) {
    BatchRow b{g, 0u, 0ull, 0ull};
    if (i >= cnt) return b;
    if (MODE == M_PR || MODE == M_PR_AMEND) {
        const RowDesc d = a.s.out[g];
        const float x = MODE == M_PR ? a.delta[i] : a.rank[g] * a.sign;
        const uint32_t live = row_live(d);
        if (live && x != 0.f) {
            b.deg = d.ext;
            b.off = d.off;
            b.val = __float_as_uint(0.85f * x / (float)live);
        }
    } else {
        uint32_t q, v;
        decode_g(a, g, q, v);
        const RowDesc d = MODE == M_PULL ? a.s.in[v] : a.s.out[v];
        b.val = MODE == M_PULL ? 0ull : a.vs[g].st;
        if (d.ext && (MODE == M_PULL || (uint32_t)(b.val >> 32) != kInf)) {
            b.deg = d.ext;
            b.off = d.off;
        }
    }
    return b;
}

template <class PS, uint32_t MODE>
__device__ void phase_a(const TravArgs &a, uint32_t r, const Team &t, WaveLds &L, OutLds *O, uint64_t &edges,
                        uint64_t &expansions) {
    const uint32_t cur = r % 3, nxt = (r + 1) % 3;
    const uint32_t lane = lane_id();
    const uint32_t *in = r == 0 ? a.in0 : ((r & 1) ? a.qa : a.qb);
    uint32_t *out = (r & 1) ? a.qb : a.qa;
    const uint32_t stamp = stamp_base(a) + r;
    const uint32_t cnt = front_count(a, cur);
    Stage sn{L.stage_near, 0}, sf{L.stage_far, 0};
    // the first batch of every wave is static; then kClaim per cursor atomic
    uint32_t spare = 0;
    auto claim = [&](uint32_t b) -> uint32_t {
        if (spare) {
            --spare;
            return b + 1;
        }
        uint32_t nb = 0;
        if (lane == 0) nb = t.nwaves + atomicAdd(&a.ctl->cursor[cur], kClaim);
        spare = kClaim - 1;
        return __shfl(nb, 0, kWave);
    };
    // Rows are loaded one batch ahead: the next batch is claimed and its entries
    // read before this batch's scan, its descriptors and state words before
    // this batch's edges, so their latency overlaps this batch's work.
    uint32_t batch = t.wave;
    const uint32_t i0 = batch * kWave + lane;
    BatchRow row = batch_row<MODE>(a, i0, cnt, i0 < cnt ? in[i0] : 0u);
    while (batch * kWave < cnt) {
        const uint32_t nbatch = claim(batch);
        const uint32_t ni = nbatch * kWave + lane;
        const uint32_t ng = ni < cnt ? in[ni] : 0u;
        const uint32_t g = row.g;
        uint32_t q = 0, v = g;
        if (MODE != M_PR && MODE != M_PR_AMEND) decode_g(a, g, q, v);
        uint32_t deg = row.deg;
        const uint64_t val = row.val, off = row.off;
        if (deg) mark_read(a.s, MODE == M_PULL ? a.s.hot_in : a.s.hot_out, v);
        expansions += (uint64_t)__popcll(__ballot(deg > 0));
        if (deg > a.big_row) {
            const uint64_t pieces = (deg + a.piece - 1) / a.piece;
            const unsigned long long p = atomicAdd(&a.ctl->big[cur], (1ull << kBigShift) | pieces);
            const uint64_t e = p >> kBigShift;
            if (e < a.bigcap) {
                BigEntry be;
                be.off = off;
                be.g = g;
                be.deg = deg;
                be.val = val;
                be.piece0 = (uint32_t)(p & kBigMask);
                a.big[e] = be;
            } else {
                a.ctl->overflow = 1;
            }
            deg = 0;
        }
        const uint32_t incl = wave_incl_scan(deg);
        const uint32_t total = __shfl(incl, 63, kWave);
        L.pref[lane] = incl - deg;
        L.val[lane] = val;
        L.owner[lane] = v;
        L.q[lane] = q;
        L.base[lane] = deg ? row_base(a.s, off) : nullptr;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        edges += total;
        const BatchRow nrow = batch_row<MODE>(a, ni, cnt, ng);   // in flight during this batch's edges
        // entry k of the wave's edges: its row entry and owner slot in L
        auto fetch_x = [&](uint32_t k, uint32_t &x, uint32_t &lo) {
            x = kTomb;   // no edge (a tombstone is skipped the same way)
            lo = 0;
            if (k < total) {
                uint32_t hi = 64;   // last owner with pref <= k
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (L.pref[mid] <= k) lo = mid;
                    else hi = mid;
                }
                x = L.base[lo][k - L.pref[lo]];
            }
        };
        auto fetch_pre = [&](uint32_t x, uint32_t lo) {
            return (preloads<MODE>() && !(x & kTomb)) ? load_pre(a, L.q[lo], x) : DstPre{};
        };
        uint32_t x1, lo1, x2, lo2;
        fetch_x(lane, x1, lo1);
        fetch_x(kWave + lane, x2, lo2);
        DstPre p1 = fetch_pre(x1, lo1);
        for (uint32_t k0 = 0; k0 < total; k0 += kWave) {
            uint32_t x3, lo3;
            fetch_x(k0 + 2 * kWave + lane, x3, lo3);
            const DstPre p2 = fetch_pre(x2, lo2);
            uint32_t kind = 0;
            uint32_t og = 0;
            if (!(x1 & kTomb))
                kind = edge_op<PS, MODE>(a, L.q[lo1], L.owner[lo1], L.val[lo1], x1, stamp, og, preloads<MODE>(), p1);
            if (MODE != M_PULL) emit_b<MODE>(O, sn, a, out, &a.ctl->count[nxt], kind == 1, og);
            if (MODE == M_PUSH && a.far0) emit_far_b(O, sf, a, kind == 2, og);
            x1 = x2;
            lo1 = lo2;
            p1 = p2;
            x2 = x3;
            lo2 = lo3;
        }
        __builtin_amdgcn_wave_barrier();
        batch = nbatch;
        row = nrow;
    }
    stage_close<MODE>(a, sn, sf, out, &a.ctl->count[nxt]);
}

// Phase B of round r: the pieces of the large rows (first piece of every wave
// static, then an atomic cursor).
template <class PS, uint32_t MODE>
__device__ void phase_b(const TravArgs &a, uint32_t r, const Team &t, WaveLds &L, OutLds *O, uint64_t &edges) {
    const uint32_t cur = r % 3, nxt = (r + 1) % 3;
    const uint32_t lane = lane_id();
    uint32_t *out = (r & 1) ? a.qb : a.qa;
    const uint32_t stamp = stamp_base(a) + r;
    Stage sn{L.stage_near, 0}, sf{L.stage_far, 0};
    const unsigned long long bp = *(volatile unsigned long long *)&a.ctl->big[cur];
    uint64_t nb = bp >> kBigShift;
    if (nb > a.bigcap) nb = a.bigcap;
    const uint64_t np = bp & kBigMask;
    if (nb == 0) return;
    uint32_t p = t.wave;
    while (p < np) {
        uint64_t lo = 0, hi = nb;
        while (hi - lo > 1) {
            const uint64_t mid = (lo + hi) >> 1;
            if (a.big[mid].piece0 <= p) lo = mid;
            else hi = mid;
        }
        const BigEntry be = a.big[lo];
        const uint32_t start = (p - be.piece0) * a.piece;
        if (start < be.deg) {
            const uint32_t len = be.deg - start < a.piece ? be.deg - start : a.piece;
            const uint32_t *base = row_base(a.s, be.off + start);
            uint32_t q = 0, v = 0;
            if (MODE == M_PR || MODE == M_PR_AMEND) v = be.g;
            else decode_g(a, be.g, q, v);
            edges += lane == 0 ? len : 0;
            // pipelined as in phase A: entries two steps ahead, slots one step ahead
            auto fetch_x = [&](uint32_t k) { return k < len ? base[k] : kTomb; };
            auto fetch_pre = [&](uint32_t x) {
                return (preloads<MODE>() && !(x & kTomb)) ? load_pre(a, q, x) : DstPre{};
            };
            uint32_t x1 = fetch_x(lane), x2 = fetch_x(kWave + lane);
            DstPre p1 = fetch_pre(x1);
            if (MODE == M_PULL && rooted_query<PS>(a, q)) {
                for (uint32_t k0 = 0; k0 < len; k0 += kWave) {
                    const uint32_t x3 = fetch_x(k0 + 2 * kWave + lane);
                    const DstPre p2 = fetch_pre(x2);
                    if (!(x1 & kTomb)) pull_edge<PS>(a, q, v, x1, true, p1);
                    x1 = x2;
                    p1 = p2;
                    x2 = x3;
                }
            } else if (MODE == M_PULL) {
                uint64_t best = ~0ull;
                for (uint32_t k0 = 0; k0 < len; k0 += kWave) {
                    const uint32_t x3 = fetch_x(k0 + 2 * kWave + lane);
                    const DstPre p2 = fetch_pre(x2);
                    const uint64_t c = (x1 & kTomb) ? ~0ull : pull_cand<PS>(a, q, v, x1, true, p1);
                    best = c < best ? c : best;
                    x1 = x2;
                    p1 = p2;
                    x2 = x3;
                }
                for (int d = 32; d >= 1; d >>= 1) {
                    const uint64_t o = __shfl_xor(best, d, kWave);
                    best = o < best ? o : best;
                }
                if (lane == 0) pull_apply(a, q, v, best);
            } else {
                for (uint32_t k0 = 0; k0 < len; k0 += kWave) {
                    const uint32_t x3 = fetch_x(k0 + 2 * kWave + lane);
                    const DstPre p2 = fetch_pre(x2);
                    uint32_t kind = 0;
                    uint32_t og = 0;
                    if (!(x1 & kTomb))
                        kind = edge_op<PS, MODE>(a, q, v, be.val, x1, stamp, og, preloads<MODE>(), p1);
                    emit_b<MODE>(O, sn, a, out, &a.ctl->count[nxt], kind == 1, og);
                    if (MODE == M_PUSH && a.far0) emit_far_b(O, sf, a, kind == 2, og);
                    x1 = x2;
                    p1 = p2;
                    x2 = x3;
                }
            }
        }
        uint32_t np2 = 0;
        if (lane == 0) np2 = t.nwaves + atomicAdd(&a.ctl->big_cursor[cur], 1u);
        p = __shfl(np2, 0, kWave);
    }
    stage_close<MODE>(a, sn, sf, out, &a.ctl->count[nxt]);
}

// Start of round r: ring counters for rounds r+1 / r+2, and PR's consume step.
template <class PS, uint32_t MODE>
__device__ void round_open(const TravArgs &a, uint32_t r, const Team &t) {
    const uint32_t cur = r % 3, nxt = (r + 1) % 3, nn = (r + 2) % 3;
    if (t.tid == 0) {
        a.ctl->count[nn] = 0;
        a.ctl->cursor[nxt] = 0;
        a.ctl->big[nxt] = 0;
        a.ctl->big_cursor[nxt] = 0;
        a.ctl->rounds = r + 1;
    }
    if (MODE == M_PR) {
        // Consume (pr_residual.cuh:53-69): fold residuals above eps into ranks
        // through the program's IsActiveNode / CombineValueBuffer hooks.
        const apps::PageRank pr(a.eps, 0.f);
        const uint32_t cnt = front_count(a, cur);
        const uint32_t *in = r == 0 ? a.in0 : ((r & 1) ? a.qa : a.qb);
        for (uint32_t i = t.tid; i < cnt; i += t.nthreads) {
            const uint32_t v = in[i];
            const float res = a.res[v];
            a.queued[v] = 0;
            float dl = 0.f;
            if (!isfinite(res) || !isfinite(a.rank[v])) atomicExch((unsigned int *)&a.ctl->bad, 1u);
            if (pr.IsActiveNode(v, res, 0.f)) dl = pr.CombineValueBuffer(v, &a.rank[v], &a.res[v]);
            a.delta[i] = dl;
        }
    }
}

__device__ __forceinline__ bool pieces_pending(const TravArgs &a, uint32_t r) {
    return (*(volatile unsigned long long *)&a.ctl->big[r % 3] & kBigMask) != 0;
}

// Grid barrier of a persistent launch (co-resident workgroups: the launch is
// cooperative).  Agent-scope release before arriving and acquire after the
// generation moves publish every workgroup's writes to all XCDs.  Grids of
// kBarTreeMin or more workgroups arrive in two levels (the group counters
// spread the atomics that one counter would serialise).  Returns false when the
// barrier gave up: some workgroup waited longer than a.bar_spin wall-clock
// ticks (~10 s by default) and set bar_timeout.  Every caller then ends its
// round and the launch, so no workgroup reads entries that workgroups which
// never arrived would have written; the host reports the timeout as an error.
__device__ __forceinline__ uint32_t ld_agent(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t add_agent(uint32_t *p, uint32_t v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool gsync(const TravArgs &a) {
    __shared__ uint32_t passed;
    __syncthreads();
    if (threadIdx.x == 0) {
        Ctl *c = a.ctl;
        uint32_t ok = 0;
        if (!ld_agent(&c->bar_timeout)) {
            const uint32_t g = ld_agent(&c->bar_gen.w);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            uint32_t top = gridDim.x;
            bool arrive_top = true;
            if (gridDim.x >= kBarTreeMin) {
                const uint32_t grp = blockIdx.x % kBarGroups;
                const uint32_t members = gridDim.x / kBarGroups + (grp < gridDim.x % kBarGroups ? 1u : 0u);
                top = kBarGroups;
                arrive_top = add_agent(&c->bar_group[grp].w, 1u) == members - 1;
                if (arrive_top) {
                    st_agent(&c->bar_group[grp].w, 0u);
                    // the group's writes travel on with this workgroup's arrival
                    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
                }
            }
            if (arrive_top && add_agent(&c->bar_arrive.w, 1u) == top - 1) {
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                st_agent(&c->bar_arrive.w, 0u);
                __hip_atomic_store(&c->bar_gen.w, g + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                const uint64_t t0 = wall_clock64();
                while (ld_agent(&c->bar_gen.w) == g) {
                    __builtin_amdgcn_s_sleep(1);
                    if (wall_clock64() - t0 > a.bar_spin) {
                        st_agent(&c->bar_timeout, 1u);
                        break;
                    }
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            ok = ld_agent(&c->bar_timeout) ? 0u : 1u;
        }
        passed = ok;
    }
    __syncthreads();
    return passed != 0;
}

__device__ __forceinline__ uint32_t thr_of(uint32_t minv, uint32_t delta) {
    if (delta == kInf) return kInf;
    const uint64_t t = (uint64_t)minv + delta;
    return t >= kInf ? kInf : (uint32_t)t;
}

// Bucket advance of a push closure whose near frontier drained (round r found
// no entries) while the far pile holds some: every query's threshold moves to
// its smallest far value + its bucket width, entries now below it become round
// r's frontier and the rest move to the other pile for the next epoch.  Called
// by the whole grid; returns false when the far pile is empty (closure done).
__device__ inline bool far_advance(const TravArgs &a, uint32_t r, const Team &t, WaveLds &L) {
    const uint32_t fp = *(volatile uint32_t *)&a.ctl->far_par;
    uint32_t fc = *(volatile uint32_t *)&a.ctl->far_count[fp];
    if (fc == 0 || r == 0) return false;
    if (fc > a.qcap) fc = (uint32_t)a.qcap;
    const uint32_t *src = fp ? a.far1 : a.far0;
    uint32_t *dst = fp ? a.far0 : a.far1;
    uint32_t *mf = a.minfar + (uint64_t)fp * a.Q;
    // per-query minima in LDS first, then one global atomic per workgroup and
    // query (one global atomic per entry on Q words serialised at the L2:
    // ~30 M per snapshot on the 65-query grid)
    __shared__ uint32_t qmin[kLdsQueries];
    const bool lds = a.Q <= kLdsQueries;
    if (lds) {
        for (uint32_t q = threadIdx.x; q < a.Q; q += blockDim.x) qmin[q] = kInf;
        __syncthreads();
    }
    for (uint32_t i = t.tid; i < fc; i += t.nthreads) {
        uint32_t q, v;
        const uint32_t g = src[i];
        decode_g(a, g, q, v);
        const uint32_t val = (uint32_t)(a.vs[g].st >> 32);
        if (lds) atomicMin(&qmin[q], val);
        else atomicMin(&mf[q], val);
    }
    if (lds) {
        __syncthreads();
        for (uint32_t q = threadIdx.x; q < a.Q; q += blockDim.x)
            if (qmin[q] != kInf) atomicMin(&mf[q], qmin[q]);
    }
    if (!gsync(a)) return true;   // the caller's next loop test ends the launch
    const uint32_t ep = *(volatile uint32_t *)a.epoch + 1;
    if (t.tid == 0) {
        a.ctl->far_count[fp] = 0;
        a.ctl->far_par = fp ^ 1;
        a.ctl->advances += 1;
        a.ctl->far_scanned += fc;
    }
    for (uint32_t q = t.tid; q < a.Q; q += t.nthreads) {
        a.thr[q] = thr_of(mf[q], a.fdelta[q]);
        a.minfar[(uint64_t)(fp ^ 1) * a.Q + q] = kInf;
    }
    uint32_t *in = (r & 1) ? a.qa : a.qb;   // round r's input queue
    const uint32_t lane = lane_id();
    // both lists through the wave's staging runs: one atomic per ~200 entries
    // on each list's counter
    Stage sn{L.stage_near, 0}, sf{L.stage_far, 0};
    uint32_t *ncount = &a.ctl->count[r % 3], *fcount = &a.ctl->far_count[fp ^ 1];
    for (uint32_t b = t.wave * kWave; b < fc; b += t.nwaves * kWave) {
        const uint32_t i = b + lane;
        bool near = false, far = false;
        uint32_t g = 0;
        if (i < fc) {
            uint32_t q, v;
            g = src[i];
            decode_g(a, g, q, v);
            near = (uint32_t)(a.vs[g].st >> 32) < thr_of(mf[q], a.fdelta[q]);
            far = !near;
            if (far) a.fstamp[g] = ep;
        }
        const uint64_t nm = __ballot(near), fm = __ballot(far);
        if (nm) stage_put<M_PUSH>(a, sn, nm, near, g, in, ncount);
        if (fm) stage_put<M_PUSH>(a, sf, fm, far, g, dst, fcount);
    }
    stage_flush<M_PUSH>(a, sn, in, ncount);
    stage_flush<M_PUSH>(a, sf, dst, fcount);
    if (t.tid == 0) *(volatile uint32_t *)a.epoch = ep;
    gsync(a);
    return true;
}

constexpr uint32_t kSoloPieces = kPBlock / kWave;   // more pending pieces than one per wave end a solo stretch

// One cooperative launch (1,024-thread workgroups) runs rounds until the
// frontier is empty, with the gsync grid barrier between rounds.  While the
// frontier is small, workgroup 0 runs rounds by itself with workgroup barriers
// (the rest of the grid waits at one grid barrier); otherwise all workgroups
// share each round, separated by grid barriers (a second one only when the
// round queued pieces of large rows).
// TAG only separates launch names: 1 for static convergence (dcg_start, the
// autotuner), 0 for incremental closures, so profiles list them apart.
template <class PS, uint32_t MODE, uint32_t TAG = 0>
__global__ __launch_bounds__(kPBlock) void k_persistent(TravArgs a) {
    __shared__ WaveLds lds_all[kPBlock / kWave];
    __shared__ OutLds O;
    __shared__ WgSum<2> tot;
    if (threadIdx.x == 0) {
        O.n = O.nf = 0;
        atomicMin(&a.ctl->t_first, (unsigned long long)wall_clock64());
    }
    wg_sum_init(tot);
    __syncthreads();
    // publish round rr's buffered entries (modes that emit; whole-grid rounds
    // only: workgroup 0 alone appends directly); a workgroup barrier either way
    auto flush = [&](uint32_t rr) {
        if (MODE != M_PULL)
            flush_out<MODE>(O, a, (rr & 1) ? a.qb : a.qa, &a.ctl->count[(rr + 1) % 3]);
        else
            __syncthreads();
    };
    WaveLds &L = lds_all[threadIdx.x / kWave];
    const Team gt{(blockIdx.x * blockDim.x + threadIdx.x) / kWave, gridDim.x * blockDim.x / kWave,
                  blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x};
    const Team bt{threadIdx.x / kWave, blockDim.x / kWave, threadIdx.x, blockDim.x};
    uint64_t edges = 0, expansions = 0;
    uint32_t r = 0, idle_advances = 0;
    while (r < a.max_rounds) {
        if (*(volatile uint32_t *)&a.ctl->bar_timeout) break;
        const uint32_t cnt = *(volatile uint32_t *)&a.ctl->count[r % 3];
        if (cnt == 0) {
            // an advance always moves the smallest far entry below the new
            // threshold; two in a row without a round would mean none moved
            if (MODE == M_PUSH && a.far0 && idle_advances < 2 && far_advance(a, r, gt, L)) {
                ++idle_advances;
                if (blockIdx.x == 0 && threadIdx.x == 0)
                    trace_round(a, r, *(volatile uint32_t *)&a.ctl->count[r % 3], DCG_ROUND_ADVANCE);
                continue;
            }
            if (idle_advances >= 2 && threadIdx.x == 0 && blockIdx.x == 0) a.ctl->overflow = 1;
            break;
        }
        idle_advances = 0;
        if (cnt <= a.solo_max && gridDim.x > 1) {
            if (blockIdx.x == 0) {
                // solo stretch: rounds on workgroup 0 until the frontier grows,
                // empties, or a round queues many large-row pieces
                uint32_t resume_b = 0;
                for (;;) {
                    if (threadIdx.x == 0) trace_round(a, r, front_count(a, r % 3), DCG_ROUND_SOLO);
                    round_open<PS, MODE>(a, r, bt);
                    __syncthreads();
                    phase_a<PS, MODE>(a, r, bt, L, nullptr, edges, expansions);
                    __syncthreads();
                    const uint64_t np = *(volatile unsigned long long *)&a.ctl->big[r % 3] & kBigMask;
                    if (np > kSoloPieces) {
                        resume_b = 1;   // the whole grid takes this round's pieces
                        break;
                    }
                    if (np) {
                        phase_b<PS, MODE>(a, r, bt, L, nullptr, edges);
                        __syncthreads();
                    }
                    ++r;
                    const uint32_t c2 = *(volatile uint32_t *)&a.ctl->count[r % 3];
                    if (c2 == 0 || c2 > a.solo_max || r >= a.max_rounds) break;
                }
                if (threadIdx.x == 0) {
                    a.ctl->solo_round = r;
                    a.ctl->solo_resume_b = resume_b;
                }
            }
            if (!gsync(a)) break;
            r = *(volatile uint32_t *)&a.ctl->solo_round;
            if (*(volatile uint32_t *)&a.ctl->solo_resume_b) {
                phase_b<PS, MODE>(a, r, gt, L, &O, edges);
                flush(r);
                if (!gsync(a)) break;
                ++r;
            }
            continue;
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) trace_round(a, r, cnt, 0u);
        round_open<PS, MODE>(a, r, gt);
        if (MODE == M_PR && !gsync(a)) break;
        phase_a<PS, MODE>(a, r, gt, L, &O, edges, expansions);
        flush(r);
        if (!gsync(a)) break;
        if (pieces_pending(a, r)) {
            phase_b<PS, MODE>(a, r, gt, L, &O, edges);
            flush(r);
            if (!gsync(a)) break;
        }
        ++r;
    }
    // every lane of a wave holds the same totals; lane 0 reports them, one
    // global atomic per workgroup
    unsigned long long *const dst[2] = {&a.ctl->edges, &a.ctl->expansions};
    const unsigned long long x[2] = {edges, expansions};
    wg_sum_flush(tot, x, dst);
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(&a.ctl->t_last, (unsigned long long)wall_clock64());
    if (blockIdx.x == 0 && threadIdx.x == 0) trace_round(a, r, 0u, DCG_ROUND_END);
}

// ------------------------------------------------------------- plain kernels
// Initial state (driver.cuh / GetInitBuffer): a program's seed value at depth 0
// (rooted programs, parent INF) or with parent INF; IDENTITY elsewhere.  A
// program without kVertexSeeds seeds its query's source only.
template <class PS>
__global__ void k_init_state(TravArgs a) {
    const uint64_t total = (uint64_t)a.n * a.Q;
    for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < total;
         g += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t q = (uint32_t)(g / a.n), v = (uint32_t)(g - (uint64_t)q * a.n);
        a.vs[g].st = Visit<PS>::on(a, q, [&](const auto &app) -> uint64_t {
            using App = app_t<decltype(app)>;
            const uint32_t seed = (App::kVertexSeeds || v == a.sources[q]) ? app.GetInitBuffer(v) : kInf;
            if constexpr (App::kRootedDeletionWitness) {
                a.vs[g].par = kInf;
                return seed == kInf ? kInfState : ((uint64_t)seed << 32);
            } else {
                return seed == kInf ? kInfState : (((uint64_t)seed << 32) | kInf);
            }
        });
        a.vs[g].tick = 0;
    }
}

template <class PS>
__device__ __forceinline__ bool vertex_seeded(const TravArgs &a, uint32_t q) {
    return Visit<PS>::on(a, q, [](const auto &app) -> bool { return app_t<decltype(app)>::kVertexSeeds; });
}

// Initial frontier: a query of a vertex-seeded program starts at every vertex,
// any other at its source.  `slot0[q]` is the query's first position in q0
// (host prefix sum).
template <class PS>
__global__ void k_iota_sources(TravArgs a, uint32_t *q0, const uint64_t *slot0) {
    for (uint32_t q = 0; q < a.Q; ++q) {
        if (vertex_seeded<PS>(a, q)) {
            for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < a.n;
                 v += (uint64_t)gridDim.x * blockDim.x)
                q0[slot0[q] + v] = (uint32_t)((uint64_t)q * a.n + v);
        } else if (blockIdx.x == 0 && threadIdx.x == 0) {
            q0[slot0[q]] = (uint32_t)((uint64_t)q * a.n + a.sources[q]);
        }
    }
}

// Delete seeds (driver.cuh:394-441): test every deleted edge against every query.
template <class PS>
__global__ __launch_bounds__(1024) void k_seed_dels(TravArgs a) {
    __shared__ uint32_t wcnt[16], wbase;   // list positions: one atomic per workgroup and step
    const uint64_t count = a.step->count;
    const dcg_edge *recs = a.step->recs;
    const uint64_t total = count * a.Q;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t b = blockIdx.x * (uint64_t)blockDim.x; b < total; b += stride) {
        const uint64_t i = b + threadIdx.x;
        bool want = false;
        uint32_t og = 0;
        if (i < total) {
            const uint32_t q = (uint32_t)(i / count);
            const dcg_edge e = recs[i - (uint64_t)q * count];
            if (e.u < a.n && e.v < a.n) want = invalidate_op<PS>(a, q, e.u, a.vs[(uint64_t)q * a.n + e.u].st, e.v, og);
        }
        const uint32_t pos = wg_append(want, &a.ctl->aff_count, wcnt, wbase);
        if (want) {
            if (pos < a.affcap) a.aff[pos] = og;
            else a.ctl->overflow = 1;
        }
    }
}

// Reset affected values (driver.cuh:443-463): the program's seed (a rooted
// vertex-seeded program's vertex becomes its own root: CC's own label), else
// IDENTITY.  The affected list and its length come from a.aff / a.ctl.
template <class PS>
__global__ void k_reset_affected(TravArgs a) {
    const uint32_t count = a.ctl->aff_count;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t g = a.aff[i];
        uint32_t q, v;
        decode_g(a, g, q, v);
        a.vs[g].st = Visit<PS>::on(a, q, [&](const auto &app) -> uint64_t {
            using App = app_t<decltype(app)>;
            if constexpr (App::kRootedDeletionWitness && App::kVertexSeeds) {
                const uint32_t seed = app.GetInitBuffer(v);   // parent already INF
                return seed == kInf ? kInfState : ((uint64_t)seed << 32);
            } else {
                return kInfState;
            }
        });
    }
}

// Rooted programs write parents after the winning atomicMin of the state word,
// so two improvers of one vertex can leave the earlier one's parent behind (the
// invalidation rule treats such a parent as not tight, which only widens it).
// After a closure every rooted vertex it touched (tick >= stamp_lo, or on the
// affected list) has its parent checked: kept if its offer still equals the
// vertex's (value, depth) word, else replaced by the first in-neighbour whose
// offer does; depth 0 is a root (none), and an unreached vertex has none.
template <class PS>
__device__ __forceinline__ void cc_fix_one(const TravArgs &a, uint32_t g) {
    uint32_t q, v;
    decode_g(a, g, q, v);
    Visit<PS>::on(a, q, [&](const auto &app) -> int {
        using App = app_t<decltype(app)>;
        if constexpr (App::kRootedDeletionWitness) {
            const uint64_t qoff = (uint64_t)q * a.n;
            const uint64_t sv = a.vs[g].st;
            const uint32_t p = a.vs[g].par;
            if ((uint32_t)sv == 0 || (uint32_t)(sv >> 32) == kInf) {
                if (p != kInf) a.vs[g].par = kInf;
                return 0;
            }
            if (p != kInf && rooted_offer(app, p, a.vs[qoff + p].st, v) == sv) return 0;
            const RowDesc d = a.s.in[v];
            uint32_t best = kInf;
            for (uint32_t j = 0; j < d.ext; ++j) {
                const uint32_t x = *row_base(a.s, d.off + j);
                if (!(x & kTomb) && rooted_offer(app, x, a.vs[qoff + x].st, v) == sv) {
                    best = x;
                    break;
                }
            }
            a.vs[g].par = best;
        }
        return 0;
    });
}

// With a list: the listed slots (one thread each).  Without: query slots
// [base, base + n) stamped since stamp_lo, visiting only the 1,024-slot blocks
// that a rooted improvement marked (one workgroup per block); the marks are
// cleared after every rooted query of the stage was checked.
template <class PS>
__global__ void k_cc_fix_parents(TravArgs a, const uint32_t *list, const Ctl *list_ctl, uint32_t base) {
    const uint32_t stamp_lo = stamp_base(a);
    if (list) {
        const uint64_t total = (uint64_t)list_ctl->aff_count;
        for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total;
             i += (uint64_t)gridDim.x * blockDim.x)
            cc_fix_one<PS>(a, list[i]);
        return;
    }
    const uint64_t lo = base, hi = (uint64_t)base + a.n;
    for (uint64_t b = (lo >> kDirtyLog2) + blockIdx.x; (b << kDirtyLog2) < hi; b += gridDim.x) {
        if (!a.cc_dirty[b]) continue;   // the same byte for the whole workgroup
        for (uint32_t j = threadIdx.x; j < (1u << kDirtyLog2); j += blockDim.x) {
            const uint64_t g = (b << kDirtyLog2) + j;
            if (g < lo || g >= hi || a.vs[g].tick < stamp_lo) continue;
            cc_fix_one<PS>(a, (uint32_t)g);
        }
    }
}

// Insertion seeds (framework.cuh:297-370): relax every added edge for every query.
template <class PS>
__global__ __launch_bounds__(1024) void k_seed_adds(TravArgs a, uint32_t *outq) {
    __shared__ uint32_t wcnt[16], wbase;   // list positions: one atomic per workgroup and step
    const uint64_t count = a.step->count;
    const dcg_edge *recs = a.step->recs;
    const uint32_t stamp0 = stamp_base(a);
    const uint64_t total = count * a.Q;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t b = blockIdx.x * (uint64_t)blockDim.x; b < total; b += stride) {
        const uint64_t i = b + threadIdx.x;
        bool want = false;
        uint32_t og = 0;
        if (i < total) {
            const uint32_t q = (uint32_t)(i / count);
            const dcg_edge e = recs[i - (uint64_t)q * count];
            if (e.u < a.n && e.v < a.n) {
                const uint64_t sv = a.vs[(uint64_t)q * a.n + e.u].st;
                if ((uint32_t)(sv >> 32) != kInf) want = push_op<PS>(a, q, e.u, sv, e.v, stamp0, og, false) == 1;
            }
        }
        const uint32_t pos = wg_append(want, &a.ctl->aff_count, wcnt, wbase);
        if (want) {
            if (pos < a.qcap) outq[pos] = og;
            else a.ctl->overflow = 1;
        }
    }
}

// ---- program tables ------------------------------------------------------------
// Layout tag of TravArgs and Ctl: a table built against another layout is refused.
constexpr uint32_t kTravArgsTag = (uint32_t)(sizeof(TravArgs) << 16) | (uint32_t)sizeof(Ctl);

template <class P>
constexpr uint32_t program_flags() {
    return (P::kRootedDeletionWitness ? (uint32_t)DCG_PROGRAM_ROOTED : 0u) |
           (P::kVertexSeeds ? (uint32_t)DCG_PROGRAM_VERTEX_SEEDS : 0u);
}

template <class PS>
struct ProgramTable;
template <class... P>
struct ProgramTable<ProgramSet<P...>> {
    using PS = ProgramSet<P...>;
    static_assert(sizeof...(P) >= 1 && sizeof...(P) <= DCG_PROGRAM_MAX, "1..DCG_PROGRAM_MAX programs per set");
    static uint32_t bucket(uint32_t prog, uint32_t quarters, uint32_t wmax) {
        const uint32_t w[] = {P::OrderedBucketWidth(quarters, wmax)...};
        return prog < sizeof...(P) ? w[prog] : kInf;
    }
    static dcg_program make(const char *name) {
        dcg_program t{};
        t.abi = DCG_PROGRAM_ABI;
        t.num_programs = (uint32_t)sizeof...(P);
        const uint32_t f[] = {program_flags<P>()...};
        for (uint32_t i = 0; i < sizeof...(P); ++i) t.flags[i] = f[i];
        t.name = name;
        t.kernels[DCG_PK_PUSH] = (const void *)&k_persistent<PS, M_PUSH, 0>;
        t.kernels[DCG_PK_PULL] = (const void *)&k_persistent<PS, M_PULL, 0>;
        t.kernels[DCG_PK_INVALIDATE] = (const void *)&k_persistent<PS, M_INVALIDATE, 0>;
        t.kernels[DCG_PK_STATIC] = (const void *)&k_persistent<PS, M_PUSH, 1>;
        t.kernels[DCG_PK_SEED_DELS] = (const void *)&k_seed_dels<PS>;
        t.kernels[DCG_PK_SEED_ADDS] = (const void *)&k_seed_adds<PS>;
        t.kernels[DCG_PK_RESET] = (const void *)&k_reset_affected<PS>;
        t.kernels[DCG_PK_FIX_PARENTS] = (const void *)&k_cc_fix_parents<PS>;
        t.kernels[DCG_PK_INIT] = (const void *)&k_init_state<PS>;
        t.kernels[DCG_PK_FRONTIER] = (const void *)&k_iota_sources<PS>;
        t.bucket_width = &bucket;
        t.tag = kTravArgsTag;
        return t;
    }
};

}  // namespace dcg
