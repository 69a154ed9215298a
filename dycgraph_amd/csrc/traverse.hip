// traverse.hip -- persistent frontier kernels for the min-lattice vertex
// programs (SSSP, BFS, CC) and the PageRank residual push, plus the host
// orchestration of one snapshot's delete and add stages.
//
// Vertex state is one 64-bit word per (query, vertex).  SSSP/BFS: value in the
// high half, parent in the low half; a relaxation is one 64-bit atomicMin, so
// the value lattice (the reference's atomicMin + parent CAS loop,
// hybrid_sssp.cu:126-145) and the smallest-source tie-break of the pull repair
// (framework.cuh:229-244) come from the same instruction.  CC: label in the high
// half, hop depth from the label's root in the low half, parent in a separate
// array.  The label lattice is unchanged (label = smallest id that reaches the
// vertex), and the (label, depth) order makes parents a shortest-hop tree
// inside each label class, so a deletion invalidates only the part of that tree
// that lost its path to the root (a rooted witness, the exact counterpart of
// directed_label_witness.h) instead of every vertex with an equal label.
//
// One cooperative launch runs rounds until the frontier is empty:
//   phase A: waves take 64 frontier entries at a time; rows of <= 256 edges are
//            expanded edge-balanced inside the wave (wave64 prefix sum + binary
//            search of the owner); larger rows are queued as 512-edge pieces.
//   phase B: all waves share the pieces of the large rows.
// New frontier entries are appended with one atomic per wave (ballot + mbcnt).
#include "engine.hpp"
#include "mlp.hpp"

namespace dcg {

enum Mode : uint32_t { M_PUSH = 0, M_PULL = 1, M_INVALIDATE = 2, M_PR = 3, M_PR_AMEND = 4 };
constexpr uint32_t kDirtyLog2 = 10;   // CC parent-check block: 1,024 state slots
constexpr uint32_t kLdsQueries = 1024;   // queries whose far-pile minima a workgroup gathers in LDS

struct TravArgs {
    StoreView s;
    uint32_t n, Q;
    uint64_t *st;
    uint32_t *par;             // CC parents (q*n + v), null without CC queries
    uint32_t *tick;
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
};

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

__device__ __forceinline__ uint32_t qweight(const TravArgs &a, uint32_t app, uint32_t u, uint32_t v) {
    if (app == DCG_APP_BFS) return 1u;
    if (app == DCG_APP_CC) return 0u;
    return edge_weight(a.w, u, v);
}

// Dependency predicate of the delete stage (api.cuh:43-49): the parent edge,
// or a tight witness while the current parent is not tight.  For CC "tight"
// means equal label and depth + 1 (the reference's hybrid_cc.cu:60-64 marks
// every equal-label successor instead).  A dependent vertex is claimed once:
// SSSP/BFS set the parent half to INF, CC its parent entry.  `sv` is the
// state word of the edge's source.
__device__ __forceinline__ bool invalidate_op(const TravArgs &a, uint32_t q, uint32_t s, uint64_t sv, uint32_t d,
                                              uint32_t &out_g) {
    const uint64_t qoff = (uint64_t)q * a.n;
    const uint64_t gd = qoff + d;
    const uint64_t sd = a.st[gd];
    const uint32_t vd = (uint32_t)(sd >> 32), pd = (uint32_t)sd;
    const uint32_t vs = (uint32_t)(sv >> 32);
    const uint32_t app = a.qapp[q];
    bool dep;
    if (app == DCG_APP_CC) {
        const uint32_t pc = a.par[gd];
        if (pc == kInf) return false;   // a root (own label) or already claimed
        if (pc == s) dep = true;
        else if (sv + 1 == sd) dep = a.st[qoff + pc] + 1 != sd;
        else dep = false;
        if (!dep) return false;
        if (atomicExch(&a.par[gd], kInf) == kInf) return false;
        out_g = (uint32_t)gd;
        return true;
    } else {
        if (d == a.sources[q] || pd == kInf) return false;
        if (pd == s) {
            dep = true;
        } else if (vs != kInf && (uint64_t)vs + qweight(a, app, s, d) == (uint64_t)vd) {
            const uint32_t vp = (uint32_t)(a.st[qoff + pd] >> 32);
            const bool ptight = vp != kInf && (uint64_t)vp + qweight(a, app, pd, d) == (uint64_t)vd;
            dep = !ptight;
        } else {
            dep = false;
        }
    }
    if (!dep) return false;
    if (atomicExch(reinterpret_cast<uint32_t *>(&a.st[gd]), kInf) == kInf) return false;
    out_g = (uint32_t)gd;
    return true;
}

// Relax s -> d.  Returns 0 (no change to publish), 1 (d joins the next
// frontier) or 2 (d joins the far pile: bucketed closures only).
__device__ __forceinline__ uint32_t push_op(const TravArgs &a, uint32_t q, uint32_t s, uint64_t sv, uint32_t d,
                                            uint32_t stamp, uint32_t &out_g, bool bucket = true) {
    const uint32_t app = a.qapp[q];
    if (app == DCG_APP_CC) {
        // (label, depth + 1); the parent entry follows the winning word
        const uint64_t gd = (uint64_t)q * a.n + d;
        const uint64_t cand = sv + 1;
        if (cand >= a.st[gd]) return 0;
        const uint64_t old = atomicMin((unsigned long long *)&a.st[gd], (unsigned long long)cand);
        if (cand >= old) return 0;
        a.par[gd] = s;
        if (atomicExch(&a.tick[gd], stamp) == stamp) return 0;
        a.cc_dirty[gd >> kDirtyLog2] = 1;
        out_g = (uint32_t)gd;
        return 1;
    }
    const uint32_t vs = (uint32_t)(sv >> 32);
    const uint64_t cand = (uint64_t)vs + qweight(a, app, s, d);
    if (cand >= kInf) return 0;
    const uint64_t gd = (uint64_t)q * a.n + d;
    const uint64_t cur = a.st[gd];
    if (cand >= (cur >> 32)) return 0;
    const uint64_t old = atomicMin((unsigned long long *)&a.st[gd], (unsigned long long)((cand << 32) | s));
    if (cand >= (old >> 32)) return 0;
    if (bucket && a.far0 && (uint32_t)cand >= a.thr[q]) {
        const uint32_t ep = *(volatile uint32_t *)a.epoch;
        if (atomicExch(&a.fstamp[gd], ep) == ep) return 0;   // already in this epoch's pile
        out_g = (uint32_t)gd;
        return 2;
    }
    if (atomicExch(&a.tick[gd], stamp) == stamp) return 0;
    out_g = (uint32_t)gd;
    return 1;
}

// Pull: the row owner `d` is the destination, `x` an in-neighbour.
__device__ __forceinline__ uint64_t pull_cand(const TravArgs &a, uint32_t q, uint32_t d, uint32_t x) {
    const uint64_t qoff = (uint64_t)q * a.n;
    const uint32_t vx = (uint32_t)(a.st[qoff + x] >> 32);
    if (vx == kInf) return ~0ull;
    const uint64_t cand = (uint64_t)vx + qweight(a, a.qapp[q], x, d);
    if (cand >= kInf) return ~0ull;
    return (cand << 32) | x;
}

__device__ __forceinline__ void pull_apply(const TravArgs &a, uint32_t q, uint32_t d, uint64_t packed) {
    if (packed == ~0ull) return;
    const uint64_t gd = (uint64_t)q * a.n + d;
    if (packed >= a.st[gd]) return;
    atomicMin((unsigned long long *)&a.st[gd], (unsigned long long)packed);
}

// CC pull: (label, depth + 1) of in-neighbour x; the parent entry follows the
// winning word.
__device__ __forceinline__ void pull_cc(const TravArgs &a, uint32_t q, uint32_t d, uint32_t x) {
    const uint64_t qoff = (uint64_t)q * a.n;
    const uint64_t cand = a.st[qoff + x] + 1;
    const uint64_t gd = qoff + d;
    if (cand >= a.st[gd]) return;
    const uint64_t old = atomicMin((unsigned long long *)&a.st[gd], (unsigned long long)cand);
    if (cand < old) a.par[gd] = x;
}

__device__ __forceinline__ bool pr_add(const TravArgs &a, uint32_t d, float c, uint32_t &out_g) {
    // Frontier::Add (include/framework/pr_residual.cuh:21-30)
    const float old = atomicAdd(&a.res[d], c);
    const float val = old + c;
    if (!isfinite(val)) atomicExch((unsigned int *)&a.ctl->bad, 1u);
    if (fabsf(old) <= a.eps && fabsf(val) > a.eps && atomicCAS(&a.queued[d], 0u, 1u) == 0u) {
        out_g = d;
        return true;
    }
    return false;
}

template <uint32_t MODE>
// 0: nothing to publish, 1: next frontier, 2: far pile (push only).
__device__ __forceinline__ uint32_t edge_op(const TravArgs &a, uint32_t q, uint32_t owner, uint64_t val, uint32_t x,
                                            uint32_t stamp, uint32_t &out_g) {
    if (x & kTomb) return 0;   // deleted entry of a large row
    if (MODE == M_PUSH) return push_op(a, q, owner, val, x, stamp, out_g);
    if (MODE == M_INVALIDATE) return invalidate_op(a, q, owner, val, x, out_g) ? 1u : 0u;
    if (MODE == M_PR || MODE == M_PR_AMEND) return pr_add(a, x, __uint_as_float((uint32_t)val), out_g) ? 1u : 0u;
    if (a.qapp[q] == DCG_APP_CC) pull_cc(a, q, owner, x);
    else pull_apply(a, q, owner, pull_cand(a, q, owner, x));
    return 0;
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

// Workgroup output buffer of a persistent launch: waves append new frontier
// (and far-pile) entries with LDS atomics and the workgroup publishes them
// with one global atomic per phase (flush_out), instead of one global atomic
// per wave and 64-edge chunk on a single counter.  A group that finds the
// buffer full goes to global memory directly and leaves sentinels behind,
// which the flush compacts away.
constexpr uint32_t kOutCap = 4096, kFarCap = 2048;
struct OutLds {
    uint32_t n, nf, cn, cf, base, fbase, abase;
    uint32_t near[kOutCap];
    uint32_t far[kFarCap];
};

template <uint32_t MODE>
__device__ __forceinline__ void emit_b(OutLds *Op, const TravArgs &a, uint32_t *outq, uint32_t *count, bool want,
                                       uint32_t g) {
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
    emit<MODE>(a, outq, count, want, g);
}

__device__ __forceinline__ void emit_far_b(OutLds *Op, const TravArgs &a, bool want, uint32_t g) {
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
    emit_far(a, want, g);
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

template <uint32_t MODE>
__device__ void phase_a(const TravArgs &a, uint32_t r, const Team &t, WaveLds &L, OutLds *O, uint64_t &edges,
                        uint64_t &expansions) {
    const uint32_t cur = r % 3, nxt = (r + 1) % 3;
    const uint32_t lane = lane_id();
    const uint32_t cnt = *(volatile uint32_t *)&a.ctl->count[cur];
    const uint32_t *in = r == 0 ? a.in0 : ((r & 1) ? a.qa : a.qb);
    uint32_t *out = (r & 1) ? a.qb : a.qa;
    const uint32_t stamp = stamp_base(a) + r;
    uint32_t batch = t.wave, spare = 0;   // the first batch is static; then kClaim per atomic
    for (;;) {
        const uint32_t b = batch * kWave;
        if (b >= cnt) break;
        const uint32_t i = b + lane;
        const bool has = i < cnt;
        const uint32_t g = has ? in[i] : 0u;
        uint32_t q = 0, v = 0;
        uint32_t deg = 0;
        uint64_t val = 0;
        uint64_t off = 0;
        if (has) {
            if (MODE == M_PR || MODE == M_PR_AMEND) {
                v = g;
                const RowDesc d = a.s.out[v];
                const float x = MODE == M_PR ? a.delta[i] : a.rank[v] * a.sign;
                const uint32_t live = row_live(d);
                if (live && x != 0.f) {
                    deg = d.ext;
                    off = d.off;
                    val = __float_as_uint(0.85f * x / (float)live);
                    mark_read(a.s, a.s.hot_out, v);
                }
            } else {
                decode_g(a, g, q, v);
                const RowDesc d = MODE == M_PULL ? a.s.in[v] : a.s.out[v];
                val = MODE == M_PULL ? 0ull : a.st[g];
                if (d.ext && (MODE == M_PULL || (uint32_t)(val >> 32) != kInf)) {
                    deg = d.ext;
                    off = d.off;
                    mark_read(a.s, MODE == M_PULL ? a.s.hot_in : a.s.hot_out, v);
                }
            }
        }
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
        for (uint32_t k0 = 0; k0 < total; k0 += kWave) {
            const uint32_t k = k0 + lane;
            uint32_t kind = 0;
            uint32_t og = 0;
            if (k < total) {
                uint32_t lo = 0, hi = 64;   // last owner with pref <= k
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (L.pref[mid] <= k) lo = mid;
                    else hi = mid;
                }
                const uint32_t x = L.base[lo][k - L.pref[lo]];
                kind = edge_op<MODE>(a, L.q[lo], L.owner[lo], L.val[lo], x, stamp, og);
            }
            if (MODE != M_PULL) emit_b<MODE>(O, a, out, &a.ctl->count[nxt], kind == 1, og);
            if (MODE == M_PUSH && a.far0) emit_far_b(O, a, kind == 2, og);
        }
        __builtin_amdgcn_wave_barrier();
        if (spare) {
            ++batch;
            --spare;
        } else {
            uint32_t nb = 0;
            if (lane == 0) nb = t.nwaves + atomicAdd(&a.ctl->cursor[cur], kClaim);
            batch = __shfl(nb, 0, kWave);
            spare = kClaim - 1;
        }
    }
}

// Phase B of round r: the pieces of the large rows (first piece of every wave
// static, then an atomic cursor).
template <uint32_t MODE>
__device__ void phase_b(const TravArgs &a, uint32_t r, const Team &t, OutLds *O, uint64_t &edges) {
    const uint32_t cur = r % 3, nxt = (r + 1) % 3;
    const uint32_t lane = lane_id();
    uint32_t *out = (r & 1) ? a.qb : a.qa;
    const uint32_t stamp = stamp_base(a) + r;
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
            if (MODE == M_PULL && a.qapp[q] == DCG_APP_CC) {
                for (uint32_t k = lane; k < len; k += kWave) {
                    const uint32_t x = base[k];
                    if (!(x & kTomb)) pull_cc(a, q, v, x);
                }
            } else if (MODE == M_PULL) {
                uint64_t best = ~0ull;
                for (uint32_t k = lane; k < len; k += kWave) {
                    const uint32_t x = base[k];
                    const uint64_t c = (x & kTomb) ? ~0ull : pull_cand(a, q, v, x);
                    best = c < best ? c : best;
                }
                for (int d = 32; d >= 1; d >>= 1) {
                    const uint64_t o = __shfl_xor(best, d, kWave);
                    best = o < best ? o : best;
                }
                if (lane == 0) pull_apply(a, q, v, best);
            } else {
                for (uint32_t k0 = 0; k0 < len; k0 += kWave) {
                    const uint32_t k = k0 + lane;
                    uint32_t kind = 0;
                    uint32_t og = 0;
                    if (k < len) kind = edge_op<MODE>(a, q, v, be.val, base[k], stamp, og);
                    emit_b<MODE>(O, a, out, &a.ctl->count[nxt], kind == 1, og);
                    if (MODE == M_PUSH && a.far0) emit_far_b(O, a, kind == 2, og);
                }
            }
        }
        uint32_t np2 = 0;
        if (lane == 0) np2 = t.nwaves + atomicAdd(&a.ctl->big_cursor[cur], 1u);
        p = __shfl(np2, 0, kWave);
    }
}

// Start of round r: ring counters for rounds r+1 / r+2, and PR's consume step.
template <uint32_t MODE>
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
        // Consume (pr_residual.cuh:53-69): fold residuals above eps into ranks.
        const uint32_t cnt = *(volatile uint32_t *)&a.ctl->count[cur];
        const uint32_t *in = r == 0 ? a.in0 : ((r & 1) ? a.qa : a.qb);
        for (uint32_t i = t.tid; i < cnt; i += t.nthreads) {
            const uint32_t v = in[i];
            const float res = a.res[v];
            a.queued[v] = 0;
            float dl = 0.f;
            if (!isfinite(res) || !isfinite(a.rank[v])) atomicExch((unsigned int *)&a.ctl->bad, 1u);
            if (fabsf(res) > a.eps) {
                a.rank[v] += res;
                a.res[v] = 0.f;
                dl = res;
            }
            a.delta[i] = dl;
        }
    }
}

__device__ __forceinline__ bool pieces_pending(const TravArgs &a, uint32_t r) {
    return (*(volatile unsigned long long *)&a.ctl->big[r % 3] & kBigMask) != 0;
}

// Grid barrier of a persistent launch (co-resident workgroups: the launch is
// cooperative).  One arrival counter and a generation word per launch in its
// control block; agent-scope release before arriving and acquire after the
// generation moves publish every workgroup's writes to all XCDs.  A waiter that
// sees no progress for ~10 s (100 MHz wall clock) marks the launch timed out
// and every later barrier returns at once, so the launch always ends.
__device__ __forceinline__ void gsync(const TravArgs &a) {
    __syncthreads();
    if (threadIdx.x == 0 && !__hip_atomic_load(&a.ctl->bar_timeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        uint32_t *arrive = &a.ctl->bar_arrive, *gen = &a.ctl->bar_gen;
        const uint32_t g = __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        if (__hip_atomic_fetch_add(arrive, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1) {
            __hip_atomic_store(arrive, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(gen, g + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const uint64_t t0 = wall_clock64();
            while (__hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g) {
                __builtin_amdgcn_s_sleep(1);
                if (wall_clock64() - t0 > 1000000000ull) {
                    __hip_atomic_store(&a.ctl->bar_timeout, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
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
__device__ bool far_advance(const TravArgs &a, uint32_t r, const Team &t) {
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
        const uint32_t val = (uint32_t)(a.st[g] >> 32);
        if (lds) atomicMin(&qmin[q], val);
        else atomicMin(&mf[q], val);
    }
    if (lds) {
        __syncthreads();
        for (uint32_t q = threadIdx.x; q < a.Q; q += blockDim.x)
            if (qmin[q] != kInf) atomicMin(&mf[q], qmin[q]);
    }
    gsync(a);
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
    for (uint32_t b = t.wave * kWave; b < fc; b += t.nwaves * kWave) {
        const uint32_t i = b + lane;
        bool near = false, far = false;
        uint32_t g = 0;
        if (i < fc) {
            uint32_t q, v;
            g = src[i];
            decode_g(a, g, q, v);
            near = (uint32_t)(a.st[g] >> 32) < thr_of(mf[q], a.fdelta[q]);
            far = !near;
            if (far) a.fstamp[g] = ep;
        }
        emit<M_PUSH>(a, in, &a.ctl->count[r % 3], near, g);
        emit_to(a, dst, &a.ctl->far_count[fp ^ 1], far, g);
    }
    if (t.tid == 0) *(volatile uint32_t *)a.epoch = ep;
    gsync(a);
    return true;
}

constexpr uint32_t kSoloMax = 2048;      // frontier entries workgroup 0 handles alone
constexpr uint32_t kSoloPieces = 64;     // more pending pieces than this end a solo stretch

// One cooperative launch (1,024-thread workgroups) runs rounds until the
// frontier is empty, with the gsync grid barrier between rounds.  While the
// frontier is small, workgroup 0 runs rounds by itself with workgroup barriers
// (the rest of the grid waits at one grid barrier); otherwise all workgroups
// share each round, separated by grid barriers (a second one only when the
// round queued pieces of large rows).
// TAG only separates launch names: 1 for static convergence (dcg_start, the
// autotuner), 0 for incremental closures, so profiles list them apart.
template <uint32_t MODE, uint32_t TAG = 0>
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
            if (MODE == M_PUSH && a.far0 && idle_advances < 2 && far_advance(a, r, gt)) {
                ++idle_advances;
                continue;
            }
            if (idle_advances >= 2 && threadIdx.x == 0 && blockIdx.x == 0) a.ctl->overflow = 1;
            break;
        }
        idle_advances = 0;
        if (cnt <= kSoloMax && gridDim.x > 1) {
            if (blockIdx.x == 0) {
                // solo stretch: rounds on workgroup 0 until the frontier grows,
                // empties, or a round queues many large-row pieces
                uint32_t resume_b = 0;
                for (;;) {
                    round_open<MODE>(a, r, bt);
                    __syncthreads();
                    phase_a<MODE>(a, r, bt, L, nullptr, edges, expansions);
                    __syncthreads();
                    const uint64_t np = *(volatile unsigned long long *)&a.ctl->big[r % 3] & kBigMask;
                    if (np > kSoloPieces) {
                        resume_b = 1;   // the whole grid takes this round's pieces
                        break;
                    }
                    if (np) {
                        phase_b<MODE>(a, r, bt, nullptr, edges);
                        __syncthreads();
                    }
                    ++r;
                    const uint32_t c2 = *(volatile uint32_t *)&a.ctl->count[r % 3];
                    if (c2 == 0 || c2 > kSoloMax || r >= a.max_rounds) break;
                }
                if (threadIdx.x == 0) {
                    a.ctl->solo_round = r;
                    a.ctl->solo_resume_b = resume_b;
                }
            }
            gsync(a);
            r = *(volatile uint32_t *)&a.ctl->solo_round;
            if (*(volatile uint32_t *)&a.ctl->solo_resume_b) {
                phase_b<MODE>(a, r, gt, &O, edges);
                flush(r);
                gsync(a);
                ++r;
            }
            continue;
        }
        round_open<MODE>(a, r, gt);
        if (MODE == M_PR) gsync(a);
        phase_a<MODE>(a, r, gt, L, &O, edges, expansions);
        flush(r);
        gsync(a);
        if (pieces_pending(a, r)) {
            phase_b<MODE>(a, r, gt, &O, edges);
            flush(r);
            gsync(a);
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
}

// ------------------------------------------------------------- plain kernels
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
    c->bar_arrive = c->bar_gen = c->bar_timeout = 0;
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

__global__ void k_init_state(uint64_t *st, uint32_t *par, uint32_t *tick, uint32_t n, uint32_t Q,
                             const uint32_t *qapp, const uint32_t *sources) {
    const uint64_t total = (uint64_t)n * Q;
    for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < total;
         g += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t q = (uint32_t)(g / n), v = (uint32_t)(g - (uint64_t)q * n);
        uint64_t s = kInfState;
        if (qapp[q] == DCG_APP_CC) {
            s = (uint64_t)v << 32;   // own label at depth 0, no parent
            par[g] = kInf;
        } else if (v == sources[q]) {
            s = (uint64_t)kInf;   // value 0, parent INF
        }
        st[g] = s;
        tick[g] = 0;
    }
}

// Initial frontier: a traversal query starts at its source, a CC query at every
// vertex.  `slot0[q]` is the query's first position in q0 (host prefix sum).
__global__ void k_iota_sources(uint32_t *q0, uint32_t n, uint32_t Q, const uint32_t *qapp, const uint32_t *sources,
                               const uint64_t *slot0) {
    for (uint32_t q = 0; q < Q; ++q) {
        if (qapp[q] == DCG_APP_CC) {
            for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n;
                 v += (uint64_t)gridDim.x * blockDim.x)
                q0[slot0[q] + v] = (uint32_t)((uint64_t)q * n + v);
        } else if (blockIdx.x == 0 && threadIdx.x == 0) {
            q0[slot0[q]] = (uint32_t)((uint64_t)q * n + sources[q]);
        }
    }
}

// Delete seeds (driver.cuh:394-441): test every deleted edge against every query.
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
            if (e.u < a.n && e.v < a.n) want = invalidate_op(a, q, e.u, a.st[(uint64_t)q * a.n + e.u], e.v, og);
        }
        const uint32_t pos = wg_append(want, &a.ctl->aff_count, wcnt, wbase);
        if (want) {
            if (pos < a.affcap) a.aff[pos] = og;
            else a.ctl->overflow = 1;
        }
    }
}

// Reset affected values (driver.cuh:443-463): INF, or the own id for CC.
__global__ void k_reset_affected(uint64_t *st, const uint32_t *aff, const Ctl *c, uint32_t n, uint32_t Q,
                                 const uint32_t *qapp) {
    const uint32_t count = c->aff_count;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t g = aff[i];
        const uint32_t q = Q == 1 ? 0u : g / n;
        const uint32_t v = g - q * n;
        st[g] = qapp[q] == DCG_APP_CC ? ((uint64_t)v << 32) : kInfState;   // CC: own root (parent already INF)
    }
}

// CC parents are written after the winning atomicMin of the state word, so two
// improvers of one vertex can leave the earlier one's parent behind (the
// invalidation rule treats such a parent as not tight, which only widens it).
// After a closure every CC vertex it touched (tick >= stamp_lo, or on the
// affected list) has its parent checked: kept if tight (equal label, depth - 1),
// else replaced by the first tight in-neighbour; depth 0 is a root (none).
__device__ __forceinline__ void cc_fix_one(const TravArgs &a, uint32_t g) {
    uint32_t q, v;
    decode_g(a, g, q, v);
    if (a.qapp[q] != DCG_APP_CC) return;
    const uint64_t qoff = (uint64_t)q * a.n;
    const uint64_t sv = a.st[g];
    const uint32_t p = a.par[g];
    if ((uint32_t)sv == 0) {
        if (p != kInf) a.par[g] = kInf;
        return;
    }
    if (p != kInf && a.st[qoff + p] + 1 == sv) return;
    const RowDesc d = a.s.in[v];
    uint32_t best = kInf;
    for (uint32_t j = 0; j < d.ext; ++j) {
        const uint32_t x = *row_base(a.s, d.off + j);
        if (!(x & kTomb) && a.st[qoff + x] + 1 == sv) {
            best = x;
            break;
        }
    }
    a.par[g] = best;
}

// With a list: the listed slots (one thread each).  Without: query slots
// [base, base + n) stamped since stamp_lo, visiting only the 1,024-slot blocks
// that a CC improvement marked (one workgroup per block); the marks are
// cleared after every CC query of the stage was checked.
__global__ void k_cc_fix_parents(TravArgs a, const uint32_t *list, const Ctl *list_ctl, uint32_t base) {
    const uint32_t stamp_lo = stamp_base(a);
    if (list) {
        const uint64_t total = (uint64_t)list_ctl->aff_count;
        for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total;
             i += (uint64_t)gridDim.x * blockDim.x)
            cc_fix_one(a, list[i]);
        return;
    }
    const uint64_t lo = base, hi = (uint64_t)base + a.n;
    for (uint64_t b = (lo >> kDirtyLog2) + blockIdx.x; (b << kDirtyLog2) < hi; b += gridDim.x) {
        if (!a.cc_dirty[b]) continue;   // the same byte for the whole workgroup
        for (uint32_t j = threadIdx.x; j < (1u << kDirtyLog2); j += blockDim.x) {
            const uint64_t g = (b << kDirtyLog2) + j;
            if (g < lo || g >= hi || a.tick[g] < stamp_lo) continue;
            cc_fix_one(a, (uint32_t)g);
        }
    }
}

// Insertion seeds (framework.cuh:297-370): relax every added edge for every query.
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
                const uint64_t sv = a.st[(uint64_t)q * a.n + e.u];
                if ((uint32_t)(sv >> 32) != kInf) want = push_op(a, q, e.u, sv, e.v, stamp0, og, false) == 1;
            }
        }
        const uint32_t pos = wg_append(want, &a.ctl->aff_count, wcnt, wbase);
        if (want) {
            if (pos < a.qcap) outq[pos] = og;
            else a.ctl->overflow = 1;
        }
    }
}

// ------------------------------------------------------------------- host side
static void check_launch(const Ctl &c, const char *stage) {
    if (c.bar_timeout) throw Error(DCG_ERR_HIP, std::string("a grid barrier timed out in ") + stage);
    if (c.overflow) throw Error(DCG_ERR_CAPACITY, std::string("frontier overflow in ") + stage);
}

// `a`'s stamp base (stamp0, or its step slot) is the first stamp of the phase.
// With a list: the listed vertex-query slots.  Without: the slots of each CC
// query stamped in this stage, in the 1,024-slot blocks a CC improvement
// marked; the marks are cleared once every CC query was checked.
static void cc_fix_parents(Engine &e, const TravArgs &a, const uint32_t *list, const Ctl *lc) {
    if (!e.d_par.p) return;
    if (list) {
        const uint64_t work = (uint64_t)e.n * e.Q;
        hipLaunchKernelGGL(k_cc_fix_parents, dim3(grid_for(work, 256, 16384)), dim3(256), 0, e.stream, a, list, lc,
                           0u);
        return;
    }
    const uint64_t blocks = (e.n >> kDirtyLog2) + 2;   // a query's range spans at most this many blocks
    for (uint32_t q = 0; q < e.Q; ++q)
        if (e.qapps[q] == DCG_APP_CC)
            hipLaunchKernelGGL(k_cc_fix_parents, dim3(grid_for(blocks, 1, 16384)), dim3(256), 0, e.stream, a, list,
                               lc, q * e.n);
    fill<uint8_t>(e.stream, e.d_cc_dirty.p, e.d_cc_dirty.n, (uint8_t)0);
}

uint32_t Engine::coop_blocks(const void *kernel) {
    int per_cu = 0;
    DCG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kPBlock, 0));
    if (per_cu < 1) per_cu = 1;
    const int want = tune.blocks_per_cu ? (int)tune.blocks_per_cu : 1;
    if (per_cu > want) per_cu = want;
    return (uint32_t)(per_cu * cu_count);
}

static const void *persistent_fn(uint32_t mode, bool is_static) {
    if (is_static) return mode == M_PR ? (const void *)&k_persistent<M_PR, 1> : (const void *)&k_persistent<M_PUSH, 1>;
    switch (mode) {
    case M_PUSH: return (const void *)&k_persistent<M_PUSH>;
    case M_PULL: return (const void *)&k_persistent<M_PULL>;
    case M_INVALIDATE: return (const void *)&k_persistent<M_INVALIDATE>;
    case M_PR: return (const void *)&k_persistent<M_PR>;
    default: return (const void *)&k_persistent<M_PR_AMEND>;
    }
}

void launch_persistent(Engine &e, uint32_t mode, TravArgs &a, bool is_static = false) {
    const void *fn = persistent_fn(mode, is_static);
    const uint32_t gi = mode + (is_static ? 5u : 0u);   // grid sizes per instantiation
    if (!e.coop_grid[gi]) e.coop_grid[gi] = e.coop_blocks(fn);
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
    const int slot = e.time_begin(k, (int)(a.ctl - e.d_ctl.p));
    DCG_HIP(hipLaunchCooperativeKernel(fn, dim3(e.coop_grid[gi]), dim3(kPBlock), args, 0, e.stream));
    e.time_end(slot);
}

// Stamps: each launch owns a window of 2^20 round stamps; the tick array is
// cleared when the 32-bit stamp space would wrap.
// Before a phase: keep `windows` stamp windows ahead without a wrap, so every
// stamp the phase takes is above the first one (the CC parent check below
// selects the phase's vertices by stamp).
static void reserve_stamps(Engine &e, uint32_t windows) {
    constexpr uint64_t kWindow = 1ull << 20;
    if ((uint64_t)e.stamp + (uint64_t)(windows + 2) * kWindow > 0xFFFFFFFFull) {
        fill<uint32_t>(e.stream, e.d_tick.p, (uint64_t)e.n * e.Q, 0u);
        e.stamp = 1;
    }
}

static uint32_t take_stamps(Engine &e) {
    constexpr uint32_t kWindow = 1u << 20;
    if (e.stamp > 0xFFFFFFFFu - 2 * kWindow) {
        fill<uint32_t>(e.stream, e.d_tick.p, (uint64_t)e.n * e.Q, 0u);
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
    a.st = e.d_st.p;
    a.par = e.d_par.p;
    a.tick = e.d_tick.p;
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
    for (uint32_t q = 0; q < Q; ++q) {
        if (qapps[q] == DCG_APP_CC || !qt) h[q] = kInf;
        else if (qapps[q] == DCG_APP_BFS) h[q] = qt;
        else h[q] = std::max<uint32_t>(1u, (uint32_t)(((uint64_t)wmax * qt) / 4));
    }
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
    d_tick.alloc(cap);
    fill<uint32_t>(stream, d_tick.p, cap, 0u);
    if (app == DCG_APP_PR) {
        d_rank.alloc(cap);
        d_res.alloc(cap);
        d_delta.alloc(cap);
        d_queued.alloc(cap);
    } else {
        d_st.alloc(cap);
        d_aff.alloc(cap);
        bool cc = false, ordered = false;
        for (uint32_t q = 0; q < Q; ++q) (qapps[q] == DCG_APP_CC ? cc : ordered) = true;
        if (cc) {
            d_par.alloc(cap);
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
    double best = measure();
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
    }
}

void Engine::start_min_lattice() {
    const uint64_t qn = (uint64_t)n * Q;
    hipLaunchKernelGGL(k_init_state, dim3(grid_for(qn)), dim3(256), 0, stream, d_st.p, d_par.p, d_tick.p, n, Q,
                       d_qapps.p,
                       d_sources.p);
    stamp = 1;
    std::vector<uint64_t> slot0(Q);
    uint64_t count0 = 0;
    for (uint32_t q = 0; q < Q; ++q) {
        slot0[q] = count0;
        count0 += qapps[q] == DCG_APP_CC ? n : 1;
    }
    DevBuf<uint64_t> d_slot0;
    d_slot0.alloc(Q);
    DCG_HIP(hipMemcpyAsync(d_slot0.p, slot0.data(), Q * 8, hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(k_iota_sources, dim3(grid_for(n)), dim3(256), 0, stream, d_qc.p, n, Q, d_qapps.p, d_sources.p,
                       d_slot0.p);
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
void Engine::delete_phase(uint64_t begin, uint64_t count, dcg_phase_stats *st) {
    const dcg_edge *recs = batch_records(kDels, begin, count);
    mutate_delete(recs, count, st);
    mlp_release_training(*this);   // the learned policy trains beside the repair, not the mutation
    if (!count) return;
    reserve_stamps(*this, 2);
    h_step[0] = StepArgs{recs, count, {take_stamps(*this), take_stamps(*this), 0u, 0u}};
    DCG_HIP(hipMemcpyAsync(d_step.p, h_step, sizeof(StepArgs), hipMemcpyHostToDevice, stream));
    const int rs = time_begin(DCG_K_REPAIR_DELETE, -1);
    run_stage(0);
    time_end(rs);
    mark_batches_used();
    fold_enqueue();
    Ctl *c = d_ctl.p;
    DCG_HIP(hipGetLastError());
    DCG_HIP(hipMemcpyAsync(h_ctl, c, 4 * sizeof(Ctl), hipMemcpyDeviceToHost, stream));
    sync();   // one wait per stage: mutation counters, repair control blocks, read fold
    finish_delete(count, st);
    for (int i = 0; i < 4; ++i)
        check_launch(h_ctl[i], "the delete stage");
    timed[rs].bytes = stage_bytes(h_ctl, 1, 4);
    if (graph_exec[0]) collect_graph_timers(0);
    if (st) {
        st->affected = h_ctl[1].aff_count;
        st->rounds = h_ctl[1].rounds + h_ctl[3].rounds;
        st->bucket_advances = h_ctl[3].advances;
        st->far_entries = h_ctl[3].far_scanned;
        st->edges_scanned = h_ctl[1].edges + h_ctl[2].edges + h_ctl[3].edges;
    }
    fold_finish(st);
}

// Add stage (framework.cuh:3525-3756): append, seed from the added edges, close.
void Engine::add_phase(uint64_t begin, uint64_t count, dcg_phase_stats *st) {
    const dcg_edge *recs = batch_records(kAdds, begin, count);
    mutate_add(recs, count, st);
    if (!count) return;
    reserve_stamps(*this, 2);
    h_step[1] = StepArgs{recs, count, {take_stamps(*this), take_stamps(*this), 0u, 0u}};
    DCG_HIP(hipMemcpyAsync(d_step.p + 1, h_step + 1, sizeof(StepArgs), hipMemcpyHostToDevice, stream));
    const int rs = time_begin(DCG_K_REPAIR_ADD, -1);
    run_stage(1);
    time_end(rs);
    mark_batches_used();
    fold_enqueue();
    Ctl *c = d_ctl.p;
    DCG_HIP(hipGetLastError());
    DCG_HIP(hipMemcpyAsync(h_ctl + 4, c + 4, 2 * sizeof(Ctl), hipMemcpyDeviceToHost, stream));
    sync();   // one wait per stage: mutation counters, repair control blocks, read fold
    finish_add(count, st);
    check_launch(h_ctl[4], "the add stage");
    check_launch(h_ctl[5], "the add stage");
    timed[rs].bytes = stage_bytes(h_ctl, 5, 6);
    if (graph_exec[1]) collect_graph_timers(1);
    if (st) {
        st->rounds = h_ctl[5].rounds;
        st->bucket_advances = h_ctl[5].advances;
        st->far_entries = h_ctl[5].far_scanned;
        st->edges_scanned = h_ctl[5].edges;
    }
    fold_finish(st);
}

// The repair launches of a stage, reading the batch and the round stamps from
// d_step[stage].  Delete (framework.cuh:4811-4957, re-ordered: the physical
// deletion has run; the invalidation scans post-delete rows, which only drops
// deleted edges that the seed step already examined): seeds -> invalidation
// closure -> reset -> one pull pass -> push closure.  Add: seeds -> push closure.
void Engine::enqueue_repair(int stage) {
    Ctl *c = d_ctl.p;
    TravArgs a = base_args(*this);
    a.step = d_step.p + stage;
    constexpr uint32_t kSeedBlocks = 512;   // 1,024-thread workgroups, grid-stride over count * Q records
    if (stage == 0) {
        hipLaunchKernelGGL(k_ctl_setup, dim3(1), dim3(64), 0, stream, c + 0, (const uint32_t *)nullptr, 0u, 0u,
                           (const uint32_t *)nullptr);
        a.ctl = c + 0;
        hipLaunchKernelGGL(k_seed_dels, dim3(kSeedBlocks), dim3(1024), 0, stream, a);
        hipLaunchKernelGGL(k_ctl_setup, dim3(1), dim3(64), 0, stream, c + 1, &c[0].aff_count, 0u, 0u,
                           &c[0].aff_count);
        a.ctl = c + 1;
        a.in0 = d_aff.p;
        a.stamp_slot = 0;
        launch_persistent(*this, M_INVALIDATE, a);
        hipLaunchKernelGGL(k_reset_affected, dim3(grid_for((uint64_t)n * Q, 256, 8192)), dim3(256), 0, stream,
                           d_st.p, d_aff.p, c + 1, n, Q, d_qapps.p);
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
        hipLaunchKernelGGL(k_seed_adds, dim3(kSeedBlocks), dim3(1024), 0, stream, a, d_qc.p);
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

// Run a stage's repair launches: replay its captured graph (captured at the
// first call; a capture the runtime rejects turns graphs off for the engine and
// the launches are issued directly), or issue them directly.
void Engine::run_stage(int stage) {
    if (!tune.graphs) {
        enqueue_repair(stage);
        return;
    }
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
            enqueue_repair(stage);
            return;
        }
    }
    DCG_HIP(hipGraphLaunch(graph_exec[stage], stream));
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
                             uint32_t *tick, uint32_t stamp, uint32_t *out, Ctl *c) {
    const uint64_t total = na + nd;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t b = blockIdx.x * (uint64_t)blockDim.x; b < total; b += stride) {
        const uint64_t i = b + threadIdx.x;
        bool want = false;
        uint32_t u = 0;
        if (i < total) {
            const dcg_edge e = i < na ? adds[i] : dels[i - na];
            u = e.u;
            if (e.u < n && e.v < n) want = atomicExch(&tick[u], stamp) != stamp;
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
constexpr uint32_t kDenseSmall = 64;     // in-rows up to this many slots: a quarter-wave each
constexpr uint32_t kDensePiece = 2048;   // larger in-rows: one wave per piece of this many slots

// In-rows by size: small rows listed, larger rows cut into pieces; the total
// slot count prices a round (edges).
__global__ __launch_bounds__(1024) void k_prd_split(StoreView s, uint32_t n, uint32_t *small, uint2 *big, Ctl *c) {
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
        const uint32_t np = ext > kDenseSmall ? (ext + kDensePiece - 1) / kDensePiece : 0u;
        const uint32_t ps = wg_reserve(sm ? 1u : 0u, &c->count[0], wcnt, wbase);
        if (sm) small[ps] = v;
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
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        const float x = res[v];
        if (!isfinite(x) || !isfinite(rank[v])) atomicExch((unsigned long long *)&c->bad, 1ull);
        queued[v] = 0;
        rank[v] += x;
        res[v] = 0.f;
        const uint32_t live = row_live(s.out[v]);
        delta[v] = live ? 0.85f * x / (float)live : 0.f;
    }
}

// Small in-rows: 16 lanes read a row 16 slots at a time and add the shares.
__global__ __launch_bounds__(256) void k_prd_pull_small(StoreView s, const uint32_t *rows, const Ctl *c,
                                                        const float *delta, float *res) {
    if (*(volatile const uint32_t *)&c->done) return;
    const uint32_t nrows = c->count[0];
    const uint32_t lane = lane_id(), ql = lane & 15;
    const uint64_t quarters = (uint64_t)gridDim.x * blockDim.x / 16;
    for (uint64_t i = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / 16; i < nrows; i += quarters) {
        const uint32_t v = rows[i];
        const RowDesc d = s.in[v];
        const uint32_t *base = row_base(s, d.off);   // a block of at most 64 slots lies in one chunk
        float sum = 0.f;
        for (uint32_t j = ql; j < d.ext; j += 16) {
            const uint32_t x = base[j];
            if (!(x & kTomb)) sum += delta[x];
        }
        for (int o = 8; o >= 1; o >>= 1) sum += __shfl_xor(sum, o, 16);
        if (ql == 0) {
            res[v] = sum;
            mark_read(s, s.hot_in, v);
        }
    }
}

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
        const uint32_t end = d.ext - pc.y < kDensePiece ? d.ext : pc.y + kDensePiece;
        float sum = 0.f;
        for (uint32_t j = pc.y + lane; j < end; j += kWave) {
            const uint32_t x = *row_base(s, d.off + j);
            if (!(x & kTomb)) sum += delta[x];
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
    d_prd_big.ensure(slots_bound);
    hipLaunchKernelGGL(k_ctl_setup, dim3(1), dim3(64), 0, stream, c, (const uint32_t *)nullptr, 0u, 0u,
                       (const uint32_t *)nullptr);
    fill<unsigned long long>(stream, &c->far_scanned, 1, 0ull);
    fill<uint32_t>(stream, &c->done, 1, 0u);
    fill<uint32_t>(stream, c->rmax, 2, 0u);
    const StoreView v = view();
    hipLaunchKernelGGL(k_prd_split, dim3(grid_for(n, 1024, 4096)), dim3(1024), 0, stream, v, n, d_prd_small.p,
                       d_prd_big.p, c);
    const uint32_t max_rounds = cfg.pr_max_rounds;   // 1..100, checked at creation
    const uint32_t cg = grid_for(n, 256, 4096), sg = grid_for((uint64_t)n * 16, 256, 65536);
    const uint32_t rg = grid_for(n, 256, 1024);   // residual maximum: ~16 entries per thread at n = 2^22
    const float eps = (float)cfg.pr_epsilon;
    // rounds in chunks of 16 between host checks; launches past convergence return at once
    const int slot = time_begin(DCG_K_PR_SOLVE, -1);
    for (uint32_t r0 = 0;; r0 += 16) {
        for (uint32_t r = r0; r < r0 + 16; ++r) {
            hipLaunchKernelGGL(k_prd_consume, dim3(cg), dim3(256), 0, stream, v, n, r, max_rounds, eps, d_rank.p,
                               d_res.p, d_delta.p, d_queued.p, c);
            hipLaunchKernelGGL(k_prd_pull_small, dim3(sg), dim3(256), 0, stream, v, d_prd_small.p, c, d_delta.p,
                               d_res.p);
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
                       dc, n, d_tick.p, stmp, srcs, c + 0);
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
