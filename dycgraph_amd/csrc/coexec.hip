// coexec.hip -- CPU/GPU co-execution of the insertion closure.
//
// The reference's owner runtime (framework.cuh:923-1155, 5198-5517;
// include/framework/dual_domain_event_runtime.h:81-268) gives host threads a
// set of CPU-owned vertices: during an insertion epoch their state lives on
// the host, the GPU closure turns every relaxation into a CPU-owned vertex
// into a (dst, value, parent) proposal, the CPU relaxes its own vertices' rows
// and proposes back for GPU-owned destinations, and the epoch ends when no
// proposal is outstanding on either side.  Here the exchange is bulk
// synchronous: the GPU closure runs to quiescence, its proposals (already the
// minimum per destination, bprop) go to the host, the host threads run their
// closure to quiescence, their proposals (deduplicated to the minimum per
// destination) seed the next GPU closure, and so on until a side has nothing
// to send.  Every proposal is a min-lattice offer, so the fixed point is the
// same as the all-GPU closure's.  Co-execution covers the non-rooted programs
// (SSSP, BFS) of the built-in set; rooted queries (CC) run on the GPU alone.
#include <algorithm>
#include <thread>

#include "engine.hpp"
#include "../../include/dcg_program.hpp"

namespace dcg {

// out[q * C + i] = st[q * n + idx[i]] for one query's slice (i < C).
__global__ void k_cx_gather(const VSlot *vs, uint64_t base, const uint32_t *idx, uint32_t C, uint64_t *out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < C; i += gridDim.x * blockDim.x)
        out[i] = vs[base + idx[i]].st;
}

__global__ void k_cx_scatter(VSlot *vs, uint64_t base, const uint32_t *idx, uint32_t C, const uint64_t *in) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < C; i += gridDim.x * blockDim.x)
        vs[base + idx[i]].st = in[i];
}

// The GPU's proposals to the CPU: (slot, packed) of every listed slot, and the
// slot's minimum reset for the next round.
__global__ void k_cx_take(unsigned long long *bprop, const uint32_t *blist, uint32_t count, uint64_t *packed) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const uint32_t g = blist[i];
        packed[i] = bprop[g];
        bprop[g] = ~0ull;
    }
}

// The CPU's proposals to the GPU: a strict improvement of the value lowers the
// state word and enqueues the slot once per stamp (the seeds of the next
// closure), appended with one atomic per wave.
__global__ void k_cx_apply(VSlot *vs, uint32_t stamp, const uint64_t *packed, const uint32_t *gs, uint32_t count,
                           uint32_t *outq, uint64_t qcap, Ctl *c) {
    for (uint32_t b = blockIdx.x * blockDim.x; b < count; b += gridDim.x * blockDim.x) {
        const uint32_t i = b + threadIdx.x;
        bool want = false;
        uint32_t g = 0;
        if (i < count) {
            g = gs[i];
            const uint64_t p = packed[i];
            if ((uint32_t)(p >> 32) < (uint32_t)(vs[g].st >> 32)) {
                const uint64_t old = atomicMin((unsigned long long *)&vs[g].st, (unsigned long long)p);
                want = (uint32_t)(p >> 32) < (uint32_t)(old >> 32) && atomicExch(&vs[g].tick, stamp) != stamp;
            }
        }
        const uint64_t m = __ballot(want);
        if (m) {
            const int leader = __ffsll((unsigned long long)m) - 1;
            uint32_t base = 0;
            if ((int)lane_id() == leader) base = atomicAdd(&c->aff_count, (uint32_t)__popcll(m));
            base = __shfl(base, leader, kWave);
            if (want) {
                const uint64_t pos = (uint64_t)base + lanes_below(m);
                if (pos < qcap) outq[pos] = g;
                else c->overflow = 1;
            }
        }
    }
}

void Engine::coexec_set(const uint8_t *cpu_owned) {
    // the proposal minima are all ~0 between stages (every listed slot is
    // reset when taken), so a same-sized array is kept across domain changes
    DevBuf<unsigned long long> bprop;
    std::swap(bprop.p, cx.d_bprop.p);
    std::swap(bprop.n, cx.d_bprop.n);
    cx.reset();
    if (!cpu_owned) return;
    cx.index.assign(n, -1);
    for (uint32_t v = 0; v < n; ++v)
        if (cpu_owned[v]) {
            cx.index[v] = (int32_t)cx.vertices.size();
            cx.vertices.push_back(v);
        }
    if (cx.vertices.empty()) return;
    const uint32_t C = (uint32_t)cx.vertices.size();
    const uint64_t qn = (uint64_t)n * Q;
    cx.d_owner.alloc(n);
    DCG_HIP(hipMemcpyAsync(cx.d_owner.p, cpu_owned, n, hipMemcpyHostToDevice, stream));
    if (bprop.n == qn) {
        std::swap(bprop.p, cx.d_bprop.p);
        std::swap(bprop.n, cx.d_bprop.n);
    } else {
        cx.d_bprop.alloc(qn);
        fill<unsigned long long>(stream, cx.d_bprop.p, qn, ~0ull);
    }
    cx.d_blist.alloc((uint64_t)Q * C);
    cx.d_bcount.alloc(1);
    fill<uint32_t>(stream, cx.d_bcount.p, 1, 0u);
    cx.d_idx.alloc(C);
    DCG_HIP(hipMemcpyAsync(cx.d_idx.p, cx.vertices.data(), (size_t)C * 4, hipMemcpyHostToDevice, stream));
    cx.d_vals.alloc((uint64_t)Q * C);
    cx.stamp.assign((size_t)Q * C, 0u);
    const unsigned hc = std::thread::hardware_concurrency();
    cx.threads = (int)std::max(1u, std::min(hc ? hc : 1u, 64u));
    sync();
    cx.on = true;
}

namespace {

// The host program of query q (SSSP or BFS of the built-in set): the
// candidate for w from u at value val, or kInf.
struct HostPrograms {
    std::vector<uint32_t> kind;   // 0 SSSP, 1 BFS, 2 other (not co-executed)
    std::vector<api::Env> env;
    uint32_t candidate(uint32_t q, uint32_t u, uint32_t w, uint32_t val) const {
        if (kind[q] == 0) {
            const apps::SSSP app(env[q]);
            return app.AccumulateBuffer(u, w, app.TraversalEdgeWeight(u, w), val);
        }
        const apps::BFS app(env[q]);
        return app.AccumulateBuffer(u, w, app.TraversalEdgeWeight(u, w), val);
    }
};

inline bool lower_value(uint64_t *slot, uint64_t cand) {
    uint64_t cur = __atomic_load_n(slot, __ATOMIC_RELAXED);
    while ((uint32_t)(cand >> 32) < (uint32_t)(cur >> 32)) {
        if (__atomic_compare_exchange_n(slot, &cur, cand, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) return true;
    }
    return false;
}

}  // namespace

// Co-executed add stage: the stage's seeds and GPU closure (relaxations into
// CPU-owned vertices held back as proposals), then alternate host and GPU
// closures on each other's proposals until one side has none.
void Engine::coexec_add_stage() {
    const uint32_t C = (uint32_t)cx.vertices.size();
    HostPrograms hp;
    hp.kind.resize(Q);
    hp.env.resize(Q);
    std::vector<uint32_t> coq;   // co-executed queries
    for (uint32_t q = 0; q < Q; ++q) {
        hp.kind[q] = qapps[q] < 2 && prog == builtin_program() ? qapps[q] : 2u;
        hp.env[q] = api::Env{sources[q], wp.kind == W_SEEDED ? (uint32_t)DCG_W_SEEDED : (uint32_t)DCG_W_REFERENCE,
                             wp.wmax, wp.seed};
        if (hp.kind[q] < 2) coq.push_back(q);
    }
    // epoch start: the CPU-owned vertices' state and out-rows to the host
    cx.st.assign((size_t)Q * C, ~0ull);
    for (uint32_t q : coq)
        hipLaunchKernelGGL(k_cx_gather, dim3(grid_for(C)), dim3(256), 0, stream, d_vs.p, (uint64_t)q * n, cx.d_idx.p,
                           C, cx.d_vals.p + (uint64_t)q * C);
    DCG_HIP(hipGetLastError());
    DCG_HIP(hipMemcpyAsync(cx.st.data(), cx.d_vals.p, (size_t)Q * C * 8, hipMemcpyDeviceToHost, stream));
    gather_rows(d_out.p, cx.vertices, cx.rp, cx.col);   // synchronises
    cx.relax = cx.to_cpu = cx.to_gpu = cx.rounds = 0;

    cx.active = true;
    enqueue_coexec_seeds();   // seeds + first GPU closure (+ CC parent check)
    std::vector<uint64_t> hv;
    std::vector<uint32_t> hg;
    for (;;) {
        // the GPU closure's proposals to the CPU
        uint32_t nb = 0;
        DCG_HIP(hipMemcpyAsync(&nb, cx.d_bcount.p, 4, hipMemcpyDeviceToHost, stream));
        DCG_HIP(hipMemcpyAsync(h_ctl + 4, d_ctl.p + 4, 2 * sizeof(Ctl), hipMemcpyDeviceToHost, stream));
        sync();
        check_launch(h_ctl[4], "the co-executed add stage");
        check_launch(h_ctl[5], "the co-executed add stage");
        ++cx.rounds;
        if (nb == 0) break;
        nb = std::min<uint32_t>(nb, (uint32_t)cx.d_blist.n);
        hg.resize(nb);
        hv.resize(nb);
        DevBuf<uint64_t> d_take;
        d_take.alloc(nb);
        hipLaunchKernelGGL(k_cx_take, dim3(grid_for(nb)), dim3(256), 0, stream, cx.d_bprop.p, cx.d_blist.p, nb,
                           d_take.p);
        DCG_HIP(hipMemcpyAsync(hg.data(), cx.d_blist.p, (size_t)nb * 4, hipMemcpyDeviceToHost, stream));
        DCG_HIP(hipMemcpyAsync(hv.data(), d_take.p, (size_t)nb * 8, hipMemcpyDeviceToHost, stream));
        fill<uint32_t>(stream, cx.d_bcount.p, 1, 0u);
        sync();
        cx.to_cpu += nb;

        // host closure over the CPU-owned vertices
        std::vector<uint64_t> frontier;   // q << 32 | i
        const uint32_t s0 = ++cx.round_stamp;
        for (uint32_t k = 0; k < nb; ++k) {
            const uint32_t q = hg[k] / n, v = hg[k] - q * n;
            const int32_t i = cx.index[v];
            if (i < 0) continue;
            const size_t slot = (size_t)q * C + (uint32_t)i;
            if (lower_value(&cx.st[slot], hv[k]) && cx.stamp[slot] != s0) {
                cx.stamp[slot] = s0;
                frontier.push_back(((uint64_t)q << 32) | (uint32_t)i);
            }
        }
        std::vector<std::vector<uint64_t>> next(cx.threads), outv(cx.threads);
        std::vector<std::vector<uint32_t>> outg(cx.threads);
        std::vector<uint64_t> relax(cx.threads, 0);
        while (!frontier.empty()) {
            const uint32_t sr = ++cx.round_stamp;
            const int T = frontier.size() < 4096 ? 1 : cx.threads;
            auto body = [&](int t) {
                const size_t lo = frontier.size() * t / T, hi = frontier.size() * (t + 1) / T;
                for (size_t f = lo; f < hi; ++f) {
                    const uint32_t q = (uint32_t)(frontier[f] >> 32), i = (uint32_t)frontier[f];
                    const uint32_t u = cx.vertices[i];
                    const uint32_t val = (uint32_t)(__atomic_load_n(&cx.st[(size_t)q * C + i], __ATOMIC_RELAXED) >> 32);
                    for (uint64_t e = cx.rp[i]; e < cx.rp[i + 1]; ++e) {
                        const uint32_t w = cx.col[e];
                        const uint32_t c = hp.candidate(q, u, w, val);
                        ++relax[t];
                        if (c == kInf) continue;
                        const uint64_t cand = ((uint64_t)c << 32) | u;
                        const int32_t j = cx.index[w];
                        if (j >= 0) {
                            const size_t slot = (size_t)q * C + (uint32_t)j;
                            if (lower_value(&cx.st[slot], cand) &&
                                __atomic_exchange_n(&cx.stamp[slot], sr, __ATOMIC_RELAXED) != sr)
                                next[t].push_back(((uint64_t)q << 32) | (uint32_t)j);
                        } else {
                            outg[t].push_back((uint32_t)((uint64_t)q * n + w));
                            outv[t].push_back(cand);
                        }
                    }
                }
            };
            std::vector<std::thread> pool;
            for (int t = 1; t < T; ++t) pool.emplace_back(body, t);
            body(0);
            for (auto &th : pool) th.join();
            frontier.clear();
            for (auto &v : next) {
                frontier.insert(frontier.end(), v.begin(), v.end());
                v.clear();
            }
        }
        for (uint64_t r : relax) cx.relax += r;
        // the host's proposals to the GPU, the minimum per slot
        std::vector<std::pair<uint32_t, uint64_t>> props;
        for (int t = 0; t < cx.threads; ++t)
            for (size_t k = 0; k < outg[t].size(); ++k) props.push_back({outg[t][k], outv[t][k]});
        if (props.empty()) break;
        std::sort(props.begin(), props.end());
        hg.clear();
        hv.clear();
        for (size_t k = 0; k < props.size(); ++k)
            if (k == 0 || props[k].first != props[k - 1].first) {
                hg.push_back(props[k].first);
                hv.push_back(props[k].second);
            }
        cx.to_gpu += hg.size();
        enqueue_coexec_round(hg, hv);   // apply them and run the GPU closure on the improved slots
    }
    cx.active = false;
    // epoch end: the CPU-owned state back into the state array
    DCG_HIP(hipMemcpyAsync(cx.d_vals.p, cx.st.data(), (size_t)Q * C * 8, hipMemcpyHostToDevice, stream));
    for (uint32_t q : coq)
        hipLaunchKernelGGL(k_cx_scatter, dim3(grid_for(C)), dim3(256), 0, stream, d_vs.p, (uint64_t)q * n, cx.d_idx.p,
                           C, cx.d_vals.p + (uint64_t)q * C);
    DCG_HIP(hipGetLastError());
}

// CPU proposals -> the GPU: apply (k_cx_apply appends the improved slots to
// qc, its count in ctl[4]) with fresh stamps, then a push closure from them.
void Engine::coexec_apply(const std::vector<uint32_t> &gs, const std::vector<uint64_t> &packed, uint32_t stamp,
                          Ctl *c4) {
    const uint32_t k = (uint32_t)gs.size();
    cx.d_props.ensure((uint64_t)k + (k + 1) / 2);
    uint64_t *d_packed = cx.d_props.p;
    uint32_t *d_g = reinterpret_cast<uint32_t *>(cx.d_props.p + k);
    DCG_HIP(hipMemcpyAsync(d_packed, packed.data(), (size_t)k * 8, hipMemcpyHostToDevice, stream));
    DCG_HIP(hipMemcpyAsync(d_g, gs.data(), (size_t)k * 4, hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(k_cx_apply, dim3(grid_for(k)), dim3(256), 0, stream, d_vs.p, stamp, d_packed, d_g, k, d_qc.p,
                       (uint64_t)d_qc.n, c4);
    DCG_HIP(hipGetLastError());
}

}  // namespace dcg
