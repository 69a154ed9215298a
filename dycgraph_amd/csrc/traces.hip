// traces.hip -- diagnostic traces in the reference's file formats.
//
// Repair snapshot (include/framework/i16_repair_snapshot.h:9-52; captured at
// framework.cuh:4189-4212 before the repair and 4668-4677 after it): the
// affected set of the first delete stage after arming, the affected vertices'
// in-rows and the states of every vertex the repair reads or writes, so the
// repair can be replayed offline.  Cache-candidate trace (framework.cuh:
// 1414-1507, reader cache_patch_trace.h:15-60): the admitted prefix of each
// hotness ranking, for replaying the cache policy offline.
#include <algorithm>
#include <cstdio>

#include "engine.hpp"

namespace dcg {

constexpr uint64_t kI16Magic = 0x4931365354415445ull;
constexpr uint64_t kF1Magic = 0x4631434143484531ull;

struct I16State {
    uint32_t vertex, value, buffer, parent;
};

// Live (non-tombstone) entries of the rows of ids[0..count).
__global__ void k_rows_live(const RowDesc *desc, const uint32_t *ids, uint32_t count, uint64_t *live) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x)
        live[i] = row_live(desc[ids[i]]);
}

// The live entries of each listed row, in row order, from offs[i] on (one wave per row).
__global__ void k_gather_rows(StoreView s, const RowDesc *desc, const uint32_t *ids, uint32_t count,
                              const uint64_t *offs, uint32_t *out) {
    const uint32_t lane = lane_id();
    const uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / kWave;
    const uint64_t nwaves = (uint64_t)gridDim.x * blockDim.x / kWave;
    for (uint64_t i = wave; i < count; i += nwaves) {
        const RowDesc d = desc[ids[i]];
        uint64_t w = offs[i];
        for (uint32_t t0 = 0; t0 < d.ext; t0 += kWave) {
            const uint32_t k = t0 + lane;
            const uint32_t x = k < d.ext ? *row_base(s, d.off + k) : kTomb;
            const uint64_t km = __ballot(!(x & kTomb));
            if (!(x & kTomb)) out[w + lanes_below(km)] = x;
            w += (uint64_t)__popcll(km);
        }
    }
}

// States of query q at ids[0..count): value and parent from the state word
// (or, for rooted programs, the slot's parent); buffer = value.
__global__ void k_gather_states(const VSlot *vs, bool rooted, uint64_t base, const uint32_t *ids, uint32_t count,
                                I16State *out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const uint32_t v = ids[i];
        const VSlot x = vs[base + v];
        const uint32_t value = (uint32_t)(x.st >> 32);
        out[i] = I16State{v, value, value, rooted ? x.par : (uint32_t)x.st};
    }
}

template <class T>
static void put(std::ostream &o, const T &v) {
    o.write(reinterpret_cast<const char *>(&v), sizeof(v));
}
template <class T>
static void put_vector(std::ostream &o, const std::vector<T> &v) {
    put(o, (uint64_t)v.size());
    o.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
}

// The live entries of the rows `ids` of `desc` (out- or in-rows), in row order,
// as a host CSR: offsets (ids.size() + 1) and entries.
void Engine::gather_rows(const RowDesc *desc, const std::vector<uint32_t> &ids, std::vector<uint64_t> &offsets,
                         std::vector<uint32_t> &out) {
    const uint32_t a = (uint32_t)ids.size();
    offsets.assign(a + 1, 0);
    out.clear();
    if (!a) return;
    DevBuf<uint32_t> d_ids;
    DevBuf<uint64_t> d_offs;
    d_ids.alloc(a);
    d_offs.alloc(a);
    DCG_HIP(hipMemcpyAsync(d_ids.p, ids.data(), (size_t)a * 4, hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(k_rows_live, dim3(grid_for(a)), dim3(256), 0, stream, desc, d_ids.p, a, d_offs.p);
    std::vector<uint64_t> live(a);
    DCG_HIP(hipMemcpyAsync(live.data(), d_offs.p, (size_t)a * 8, hipMemcpyDeviceToHost, stream));
    sync();
    for (uint32_t i = 0; i < a; ++i) offsets[i + 1] = offsets[i] + live[i];
    out.resize(offsets[a]);
    if (!offsets[a]) return;
    DCG_HIP(hipMemcpyAsync(d_offs.p, offsets.data(), (size_t)a * 8, hipMemcpyHostToDevice, stream));
    DevBuf<uint32_t> d_out_rows;
    d_out_rows.alloc(offsets[a]);
    hipLaunchKernelGGL(k_gather_rows, dim3(grid_for((uint64_t)a * kWave, 256, 4096)), dim3(256), 0, stream, view(),
                       desc, d_ids.p, a, d_offs.p, d_out_rows.p);
    DCG_HIP(hipMemcpyAsync(out.data(), d_out_rows.p, offsets[a] * 4, hipMemcpyDeviceToHost, stream));
    sync();
}

// Runs the delete stage's repair in two parts around the two captures.
void Engine::i16_capture() {
    const uint32_t q = i16_query;
    const uint64_t base = (uint64_t)q * n;
    enqueue_repair(0, 1);   // seeds, invalidation, reset of the affected set
    DCG_HIP(hipMemcpyAsync(h_ctl + 1, d_ctl.p + 1, sizeof(Ctl), hipMemcpyDeviceToHost, stream));
    sync();
    check_launch(h_ctl[1], "the snapshot's invalidation");
    const uint64_t na = std::min<uint64_t>(h_ctl[1].aff_count, d_aff.n);
    std::vector<uint32_t> aff_g(na);
    if (na) DCG_HIP(hipMemcpyAsync(aff_g.data(), d_aff.p, na * 4, hipMemcpyDeviceToHost, stream));
    sync();
    std::vector<uint32_t> affected;
    for (uint32_t g : aff_g)
        if (g >= base && g - base < n) affected.push_back((uint32_t)(g - base));
    std::sort(affected.begin(), affected.end());
    affected.erase(std::unique(affected.begin(), affected.end()), affected.end());
    // in-rows of the affected vertices (post-delete rows, tombstones skipped)
    std::vector<uint64_t> offsets;
    std::vector<uint32_t> incoming;
    gather_rows(d_in.p, affected, offsets, incoming);
    std::vector<uint32_t> ids = incoming;
    ids.insert(ids.end(), affected.begin(), affected.end());
    ids.push_back(sources[q]);
    std::sort(ids.begin(), ids.end());
    ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
    const uint32_t m = (uint32_t)ids.size();
    DevBuf<uint32_t> d_sid;
    DevBuf<I16State> d_state;
    d_sid.alloc(m);
    d_state.alloc(m);
    DCG_HIP(hipMemcpyAsync(d_sid.p, ids.data(), (size_t)m * 4, hipMemcpyHostToDevice, stream));
    const bool root = rooted(q);
    auto states = [&]() {
        std::vector<I16State> out(m);
        hipLaunchKernelGGL(k_gather_states, dim3(grid_for(m)), dim3(256), 0, stream, d_vs.p, root, base, d_sid.p, m,
                           d_state.p);
        DCG_HIP(hipMemcpyAsync(out.data(), d_state.p, (size_t)m * sizeof(I16State), hipMemcpyDeviceToHost, stream));
        sync();
        return out;
    };
    const std::vector<I16State> before = states();
    enqueue_repair(0, 2);   // pull pass, push closure, CC parent check
    const std::vector<I16State> after = states();
    const std::string tmp = i16_path + ".tmp";
    {
        std::ofstream o(tmp, std::ios::binary | std::ios::trunc);
        put(o, kI16Magic);
        put(o, (uint32_t)1);             // version
        put(o, (uint32_t)0xFFFFFFFFu);   // infinity
        // weight rule: 1 = (u + v) % 128 + 1 (the reference's); 2 = the seeded
        // weights 1 + mix64(seed ^ (u << 32 | v)) % w_max of this engine
        put(o, (uint32_t)(wp.kind == W_SEEDED ? 2 : 1));
        put(o, (uint32_t)0);             // batch
        put(o, sources[q]);
        put(o, n);
        put_vector(o, affected);
        put_vector(o, offsets);
        put_vector(o, incoming);
        put_vector(o, before);
        put_vector(o, after);
        if (!o) throw Error(DCG_ERR_IO, "repair snapshot write failed: " + tmp);
    }
    if (std::rename(tmp.c_str(), i16_path.c_str()) != 0)
        throw Error(DCG_ERR_IO, "repair snapshot rename failed: " + i16_path);
    i16_armed = false;
}

// Open (truncate) the cache-candidate trace and write its magic; an empty
// path closes it.
void Engine::f1_open(const char *path) {
    if (f1.is_open()) f1.close();
    f1_on = f1_init = false;
    f1_prev_ids.clear();
    f1_prev_deg.clear();
    f1_prev_edges = 0;
    if (!path || !*path) return;
    f1.open(path, std::ios::out | std::ios::binary | std::ios::trunc);
    put(f1, kF1Magic);
    f1.flush();
    if (!f1) throw Error(DCG_ERR_IO, std::string("cannot open cache-candidate trace ") + path);
    f1_on = true;
}

// One record of the cache-candidate trace from the last ranking.
void Engine::f1_record(uint32_t batch, dcg_trace_stats *out) {
    if (!f1_on) throw Error(DCG_ERR_STATE, "no cache-candidate trace open");
    const bool unchanged = f1_init && rank_ids == f1_prev_ids && rank_deg == f1_prev_deg;
    const uint64_t count = unchanged ? 0 : rank_ids.size();
    put(f1, batch);
    put(f1, rank_cap);
    put(f1, count);
    uint64_t edges = unchanged ? f1_prev_edges : 0;
    if (!unchanged) {
        std::vector<uint32_t> rec(2 * count);
        for (uint64_t i = 0; i < count; ++i) {
            rec[2 * i] = rank_ids[i];
            rec[2 * i + 1] = rank_deg[i];
            edges += rank_deg[i];
        }
        f1.write(reinterpret_cast<const char *>(rec.data()), (std::streamsize)(rec.size() * 4));
    }
    f1.flush();
    if (!f1) throw Error(DCG_ERR_IO, "cache-candidate trace write failed");
    if (out) {
        out->admitted_vertices = rank_ids.size();
        out->admitted_edges = edges;
        out->changed_records = count;
        out->capacity_edges = rank_cap;
        out->bytes = 4 + 8 + 8 + count * 8;
    }
    f1_prev_ids = rank_ids;
    f1_prev_deg = rank_deg;
    f1_prev_edges = edges;
    f1_init = true;
}

}  // namespace dcg
