// gen.cpp -- seeded synthetic inputs: RMAT edge sets and update streams.
//
// RMAT: edge i draws `scale` quadrant choices from a counter-based hash
// (16 bits per level, thresholds a, a+b, a+b+c in 1/65536 units), so draw i
// does not depend on any other draw; the edge set is the first `num_edges`
// distinct non-self-loop pairs in draw order, returned sorted by (src, dst).
// The same draw function runs on the GPU (gen.hip) for large scales.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <unordered_map>
#include <vector>

#include "gen_common.hpp"

namespace {

// Open-addressing set of 64-bit keys (key ~0 is the empty marker).
struct FlatSet {
    std::vector<uint64_t> t;
    uint64_t mask = 0, size = 0;
    explicit FlatSet(uint64_t expect) {
        uint64_t cap = 16;
        while (cap < expect * 2) cap <<= 1;
        t.assign(cap, ~0ull);
        mask = cap - 1;
    }
    bool insert(uint64_t k) {
        uint64_t i = dcg::mix64(k) & mask;
        for (;;) {
            if (t[i] == ~0ull) {
                t[i] = k;
                size++;
                return true;
            }
            if (t[i] == k) return false;
            i = (i + 1) & mask;
        }
    }
    bool contains(uint64_t k) const {
        uint64_t i = dcg::mix64(k) & mask;
        for (;;) {
            if (t[i] == ~0ull) return false;
            if (t[i] == k) return true;
            i = (i + 1) & mask;
        }
    }
};

}  // namespace

extern "C" int dcg_gen_rmat(uint32_t scale, uint64_t num_edges, double a, double b, double c, uint64_t seed,
                            uint32_t *src, uint32_t *dst) {
    if (scale == 0 || scale > 31 || !src || !dst || a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return DCG_ERR_ARG;
    const double n = std::ldexp(1.0, (int)scale);
    if ((double)num_edges > n * (n - 1) * 0.5) return DCG_ERR_ARG;
    const dcg::RmatParams p = dcg::rmat_params(scale, a, b, c, seed);
    try {
        FlatSet set(num_edges);
        std::vector<uint64_t> keys;
        keys.reserve(num_edges);
        for (uint64_t i = 0; keys.size() < num_edges; ++i) {
            const uint64_t k = dcg::rmat_draw(p, i);
            if ((uint32_t)(k >> 32) == (uint32_t)k) continue;
            if (set.insert(k)) keys.push_back(k);
        }
        std::sort(keys.begin(), keys.end());
        for (uint64_t i = 0; i < num_edges; ++i) {
            src[i] = (uint32_t)(keys[i] >> 32);
            dst[i] = (uint32_t)keys[i];
        }
    } catch (const std::bad_alloc &) {
        return DCG_ERR_CAPACITY;
    }
    return DCG_OK;
}

// Update stream.  Per snapshot s (counter-based draws keyed by seed and s):
//   deletions: uniform over the edges present at the snapshot start, without
//     replacement (index into original edges ++ edges inserted so far; draws
//     that hit a removed or already chosen edge are redrawn);
//   insertions: uniform (u, v), u != v, absent at the snapshot start and
//     distinct within the snapshot.
extern "C" int dcg_gen_stream(uint32_t num_nodes, const uint32_t *src, const uint32_t *dst, uint64_t num_edges,
                              uint32_t snapshots, const uint64_t *adds_per, const uint64_t *dels_per, uint64_t seed,
                              dcg_edge *adds, dcg_edge *dels) {
    if (!num_nodes || (num_edges && (!src || !dst)) || !adds_per || !dels_per) return DCG_ERR_ARG;
    try {
        // index of (u,v) among the sorted original edges, or num_edges
        auto find_original = [&](uint64_t key) -> uint64_t {
            uint64_t lo = 0, hi = num_edges;
            while (lo < hi) {
                const uint64_t mid = (lo + hi) >> 1;
                const uint64_t mk = ((uint64_t)src[mid] << 32) | dst[mid];
                if (mk < key) lo = mid + 1;
                else hi = mid;
            }
            if (lo < num_edges && (((uint64_t)src[lo] << 32) | dst[lo]) == key) return lo;
            return num_edges;
        };
        uint64_t total_adds = 0, total_dels = 0;
        for (uint32_t s = 0; s < snapshots; ++s) {
            total_adds += adds_per[s];
            total_dels += dels_per[s];
        }
        FlatSet removed_orig(total_dels + 16);          // original edge indices removed
        std::vector<uint64_t> inserted;                  // keys of inserted edges, in order
        std::vector<uint8_t> ins_alive;
        inserted.reserve(total_adds);
        ins_alive.reserve(total_adds);
        std::unordered_map<uint64_t, uint64_t> live_ins; // key -> index of its live inserted copy
        live_ins.reserve(total_adds + 16);
        uint64_t alive_orig = num_edges;
        uint64_t ao = 0, dof = 0;
        for (uint32_t s = 0; s < snapshots; ++s) {
            const uint64_t nd = dels_per[s], na = adds_per[s];
            if (nd > alive_orig + live_ins.size()) return DCG_ERR_ARG;
            FlatSet chosen(nd + 16);
            uint64_t ctr = 0;
            const uint64_t range = num_edges + inserted.size();
            for (uint64_t k = 0; k < nd;) {
                const uint64_t r = dcg::mix64(seed ^ dcg::mix64(((uint64_t)s << 40) ^ (0x5A5Aull << 56) ^ ctr++));
                const uint64_t j = r % range;
                if (j < num_edges) {
                    if (removed_orig.contains(j) || !chosen.insert(j)) continue;
                    dels[dof + k] = dcg_edge{src[j], dst[j]};
                } else {
                    const uint64_t t = j - num_edges;
                    if (!ins_alive[t] || !chosen.insert(j)) continue;
                    dels[dof + k] = dcg_edge{(uint32_t)(inserted[t] >> 32), (uint32_t)inserted[t]};
                }
                k++;
            }
            FlatSet fresh(na + 16);
            ctr = 0;
            for (uint64_t k = 0; k < na;) {
                const uint64_t r = dcg::mix64(seed ^ dcg::mix64(((uint64_t)s << 40) ^ (0xA5A5ull << 56) ^ ctr++));
                const uint32_t u = (uint32_t)((r >> 32) % num_nodes), v = (uint32_t)((uint32_t)r % num_nodes);
                if (u == v) continue;
                const uint64_t key = ((uint64_t)u << 32) | v;
                bool exists = live_ins.count(key) != 0;
                if (!exists) {
                    const uint64_t j = find_original(key);
                    exists = j < num_edges && !removed_orig.contains(j);
                }
                if (exists || !fresh.insert(key)) continue;
                adds[ao + k] = dcg_edge{u, v};
                k++;
            }
            // apply the snapshot (deletions first, then insertions)
            for (uint64_t k = 0; k < nd; ++k) {
                const uint64_t key = ((uint64_t)dels[dof + k].u << 32) | dels[dof + k].v;
                auto it = live_ins.find(key);
                if (it != live_ins.end()) {
                    ins_alive[it->second] = 0;
                    live_ins.erase(it);
                } else {
                    removed_orig.insert(find_original(key));
                    alive_orig--;
                }
            }
            for (uint64_t k = 0; k < na; ++k) {
                const uint64_t key = ((uint64_t)adds[ao + k].u << 32) | adds[ao + k].v;
                live_ins[key] = inserted.size();
                inserted.push_back(key);
                ins_alive.push_back(1);
            }
            ao += na;
            dof += nd;
        }
    } catch (const std::bad_alloc &) {
        return DCG_ERR_CAPACITY;
    }
    return DCG_OK;
}

// Road-like grid: side x side vertices (id = r*side + c), 4-neighbour arcs in
// both directions, plus for each cell with hash < diag_frac a diagonal to
// (r+1, c+1) in both directions.  Returned sorted by (src, dst).  Pass
// src == NULL to get the edge count.
extern "C" int dcg_gen_grid(uint32_t side, double diag_frac, uint64_t seed, uint32_t *src, uint32_t *dst,
                            uint64_t *num_edges) {
    if (side < 2 || (uint64_t)side * side > 0xFFFFFFFFull || !num_edges || diag_frac < 0 || diag_frac > 1)
        return DCG_ERR_ARG;
    const uint64_t thresh = (uint64_t)(diag_frac * 18446744073709551615.0);
    auto diag = [&](uint32_t r, uint32_t c) {   // cell (r, c) has a diagonal to (r+1, c+1)
        return r + 1 < side && c + 1 < side && dcg::mix64(seed ^ ((uint64_t)r * side + c)) < thresh;
    };
    uint64_t k = 0;
    for (uint32_t r = 0; r < side; ++r) {
        for (uint32_t c = 0; c < side; ++c) {
            const uint32_t u = r * side + c;
            uint32_t nb[6];
            int m = 0;
            if (r > 0 && c > 0 && diag(r - 1, c - 1)) nb[m++] = u - side - 1;
            if (r > 0) nb[m++] = u - side;
            if (c > 0) nb[m++] = u - 1;
            if (c + 1 < side) nb[m++] = u + 1;
            if (r + 1 < side) nb[m++] = u + side;
            if (diag(r, c)) nb[m++] = u + side + 1;
            for (int j = 0; j < m; ++j) {
                if (src) {
                    src[k] = u;
                    dst[k] = nb[j];
                }
                k++;
            }
        }
    }
    *num_edges = k;
    return DCG_OK;
}

// Batch sizes lognormal(mu, sigma) clipped to [lo, hi] (Box-Muller on
// splitmix64 draws).
extern "C" int dcg_gen_lognormal_sizes(uint32_t count, double mu, double sigma, uint64_t lo, uint64_t hi,
                                       uint64_t seed, uint64_t *out) {
    if (!out || lo > hi) return DCG_ERR_ARG;
    for (uint32_t i = 0; i < count; ++i) {
        const double u1 = ((dcg::mix64(seed ^ (2ull * i)) >> 11) + 1) * (1.0 / 9007199254740993.0);
        const double u2 = (dcg::mix64(seed ^ (2ull * i + 1)) >> 11) * (1.0 / 9007199254740992.0);
        const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
        double x = std::exp(mu + sigma * z);
        if (x < (double)lo) x = (double)lo;
        if (x > (double)hi) x = (double)hi;
        out[i] = (uint64_t)std::llround(x);
    }
    return DCG_OK;
}
