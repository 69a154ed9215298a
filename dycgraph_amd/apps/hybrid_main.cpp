// hybrid_main.cpp -- drop-in command-line front ends hybrid_sssp / hybrid_bfs /
// hybrid_cc / hybrid_pr over the C ABI (include/dcg.h).
//
// Accepts the reference's command lines (include/utils/app_skeleton.h:42-102 and
// the per-app flags of samples/hybrid_*/hybrid_*.cu), reads the same files
// (market_big graph, `a|d u v w` update file, `<adds> <dels>` size file), runs
// the same snapshot loop (all deletions, then all additions, then cache
// maintenance) and prints the same check/timer log lines and output files.
// Built once per app with -DDCG_CLI_APP=<0..3>.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "../../include/dcg.h"

#ifndef DCG_CLI_APP
#define DCG_CLI_APP 0
#endif

namespace {

constexpr uint32_t kInf = 0xFFFFFFFFu;
const char *kTag[] = {"SSSP", "BFS", "CC", "PR"};
const char *kName[] = {"sssp", "bfs", "cc", "pr"};

// ------------------------------------------------------------------ flags
struct Flags {
    std::map<std::string, std::string> v;
    std::set<std::string> known;
    void def(const std::string &k, const std::string &d) {
        known.insert(k);
        v[k] = d;
    }
    std::string s(const std::string &k) const { return v.at(k); }
    double d(const std::string &k) const { return std::atof(v.at(k).c_str()); }
    long long i(const std::string &k) const { return std::atoll(v.at(k).c_str()); }
    bool b(const std::string &k) const {
        const std::string &x = v.at(k);
        return x == "1" || x == "true" || x == "yes" || x == "TRUE" || x == "True";
    }
};

bool is_bool_flag(const Flags &f, const std::string &k) {
    const std::string &x = f.v.at(k);
    return x == "true" || x == "false";
}

// gflags syntax: -f=v, --f=v, -f v, --f v, -f (bool), -nof (bool).
int parse_flags(Flags &f, int argc, char **argv) {
    for (int a = 1; a < argc; ++a) {
        std::string t = argv[a];
        if (t.size() < 2 || t[0] != '-') {
            std::fprintf(stderr, "ERROR: unexpected argument '%s'\n", t.c_str());
            return -1;
        }
        t = t.substr(t[1] == '-' ? 2 : 1);
        std::string k = t, val;
        bool has = false;
        const size_t eq = t.find('=');
        if (eq != std::string::npos) {
            k = t.substr(0, eq);
            val = t.substr(eq + 1);
            has = true;
        }
        if (!f.known.count(k)) {
            if (k.rfind("no", 0) == 0 && f.known.count(k.substr(2)) && is_bool_flag(f, k.substr(2)) && !has) {
                f.v[k.substr(2)] = "false";
                continue;
            }
            std::fprintf(stderr, "ERROR: unknown command line flag '%s'\n", k.c_str());
            return -1;
        }
        if (!has) {
            if (is_bool_flag(f, k)) {
                val = "true";
            } else if (a + 1 < argc) {
                val = argv[++a];
            } else {
                std::fprintf(stderr, "ERROR: flag '%s' is missing its argument\n", k.c_str());
                return -1;
            }
        }
        f.v[k] = val;
    }
    return 0;
}

void LOG(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vprintf(fmt, ap);
    va_end(ap);
    std::fflush(stdout);
}

// ------------------------------------------------------------------ inputs
// market_big graph (src/utils/parser.cpp:296-406) or the binary edge list,
// read by the library on all host cores.
bool read_graph(const std::string &path, std::vector<uint32_t> &src, std::vector<uint32_t> &dst, uint32_t &n) {
    uint32_t *s = nullptr, *d = nullptr;
    uint64_t m = 0;
    if (dcg_read_graph(path.c_str(), &n, &s, &d, &m, 0) != DCG_OK) return false;
    src.assign(s, s + m);
    dst.assign(d, d + m);
    dcg_free(s);
    dcg_free(d);
    return true;
}

// Loader (include/framework/Loader.h:30-101).
bool read_updates(const std::string &sizes_path, const std::string &upd_path, std::vector<std::pair<uint64_t, uint64_t>> &batches,
                  std::vector<dcg_edge> &adds, std::vector<dcg_edge> &dels) {
    uint64_t *sz = nullptr, nb = 0, na = 0, nd = 0;
    dcg_edge *a = nullptr, *d = nullptr;
    if (dcg_read_updates(sizes_path.c_str(), upd_path.c_str(), &sz, &nb, &a, &na, &d, &nd, 0) != DCG_OK)
        return false;
    for (uint64_t b = 0; b < nb; ++b) batches.push_back({sz[2 * b], sz[2 * b + 1]});
    adds.assign(a, a + na);
    dels.assign(d, d + nd);
    dcg_free(sz);
    dcg_free(a);
    dcg_free(d);
    return true;
}

// ------------------------------------------------------------------ checkers
struct Csr {
    std::vector<uint64_t> rp;
    std::vector<uint32_t> col;
};

bool get_csr(dcg_engine *e, uint32_t n, Csr &g) {
    uint64_t m = 0;
    if (dcg_get_csr(e, nullptr, nullptr, &m)) return false;
    g.rp.resize(n + 1);
    g.col.resize(std::max<uint64_t>(m, 1));
    return dcg_get_csr(e, g.rp.data(), g.col.data(), &m) == 0;
}

uint32_t weight(int app, uint32_t u, uint32_t v) {
    if (app == DCG_APP_BFS) return 1;
    return (uint32_t)(((uint64_t)u + v) % 128 + 1);   // api.cuh:26-28
}

// CheckBellmanRelaxed + CheckParentWitness (samples/hybrid_sssp/hybrid_sssp.cu:228-333).
struct BellmanResult {
    uint64_t relaxable = 0, missing_witness = 0, invalid_parent = 0, missing_parent_edges = 0, wrong_parent = 0;
    uint32_t first_src = kInf, first_dst = kInf, first_missing = kInf, first_bad = kInf, first_bad_parent = kInf;
};

BellmanResult check_bellman(int app, const Csr &g, const std::vector<uint32_t> &dist, const std::vector<uint32_t> &par,
                            uint32_t source) {
    BellmanResult r;
    const uint32_t n = (uint32_t)dist.size();
    std::vector<uint8_t> tight(n, 0);
    if (source < n) tight[source] = 1;
    for (uint32_t u = 0; u < n; ++u) {
        if (dist[u] == kInf) continue;
        for (uint64_t e = g.rp[u]; e < g.rp[u + 1]; ++e) {
            const uint32_t v = g.col[e];
            const uint64_t c = (uint64_t)dist[u] + weight(app, u, v);
            if (dist[v] != kInf && c == dist[v]) tight[v] = 1;
            if (c < dist[v]) {
                if (!r.relaxable) {
                    r.first_src = u;
                    r.first_dst = v;
                }
                r.relaxable++;
            }
        }
    }
    for (uint32_t v = 0; v < n; ++v)
        if (dist[v] != kInf && !tight[v]) {
            if (!r.missing_witness) r.first_missing = v;
            r.missing_witness++;
        }
    for (uint32_t v = 0; v < n; ++v) {
        if (v == source || dist[v] == kInf) continue;
        const uint32_t p = par[v];
        bool valid = p < n && dist[p] != kInf;
        if (valid) {
            bool exists = false;
            for (uint64_t e = g.rp[p]; e < g.rp[p + 1]; ++e)
                if (g.col[e] == v) {
                    exists = true;
                    break;
                }
            const bool match = (uint64_t)dist[p] + weight(app, p, v) == dist[v];
            if (!exists) r.missing_parent_edges++;
            else if (!match) r.wrong_parent++;
            valid = exists && match;
        }
        if (!valid) {
            if (!r.invalid_parent) {
                r.first_bad = v;
                r.first_bad_parent = p;
            }
            r.invalid_parent++;
        }
    }
    return r;
}

uint64_t value_checksum(const std::vector<uint32_t> &x) {   // hybrid_sssp.cu:335-347
    uint64_t c = 1469598103934665603ull;
    for (uint32_t v = 0; v < x.size(); ++v) {
        if (x[v] == kInf) continue;
        c ^= (uint64_t)v + 0x9e3779b97f4a7c15ull + ((uint64_t)x[v] << 6) + ((uint64_t)x[v] >> 2);
        c *= 1099511628211ull;
    }
    return c;
}

// Increasing-seed flood fill (samples/hybrid_cc/hybrid_cc_host.cu:8-32).
uint64_t cc_check(const Csr &g, const std::vector<uint32_t> &labels) {
    const uint32_t n = (uint32_t)labels.size();
    std::vector<uint32_t> want(n, kInf), stack;
    for (uint32_t s = 0; s < n; ++s) {
        if (want[s] != kInf) continue;
        want[s] = s;
        stack.push_back(s);
        while (!stack.empty()) {
            const uint32_t u = stack.back();
            stack.pop_back();
            for (uint64_t e = g.rp[u]; e < g.rp[u + 1]; ++e)
                if (want[g.col[e]] == kInf) {
                    want[g.col[e]] = s;
                    stack.push_back(g.col[e]);
                }
        }
    }
    uint64_t err = 0;
    for (uint32_t v = 0; v < n; ++v) err += labels[v] != want[v];
    return err;
}

// PageRankCheck (samples/hybrid_pr/hybrid_pr_host.cu:9-56) in the caller's scale.
bool pr_check(const Csr &g, const std::vector<float> &rank, const std::vector<float> &res, double base, double eps,
              double out_scale, const char *stage, int batch) {
    const size_t n = rank.size();
    auto mult = [&](const std::vector<double> &x, std::vector<double> &y) {
        std::fill(y.begin(), y.end(), base);
        for (uint32_t u = 0; u < n; ++u) {
            const uint64_t d = g.rp[u + 1] - g.rp[u];
            if (!d) continue;
            const double c = 0.85 * x[u] / (double)d;
            for (uint64_t e = g.rp[u]; e < g.rp[u + 1]; ++e) y[g.col[e]] += c;
        }
    };
    std::vector<double> ref(n, 0), nx(n), act(rank.begin(), rank.end());
    bool conv = false;
    for (int r = 0; r < 10000; ++r) {
        mult(ref, nx);
        double delta = 0;
        for (size_t v = 0; v < n; ++v) delta += std::fabs(nx[v] - ref[v]);
        ref.swap(nx);
        if (delta <= 1e-12 * (base / 0.15) * (double)std::max<size_t>(1, n)) {
            conv = true;
            break;
        }
    }
    mult(act, nx);
    double l1 = 0, defect = 0, maxr = 0, maxe = 0, mass = 0;
    bool finite = true;
    for (size_t v = 0; v < n; ++v) {
        finite &= std::isfinite(rank[v]) && std::isfinite(res[v]) && rank[v] >= -eps;
        const double e = std::fabs(rank[v] - ref[v]);
        l1 += e;
        maxe = std::max(maxe, e);
        defect += std::fabs(nx[v] - rank[v] - res[v]);
        maxr = std::max(maxr, (double)std::fabs(res[v]));
        mass += std::fabs(ref[v]);
    }
    const double roundoff = 64 * std::numeric_limits<float>::epsilon() * std::max(base / 0.15, mass);
    const double bound = ((double)n * eps + roundoff) / 0.15;
    const bool ok = finite && conv && l1 <= bound && defect <= roundoff && maxr <= (double)(float)eps;
    std::printf("[PR-CHECK] stage=%s batch=%d %s l1_error=%.9g bound=%.9g max_error=%.9g invariant_l1=%.9g "
                "max_residual=%.9g\n",
                stage, batch, ok ? "passed" : "failed", l1 * out_scale, bound * out_scale, maxe * out_scale,
                defect * out_scale, maxr * out_scale);
    std::fflush(stdout);
    return ok;
}

double ms_since(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

}  // namespace

// Algorithm hints from the environment, validated like the reference
// (CG_ORDERED_REPAIR / CG_INSERTION_SCHEDULE: framework.cuh:3775-3783;
// CG_BATCH_MAINTENANCE: effective_update_batch.h:40-45).  Results never depend
// on them; they pick the executor: ordered = the bucketed push closure,
// block / thread = plain frontier order (one executor here: the edge-balanced
// wave expansion), batch maintenance large = the device-wide grouping sort for
// every batch (auto: for batches of at least 1,000,000 updates).
struct Hints {
    int ordered = -1;   // -1: the engine's automatic choice, 0 off, 1 on
    std::string batch = "regular";
};

bool read_hints(Hints &h, const char *tag) {
    const char *o = std::getenv("CG_ORDERED_REPAIR");
    const char *s = std::getenv("CG_INSERTION_SCHEDULE");
    const char *b = std::getenv("CG_BATCH_MAINTENANCE");
    const bool large_diameter = o && std::string(o) == "1";
    const std::string schedule = large_diameter ? "ordered" : (s ? s : "block");
    if (schedule != "block" && schedule != "thread" && schedule != "ordered") {
        std::fprintf(stderr, "[%s-ERROR] CG_INSERTION_SCHEDULE must be block, thread or ordered\n", tag);
        return false;
    }
    h.batch = b ? b : "regular";
    if (h.batch != "regular" && h.batch != "large" && h.batch != "auto") {
        std::fprintf(stderr, "[%s-ERROR] CG_BATCH_MAINTENANCE must be regular, large or auto\n", tag);
        return false;
    }
    if (schedule == "ordered") h.ordered = 1;
    else if (o || s) h.ordered = 0;   // an explicit unordered choice
    return true;
}

// Cache capacity in edges of the reference's cache-candidate ranking for its
// integer -cache flag (graph_datum.cuh:164-176).
uint64_t reference_cache_edges(int cache) {
    switch (cache) {
    case 3: return 805306365ull;
    case 2: return 536870921ull;
    case 1: return 268435455ull;
    case 0: return 0;
    default: return 1073741824ull;
    }
}

// State count of a written repair snapshot (the `before` vector's length).
uint64_t snapshot_states(const std::string &path) {
    std::ifstream in(path, std::ios::binary);
    uint64_t count = 0;
    in.seekg(32);   // magic, version, infinity, weight rule, batch, source, nodes
    for (int v = 0; v < 4 && in; ++v) {
        in.read(reinterpret_cast<char *>(&count), 8);
        if (v < 3) in.seekg((std::streamoff)(count * (v == 1 ? 8 : 4)), std::ios::cur);
    }
    return in ? count : 0;
}

int main(int argc, char **argv) {
    const int app = DCG_CLI_APP;
    const std::string nm = kName[app];
    Flags f;
    // include/utils/app_skeleton.h:42-102
    for (const char *k : {"interactive", "priority", "check", "verbose", "trace", "undirected", "stats", "gen_graph",
                          "gen_weights", "wl_sort", "wl_unique", "weight", "topology_replay_audit", "sparse",
                          "large_batch", "print_ranks"})
        f.def(k, "false");
    f.v["check"] = "true";
    f.v["verbose"] = "true";
    f.v["weight"] = "true";
    for (const char *k : {"cmdfile", "output", "lb_push", "lb_pull", "graphfile", "json", "updatefile", "update_size"})
        f.def(k, "");
    f.def("format", "gr");
    f.def("max_iteration", "50");
    f.def("prio_delta", "0");
    f.def("edge_factor", "1.0");
    f.def("alpha", "0.8");
    f.def("beta", "0.4");
    f.def("gen_nnodes", "100000");
    f.def("gen_factor", "10");
    f.def("gen_method", "0");
    f.def("gen_weight_range", "100");
    f.def("block_size", "512");
    f.def("grid_size", "20");
    f.def("wl_alloc_factor", "0.2");
    f.def("n_stream", "3");
    f.def("SEGMENT", "1");
    f.def("priority_a", "1");
    f.def("hybrid", "0");
    f.def("residence", "0");
    f.def("weight_num", "0");
    f.def("cache", "4");
    f.def("top_ranks", "10");
    // per-app flags
    f.def("source_node", "0");
    f.def(nm + "_max_batches", "10");
    f.def(nm + "_print_checksum", "false");
    f.def("sssp_cpu_partition_capacity", "0");
    f.def("sssp_cpu_domain_map", "");
    f.def("sssp_hotness_audit", "false");
    for (const char *k : {"i16_repair_snapshot", "e0b_trace_file", "f1_cache_trace_file", "f1_component_trace_file"})
        f.def(k, "");
    f.def("error", "0.000001");
    f.def("pr_max_rounds", "100");
    // engine-specific
    f.def("placement", "host");          // host: pinned chunk store + device cache (-cache GB); device: all in HBM
    f.def("cache_policy", "hotness");    // none | hotness | learned
    f.def("pr_scale", "reference");      // reference: base 0.15 per vertex; probability: base 0.15/n
    f.def("weight_mode", "reference");   // reference: (u+v)%128+1; seeded: 1+mix64 % weight_max
    f.def("weight_max", "100");
    f.def("weight_seed", "1");
    if (parse_flags(f, argc, argv)) return 1;
    Hints hints;
    if (!read_hints(hints, kTag[app])) return 1;
    if (f.s("graphfile").empty()) {
        std::fprintf(stderr, "[%s-ERROR] graphfile is required\n", kTag[app]);
        return 1;
    }
    if (app == DCG_APP_PR) {
        const double err = f.d("error");
        const long long rounds = f.i("pr_max_rounds"), mb = f.i("pr_max_batches");
        if (!std::isfinite(err) || err <= 0 || (float)err == 0.f || mb < 0 || rounds <= 0 || rounds > 100) {
            std::fprintf(stderr, "[PR-ERROR] PR requires finite error > 0, pr_max_batches >= 0 and 1 <= "
                                 "pr_max_rounds <= 100\n");
            return 1;
        }
        if (f.s("updatefile").empty() != f.s("update_size").empty()) {
            std::fprintf(stderr, "[PR-ERROR] PR requires both updatefile and update_size, or neither\n");
            return 1;
        }
    }
    // CPU/GPU co-execution (framework.cuh:923-1000, 3344-3395): a domain map (one
    // u16 per vertex, 0 = CPU owner) or the CPU's share of SEGMENT partitions
    const std::string dmap = f.s("sssp_cpu_domain_map");
    const long long cpu_cap = f.i("sssp_cpu_partition_capacity");
    const long long segs = std::max(1LL, f.i("SEGMENT"));
    std::vector<uint64_t> seg_start;   // capacity mode: partition boundaries with equal out-edge counts
    if (!dmap.empty() && cpu_cap != 0) {
        LOG("[E2-DOMAIN] protocol_error=domain_map_conflicts_with_capacity capacity=%lld\n", cpu_cap);
        return 1;
    }
    if (dmap.empty() && (cpu_cap < 0 || cpu_cap > segs)) {
        LOG("[DUAL-RUNTIME] protocol_error=invalid_cpu_capacity capacity=%lld partitions=%lld\n", cpu_cap, segs);
        return 1;
    }
    if ((!dmap.empty() || cpu_cap) && (app == DCG_APP_PR || app == DCG_APP_CC)) {
        std::fprintf(stderr, "[%s-ERROR] CPU/GPU co-execution covers SSSP and BFS\n", kTag[app]);
        return 1;
    }
    std::vector<uint32_t> src, dst;
    uint32_t n = 0;
    if (!read_graph(f.s("graphfile"), src, dst, n) || n == 0) {
        std::fprintf(stderr, "[%s-ERROR] cannot read graph %s\n", kTag[app], f.s("graphfile").c_str());
        return 1;
    }
    std::vector<uint8_t> cpu_mask;   // domain map: 1 = CPU-owned
    if (!dmap.empty()) {
        std::ifstream in(dmap, std::ios::binary | std::ios::ate);
        const long long expected = (long long)n * 2, actual = in.is_open() ? (long long)in.tellg() : -1;
        if (actual != expected) {
            LOG("[E2-DOMAIN] invalid_map path=%s expected_bytes=%lld actual_bytes=%lld\n", dmap.c_str(), expected,
                actual);
            return 1;
        }
        in.seekg(0);
        std::vector<uint16_t> dom(n);
        in.read(reinterpret_cast<char *>(dom.data()), expected);
        cpu_mask.resize(n);
        for (uint32_t v = 0; v < n; ++v) {
            if (dom[v] > 1) {
                LOG("[E2-DOMAIN] invalid_owner vertex=%u owner=%u\n", v, (unsigned)dom[v]);
                return 1;
            }
            cpu_mask[v] = dom[v] == 0;
        }
    }

    const bool check = f.b("check");
    const uint32_t source = (uint32_t)std::min<long long>(std::max(0LL, f.i("source_node")), (long long)n - 1);
    dcg_config cfg;
    dcg_config_init(&cfg);
    cfg.app = (uint32_t)app;
    cfg.sources = &source;
    cfg.num_queries = 1;
    cfg.placement = f.s("placement") == "device" ? DCG_PLACE_DEVICE : DCG_PLACE_HOST;
    cfg.cache_bytes = (uint64_t)(f.d("cache") * 1e9);
    const std::string pol = f.s("cache_policy");
    cfg.cache_policy = pol == "none" ? DCG_CACHE_NONE : pol == "learned" ? DCG_CACHE_LEARNED : DCG_CACHE_HOTNESS;
    cfg.weight_mode = f.s("weight_mode") == "seeded" ? DCG_W_SEEDED : DCG_W_REFERENCE;
    cfg.weight_max = (uint32_t)f.i("weight_max");
    cfg.weight_seed = (uint64_t)f.i("weight_seed");
    const bool prob = f.s("pr_scale") == "probability";
    const double scale = prob ? 1.0 / n : 1.0;
    cfg.pr_base = 0.15 * scale;
    cfg.pr_epsilon = f.d("error") * scale;
    cfg.pr_max_rounds = (uint32_t)std::max(1LL, f.i("pr_max_rounds"));
    dcg_engine *eng = nullptr;
    int st = dcg_create(&cfg, &eng);
    if (st) {
        std::fprintf(stderr, "[%s-ERROR] dcg_create: %s\n", kTag[app], dcg_status_string(st));
        return 1;
    }
    auto fail = [&](const char *what, int s) {
        std::fprintf(stderr, "[%s-ERROR] %s: %s (%s)\n", kTag[app], what, dcg_status_string(s), dcg_last_error(eng));
        dcg_destroy(eng);
        return 1;
    };
    if (app != DCG_APP_PR && hints.ordered >= 0 && (st = dcg_set_buckets(eng, hints.ordered ? 4u : 0u)))
        return fail("CG_ORDERED_REPAIR", st);   // 4 quarters: SSSP window w_max (128, ordered_gpu_repair.cuh:116-160)
    if (hints.batch == "large" && (st = dcg_set_batch_maintenance(eng, 1))) return fail("CG_BATCH_MAINTENANCE", st);
    if ((st = dcg_load_edges(eng, n, src.data(), dst.data(), src.size()))) return fail("load", st);
    if (!cpu_mask.empty()) {
        if ((st = dcg_set_cpu_domain(eng, cpu_mask.data(), n))) return fail("sssp_cpu_domain_map", st);
        uint64_t cpu = 0;
        for (uint8_t m : cpu_mask) cpu += m;
        LOG("[E2-DOMAIN] loaded path=%s cpu_vertices=%llu cpu_vertex_share=%.6f\n", dmap.c_str(),
            (unsigned long long)cpu, n ? (double)cpu / n : 0.0);
    } else if (cpu_cap) {
        std::vector<uint64_t> deg(n + 1, 0);
        for (uint32_t u : src) ++deg[u + 1];
        for (uint32_t v = 0; v < n; ++v) deg[v + 1] += deg[v];
        seg_start.assign(segs + 1, n);
        seg_start[0] = 0;
        for (long long s = 1; s < segs; ++s)
            seg_start[s] = (uint64_t)(std::lower_bound(deg.begin(), deg.end(), deg[n] * s / segs) - deg.begin());
    }
    auto t_init = std::chrono::steady_clock::now();
    if ((st = dcg_start(eng))) return fail("start", st);
    if (app == DCG_APP_PR)
        LOG("[PR-INITIAL] compute_ms=%.3f alpha=0.85 base=0.15 dangling=drop\n", ms_since(t_init));
    if (app == DCG_APP_CC) LOG("[CC-SEMANTICS] directed_min_label input_edges=as_is updates=as_is\n");
    dcg_cache_stats cst;
    dcg_cache_maintain(eng, &cst);   // initial LoadCache after Start
    // diagnostic traces (hybrid_sssp.cu:407-414, 499-506)
    const std::string i16 = f.s("i16_repair_snapshot"), f1 = f.s("f1_cache_trace_file");
    if (!i16.empty() && app != DCG_APP_PR && (st = dcg_set_repair_snapshot(eng, i16.c_str(), 0)))
        return fail("i16_repair_snapshot", st);
    const uint64_t f1_cap = reference_cache_edges((int)f.d("cache"));
    auto trace_candidates = [&](uint32_t batch) {
        uint64_t admitted = 0;
        dcg_trace_stats ts;
        int s = dcg_hot_candidates(eng, 0, f1_cap, nullptr, 0, &admitted);
        if (!s) s = dcg_trace_cache_candidates(eng, batch, &ts);
        if (s) return s;
        LOG("[F1-CACHE-TRACE] batch=%u admitted_vertices=%llu admitted_edges=%llu changed_records=%llu "
            "capacity_edges=%llu bytes=%llu\n",
            batch, (unsigned long long)ts.admitted_vertices, (unsigned long long)ts.admitted_edges,
            (unsigned long long)ts.changed_records, (unsigned long long)ts.capacity_edges,
            (unsigned long long)ts.bytes);
        return 0;
    };
    if (!f1.empty()) {
        if ((st = dcg_set_cache_trace(eng, f1.c_str()))) return fail("f1_cache_trace_file", st);
        if ((st = trace_candidates(std::numeric_limits<uint32_t>::max()))) return fail("f1_cache_trace", st);
    }

    std::vector<std::pair<uint64_t, uint64_t>> batches;
    std::vector<dcg_edge> adds, dels;
    if (!f.s("updatefile").empty()) {
        if (!read_updates(f.s("update_size"), f.s("updatefile"), batches, adds, dels)) {
            std::fprintf(stderr, "[%s-ERROR] cannot read update files\n", kTag[app]);
            dcg_destroy(eng);
            return 1;
        }
        if ((st = dcg_load_updates(eng, adds.data(), adds.size(), dels.data(), dels.size())))
            return fail("load_updates", st);
    }
    const long long maxb = f.i(nm + "_max_batches");
    const uint64_t nb = std::min<uint64_t>((uint64_t)std::max(0LL, maxb), batches.size());
    if ((uint64_t)std::max(0LL, maxb) > batches.size() && app != DCG_APP_PR)
        std::printf("[%s] clamp %s_max_batches from %lld to available update batches %zu\n", kTag[app], nm.c_str(),
                    maxb, batches.size());
    bool success = true;
    std::vector<uint32_t> vals(n), par(n);
    std::vector<float> rank(n), res(n);
    Csr g;
    auto min_check = [&](const char *stage, uint64_t b, bool delete_stage) {
        if (dcg_gather_values(eng, 0, vals.data(), n) || !get_csr(eng, n, g)) return false;
        if (app == DCG_APP_CC) {
            const uint64_t err = cc_check(g, vals);
            LOG("[CC-%s-CHECK][batch %u] %s errors=%llu label_checksum=%llu\n", stage, (unsigned)b,
                err ? "failed" : "passed", (unsigned long long)err, (unsigned long long)value_checksum(vals));
            return err == 0;
        }
        dcg_gather_parents(eng, 0, par.data(), n);
        const BellmanResult r = check_bellman(app, g, vals, par, source);
        const bool src_ok = vals[source] == 0;
        const bool ok = src_ok && r.relaxable == 0 && r.missing_witness == 0;
        if (delete_stage)
            LOG("[%s-%s-CHECK][batch %d] %s source_ok=%d relaxable_edges=%llu missing_tight_witnesses=%llu "
                "invalid_parent_witness=%llu missing_parent_edges=%llu wrong_parent_distances=%llu first_src=%u "
                "first_dst=%u first_missing_witness=%u first_bad_vertex=%u first_bad_parent=%u distance_checksum=%llu\n",
                kTag[app], stage, (int)b, ok ? "passed" : "failed", src_ok ? 1 : 0, (unsigned long long)r.relaxable,
                (unsigned long long)r.missing_witness, (unsigned long long)r.invalid_parent,
                (unsigned long long)r.missing_parent_edges, (unsigned long long)r.wrong_parent, r.first_src,
                r.first_dst, r.first_missing, r.first_bad, r.first_bad_parent,
                (unsigned long long)value_checksum(vals));
        else
            LOG("[%s-%s-CHECK][batch %d] %s source_ok=%d relaxable_edges=%llu missing_tight_witnesses=%llu "
                "invalid_parent_witness=%llu missing_parent_edges=%llu wrong_parent_distances=%llu distance_checksum=%llu\n",
                kTag[app], stage, (int)b, ok ? "passed" : "failed", src_ok ? 1 : 0, (unsigned long long)r.relaxable,
                (unsigned long long)r.missing_witness, (unsigned long long)r.invalid_parent,
                (unsigned long long)r.missing_parent_edges, (unsigned long long)r.wrong_parent,
                (unsigned long long)value_checksum(vals));
        return ok;
    };
    auto pr_stage_check = [&](const char *stage, int b) {
        if (dcg_gather_values(eng, 0, rank.data(), n) || dcg_gather_buffers(eng, 0, res.data(), n) ||
            !get_csr(eng, n, g))
            return false;
        return pr_check(g, rank, res, cfg.pr_base, cfg.pr_epsilon, 1.0 / scale, stage, b);
    };
    if (app == DCG_APP_PR && check && !pr_stage_check("initial", -1)) success = false;

    uint64_t ab = 0, db = 0;
    for (uint64_t b = 0; b < nb; ++b) {
        const uint64_t na = batches[b].first, nd = batches[b].second;
        if (hints.batch == "auto") dcg_set_batch_maintenance(eng, na + nd >= 1000000 ? 1 : 0);
        if (app == DCG_APP_PR) {
            auto t = std::chrono::steady_clock::now();
            dcg_phase_stats ps;
            if ((st = dcg_update_pagerank(eng, ab, na, db, nd, &ps))) return fail("update_pagerank", st);
            dcg_cache_maintain(eng, &cst);
            if (!f1.empty() && (st = trace_candidates((uint32_t)b))) return fail("f1_cache_trace", st);
            LOG("[PR-BATCH] batch=%u touched_sources=%llu rounds=%llu stop=%s active=%llu limit=%u\n", (unsigned)b,
                (unsigned long long)ps.touched_sources, (unsigned long long)ps.rounds,
                ps.affected ? "iteration_limit" : "converged", (unsigned long long)ps.affected, cfg.pr_max_rounds);
            LOG("[P0-TIMER][PR][batch %u] paper_algorithm_ms=%.3f\n", (unsigned)b, ms_since(t));
            if (check && !pr_stage_check("batch", (int)b)) success = false;
        } else {
            std::printf("batch number %llu\n", (unsigned long long)b);
            auto t = std::chrono::steady_clock::now();
            dcg_phase_stats ds, as;
            if ((st = dcg_del_edge(eng, db, nd, &ds))) return fail("del_edge", st);
            double total_ms = ms_since(t);
            if (b == 0 && !i16.empty() && nd)
                LOG("[I16-SNAPSHOT][batch 0] states=%llu diagnostic_only=1\n",
                    (unsigned long long)snapshot_states(i16));
            if (check && !min_check("DELETE-STAGE", b, true)) success = false;
            if (cpu_cap) {
                // the CPU takes the cpu_cap partitions that receive the most of this
                // batch's insertion seeds (ties: the lower partition)
                std::vector<std::pair<long long, long long>> rank(segs);
                for (long long s = 0; s < segs; ++s) rank[s] = {0, s};
                for (uint64_t k = 0; k < na; ++k) {
                    const uint32_t v = adds[ab + k].v;
                    const long long s = (long long)(std::upper_bound(seg_start.begin(), seg_start.end(), v) -
                                                    seg_start.begin()) - 1;
                    if (v < n && s >= 0 && s < segs) --rank[s].first;
                }
                std::sort(rank.begin(), rank.end());
                std::vector<uint8_t> mask(n, 0);
                for (long long k = 0; k < cpu_cap; ++k)
                    std::fill(mask.begin() + (long)seg_start[rank[k].second],
                              mask.begin() + (long)seg_start[rank[k].second + 1], (uint8_t)1);
                if ((st = dcg_set_cpu_domain(eng, mask.data(), n))) return fail("sssp_cpu_partition_capacity", st);
            }
            t = std::chrono::steady_clock::now();
            if ((st = dcg_add_edge(eng, ab, na, &as))) return fail("add_edge", st);
            if (!dmap.empty() || cpu_cap)
                LOG("[E2-COEXEC][batch %d] gpu_closures=%llu cpu_relaxations=%llu gpu_to_cpu=%llu cpu_to_gpu=%llu\n",
                    (int)b, (unsigned long long)as.coexec_rounds, (unsigned long long)as.cpu_relaxations,
                    (unsigned long long)as.proposals_to_cpu, (unsigned long long)as.proposals_to_gpu);
            if (cfg.cache_policy == DCG_CACHE_LEARNED && b + 1 < nb)
                dcg_cache_prefetch(eng, ab + na, batches[b + 1].first, db + nd, batches[b + 1].second, &cst);
            dcg_cache_maintain(eng, &cst);
            total_ms += ms_since(t);
            if (!f1.empty() && (st = trace_candidates((uint32_t)b))) return fail("f1_cache_trace", st);
            LOG("[P0-ATTR][%s][batch %d] deletion=%.3f add=%.3f cache=%.3f affected=%llu rounds=%llu total=%.3f\n",
                kTag[app], (int)b, ds.ms, as.ms, cst.ms, (unsigned long long)ds.affected,
                (unsigned long long)(ds.rounds + as.rounds), total_ms);
            LOG("[P0-TIMER][%s][batch %d] total_batch: %.3f ms\n", kTag[app], (int)b, total_ms);
            if (check && !min_check("BATCH", b, false)) success = false;
        }
        ab += na;
        db += nd;
    }

    if (app == DCG_APP_PR) {
        dcg_gather_values(eng, 0, rank.data(), n);
        dcg_gather_buffers(eng, 0, res.data(), n);
        if (check && !pr_stage_check("final", (int)nb)) success = false;
        if (!f.s("output").empty()) {
            std::ofstream o(f.s("output"));
            o.precision(9);
            for (uint32_t v = 0; v < n; ++v) o << v << ' ' << rank[v] / scale << ' ' << res[v] / scale << '\n';
        }
        LOG("[PR-FINAL] vertices=%u batches=%llu\n", n, (unsigned long long)nb);
    } else {
        dcg_gather_values(eng, 0, vals.data(), n);
        if (app == DCG_APP_CC) {
            if (check) {
                get_csr(eng, n, g);
                const uint64_t err = cc_check(g, vals);
                success = success && err == 0;
                LOG("[CC-FINAL-CHECK] %s errors=%llu\n", err ? "failed" : "passed", (unsigned long long)err);
            }
            if (f.b("cc_print_checksum"))
                LOG("[CC-FINAL-CHECK] label_checksum=%llu\n", (unsigned long long)value_checksum(vals));
            if (!f.s("output").empty()) {
                std::ofstream o(f.s("output"));
                for (uint32_t v = 0; v < n; ++v) o << v << ' ' << vals[v] << '\n';
            }
        } else {
            dcg_gather_parents(eng, 0, par.data(), n);
            uint64_t reach = 0;
            for (uint32_t v = 0; v < n; ++v) reach += vals[v] != kInf;
            if (check) {
                get_csr(eng, n, g);
                const BellmanResult r = check_bellman(app, g, vals, par, source);
                const bool ok = vals[source] == 0 && r.relaxable == 0 && r.missing_witness == 0;
                success = success && ok;
                if (ok)
                    LOG("[%s-BELLMAN-CHECK] passed reachable=%llu relaxable_edges=0 missing_tight_witnesses=0\n",
                        kTag[app], (unsigned long long)reach);
                else
                    LOG("[%s-BELLMAN-CHECK] failed relaxable_edges=%llu missing_tight_witnesses=%llu\n", kTag[app],
                        (unsigned long long)r.relaxable, (unsigned long long)r.missing_witness);
            }
            const bool pc = f.b(nm + "_print_checksum");
            LOG("[%s-FINAL-CHECK] final_reachable=%llu checksum=%s\n", kTag[app], (unsigned long long)reach,
                pc ? "enabled" : "disabled");
            if (pc) {
                uint64_t dc = 1469598103934665603ull, pcs = 1099511628211ull;
                for (uint32_t v = 0; v < n; ++v) {
                    if (vals[v] == kInf) continue;
                    dc ^= (uint64_t)v + 0x9e3779b97f4a7c15ull + ((uint64_t)vals[v] << 6) + ((uint64_t)vals[v] >> 2);
                    dc *= 1099511628211ull;
                    pcs ^= (uint64_t)v + 0x517cc1b727220a95ull + ((uint64_t)par[v] << 7) + ((uint64_t)par[v] >> 3);
                    pcs *= 1469598103934665603ull;
                }
                LOG("[%s-FINAL-CHECK] distance_checksum=%llu parent_checksum=%llu\n", kTag[app],
                    (unsigned long long)dc, (unsigned long long)pcs);
            }
            if (!f.s("output").empty()) {
                std::ofstream o(f.s("output"));
                for (uint32_t v = 0; v < n; ++v) o << v << ' ' << vals[v] << ' ' << par[v] << ' ' << vals[v] << '\n';
            }
        }
    }
    dcg_destroy(eng);
    return success ? 0 : 2;
}
