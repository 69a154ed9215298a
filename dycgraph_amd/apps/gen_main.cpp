// dcg_gen -- writes benchmark inputs as files: a graph (market_big text or the
// binary edge list) and, optionally, an update stream (update + size files in
// the reference's formats, include/framework/Loader.h:30-101).  The reference
// produces its inputs with scripts; this tool writes the same seeded inputs
// the library's generators give in memory, so a file run and an in-memory run
// see identical records.
//
//   dcg_gen --graph=rmat --scale=16 --edges=1048576 [--a=0.57 --b=0.19 --c=0.19] [--seed=1]
//   dcg_gen --graph=grid --side=4096 [--diag=0.05] [--seed=1]
//     --out=FILE [--binary]
//     [--updates=FILE --sizes=FILE --snapshots=T (--adds=A --dels=D | --lognormal=mu,sigma,lo,hi)
//      [--stream_seed=7]]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/dcg.h"

static int usage(const char *why) {
    std::fprintf(stderr, "dcg_gen: %s\n", why);
    return 1;
}

int main(int argc, char **argv) {
    std::map<std::string, std::string> f = {{"graph", "rmat"}, {"scale", "16"}, {"edges", "1048576"}, {"a", "0.57"},
                                            {"b", "0.19"}, {"c", "0.19"}, {"seed", "1"}, {"side", "64"},
                                            {"diag", "0.05"}, {"out", ""}, {"binary", "0"}, {"updates", ""},
                                            {"sizes", ""}, {"snapshots", "10"}, {"adds", "1000"}, {"dels", "1000"},
                                            {"lognormal", ""}, {"stream_seed", "7"}};
    for (int i = 1; i < argc; ++i) {
        std::string t = argv[i];
        while (!t.empty() && t[0] == '-') t.erase(0, 1);
        const size_t eq = t.find('=');
        const std::string k = t.substr(0, eq), v = eq == std::string::npos ? "1" : t.substr(eq + 1);
        if (!f.count(k)) return usage(("unknown flag " + k).c_str());
        f[k] = v;
    }
    if (f["out"].empty()) return usage("--out is required");
    auto u64 = [&](const char *k) { return (uint64_t)std::strtoull(f[k].c_str(), nullptr, 10); };
    auto dbl = [&](const char *k) { return std::strtod(f[k].c_str(), nullptr); };
    uint32_t n = 0;
    std::vector<uint32_t> src, dst;
    if (f["graph"] == "rmat") {
        const uint32_t scale = (uint32_t)u64("scale");
        const uint64_t m = u64("edges");
        if (scale < 1 || scale > 30) return usage("scale must be 1..30");
        n = 1u << scale;
        src.resize(m);
        dst.resize(m);
        if (dcg_gen_rmat(scale, m, dbl("a"), dbl("b"), dbl("c"), u64("seed"), src.data(), dst.data()))
            return usage("dcg_gen_rmat failed");
    } else if (f["graph"] == "grid") {
        const uint32_t side = (uint32_t)u64("side");
        uint64_t m = 0;
        if (dcg_gen_grid(side, dbl("diag"), u64("seed"), nullptr, nullptr, &m)) return usage("dcg_gen_grid failed");
        src.resize(m);
        dst.resize(m);
        if (dcg_gen_grid(side, dbl("diag"), u64("seed"), src.data(), dst.data(), &m)) return usage("dcg_gen_grid failed");
        n = side * side;
    } else {
        return usage("--graph must be rmat or grid");
    }
    if (dcg_write_graph(f["out"].c_str(), n, src.data(), dst.data(), src.size(), f["binary"] != "0"))
        return usage("cannot write the graph file");
    std::printf("graph %s: %u vertices, %zu edges (%s)\n", f["out"].c_str(), n, src.size(),
                f["binary"] != "0" ? "binary" : "market_big");
    if (f["updates"].empty()) return 0;
    if (f["sizes"].empty()) return usage("--updates needs --sizes");
    const uint32_t T = (uint32_t)u64("snapshots");
    std::vector<uint64_t> ap(T), dp(T);
    if (!f["lognormal"].empty()) {
        double mu = 8, sigma = 1.5;
        unsigned long long lo = 100, hi = 100000;
        if (std::sscanf(f["lognormal"].c_str(), "%lf,%lf,%llu,%llu", &mu, &sigma, &lo, &hi) != 4)
            return usage("--lognormal=mu,sigma,lo,hi");
        std::vector<uint64_t> tot(T);
        if (dcg_gen_lognormal_sizes(T, mu, sigma, lo, hi, u64("seed") + 3, tot.data()))
            return usage("dcg_gen_lognormal_sizes failed");
        for (uint32_t t = 0; t < T; ++t) {
            ap[t] = tot[t] - tot[t] / 2;
            dp[t] = tot[t] / 2;
        }
    } else {
        for (uint32_t t = 0; t < T; ++t) {
            ap[t] = u64("adds");
            dp[t] = u64("dels");
        }
    }
    uint64_t na = 0, nd = 0;
    for (uint32_t t = 0; t < T; ++t) {
        na += ap[t];
        nd += dp[t];
    }
    std::vector<dcg_edge> adds(na), dels(nd);
    if (dcg_gen_stream(n, src.data(), dst.data(), src.size(), T, ap.data(), dp.data(), u64("stream_seed"), adds.data(),
                       dels.data()))
        return usage("dcg_gen_stream failed (the graph must be sorted without duplicates)");
    std::vector<uint64_t> sizes(2 * T);
    for (uint32_t t = 0; t < T; ++t) {
        sizes[2 * t] = ap[t];
        sizes[2 * t + 1] = dp[t];
    }
    if (dcg_write_updates(f["sizes"].c_str(), f["updates"].c_str(), sizes.data(), T, adds.data(), dels.data()))
        return usage("cannot write the update files");
    std::printf("updates %s, sizes %s: %u snapshots, %llu additions, %llu deletions\n", f["updates"].c_str(),
                f["sizes"].c_str(), T, (unsigned long long)na, (unsigned long long)nd);
    return 0;
}
