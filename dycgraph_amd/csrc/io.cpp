// io.cpp -- graph and update files in the reference's formats, plus a binary
// edge list for large inputs.
//
// market_big (src/utils/parser.cpp:296-406): one "src dst [w]" line per edge,
// n = largest id + 1, edges kept in file order.  The reference reads the file
// twice on one thread; here the file is mapped once and split at line breaks
// into one range per thread: each thread counts its lines, the counts' prefix
// sum gives its output offset, and a second pass over the (cached) range
// parses straight into place, so the edge order is the file's.
//
// Updates (include/framework/Loader.h:30-101): a size file of "<adds> <dels>"
// lines, one per batch, and an update file of "a u v w" / "d u v w" lines whose
// additions and deletions are consumed in file order (the weight is replaced
// by the program's weight when the engine reads the records).
//
// Binary edge list: "DCGEDGE1" (8 bytes), u32 version 1, u32 num_nodes, u64
// num_edges, then num_edges u32 sources and num_edges u32 destinations.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/dcg.h"

namespace {

constexpr char kBinMagic[8] = {'D', 'C', 'G', 'E', 'D', 'G', 'E', '1'};

struct Mapped {
    const char *p = nullptr;
    size_t n = 0;
    int fd = -1;
    bool open(const char *path) {
        fd = ::open(path, O_RDONLY);
        if (fd < 0) return false;
        struct stat st;
        if (fstat(fd, &st) != 0) return false;
        n = (size_t)st.st_size;
        if (n == 0) return true;
        void *m = mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
        if (m == MAP_FAILED) return false;
        p = static_cast<const char *>(m);
        madvise(m, n, MADV_SEQUENTIAL);
        return true;
    }
    ~Mapped() {
        if (p) munmap(const_cast<char *>(p), n);
        if (fd >= 0) ::close(fd);
    }
};

int thread_count(int threads) {
    if (threads > 0) return std::min(threads, 256);
    const unsigned h = std::thread::hardware_concurrency();
    return h ? (int)std::min(h, 64u) : 1;
}

// Range boundaries at line starts: [cut[t], cut[t + 1]) for t < T.
std::vector<size_t> line_cuts(const char *p, size_t n, int T) {
    std::vector<size_t> cut(T + 1, n);
    cut[0] = 0;
    for (int t = 1; t < T; ++t) {
        size_t c = std::max(cut[t - 1], n / T * (size_t)t);
        while (c < n && c > 0 && p[c - 1] != '\n') ++c;
        cut[t] = c;
    }
    return cut;
}

inline bool next_uint(const char *&p, const char *e, uint64_t &x) {
    while (p < e && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
    if (p >= e || *p < '0' || *p > '9') return false;
    x = 0;
    while (p < e && *p >= '0' && *p <= '9') x = x * 10 + (uint64_t)(*p++ - '0');
    return true;
}

inline const char *line_end(const char *p, const char *e) {
    const void *nl = memchr(p, '\n', (size_t)(e - p));
    return nl ? static_cast<const char *>(nl) : e;
}

// Run body(t) on T threads (t = 0 on the caller).
template <class F>
void parallel(int T, F body) {
    std::vector<std::thread> pool;
    for (int t = 1; t < T; ++t) pool.emplace_back(body, t);
    body(0);
    for (auto &th : pool) th.join();
}

// One pass over a range of "u v ..." lines: count (out == nullptr) or write.
uint64_t edge_lines(const char *p, const char *e, uint32_t *src, uint32_t *dst, uint64_t &maxid, bool &bad) {
    uint64_t k = 0;
    while (p < e) {
        const char *le = line_end(p, e);
        const char *q = p;
        uint64_t u, v;
        if (next_uint(q, le, u) && next_uint(q, le, v)) {
            if (u > 0xFFFFFFFEull || v > 0xFFFFFFFEull) bad = true;
            if (src) {
                src[k] = (uint32_t)u;
                dst[k] = (uint32_t)v;
                maxid = std::max(maxid, std::max(u, v));
            }
            ++k;
        }
        p = le + 1;
    }
    return k;
}

template <class T>
T *alloc_array(uint64_t n) {
    return static_cast<T *>(std::malloc((size_t)std::max<uint64_t>(n, 1) * sizeof(T)));
}

}  // namespace

extern "C" {

void dcg_free(void *p) { std::free(p); }

int dcg_read_graph(const char *path, uint32_t *num_nodes, uint32_t **src, uint32_t **dst, uint64_t *num_edges,
                   int threads) {
    if (!path || !num_nodes || !src || !dst || !num_edges) return DCG_ERR_ARG;
    *src = *dst = nullptr;
    Mapped f;
    if (!f.open(path)) return DCG_ERR_IO;
    if (f.n >= 24 && std::memcmp(f.p, kBinMagic, 8) == 0) {
        uint32_t version, n;
        uint64_t m;
        std::memcpy(&version, f.p + 8, 4);
        std::memcpy(&n, f.p + 12, 4);
        std::memcpy(&m, f.p + 16, 8);
        if (version != 1 || f.n != 24 + m * 8) return DCG_ERR_IO;
        uint32_t *s = alloc_array<uint32_t>(m), *d = alloc_array<uint32_t>(m);
        if (!s || !d) {
            std::free(s);
            std::free(d);
            return DCG_ERR_CAPACITY;
        }
        const int T = thread_count(threads);
        parallel(T, [&](int t) {   // parallel copies out of the page cache
            const uint64_t lo = m * t / T, hi = m * (t + 1) / T;
            std::memcpy(s + lo, f.p + 24 + lo * 4, (hi - lo) * 4);
            std::memcpy(d + lo, f.p + 24 + m * 4 + lo * 4, (hi - lo) * 4);
        });
        for (uint64_t i = 0; i < m; ++i)
            if (s[i] >= n || d[i] >= n) {
                std::free(s);
                std::free(d);
                return DCG_ERR_IO;
            }
        *num_nodes = n;
        *num_edges = m;
        *src = s;
        *dst = d;
        return DCG_OK;
    }
    const int T = f.n < (1u << 20) ? 1 : thread_count(threads);
    const std::vector<size_t> cut = line_cuts(f.p, f.n, T);
    std::vector<uint64_t> count(T + 1, 0), maxid(T, 0);
    std::vector<char> bad(T, 0);
    parallel(T, [&](int t) {
        uint64_t mx = 0;
        bool b = false;
        count[t + 1] = edge_lines(f.p + cut[t], f.p + cut[t + 1], nullptr, nullptr, mx, b);
        bad[t] = b;
    });
    for (int t = 0; t < T; ++t) count[t + 1] += count[t];
    for (char b : bad)
        if (b) return DCG_ERR_IO;   // ids must fit 32 bits
    const uint64_t m = count[T];
    uint32_t *s = alloc_array<uint32_t>(m), *d = alloc_array<uint32_t>(m);
    if (!s || !d) {
        std::free(s);
        std::free(d);
        return DCG_ERR_CAPACITY;
    }
    parallel(T, [&](int t) {
        bool b = false;
        edge_lines(f.p + cut[t], f.p + cut[t + 1], s + count[t], d + count[t], maxid[t], b);
    });
    const uint64_t mx = m ? *std::max_element(maxid.begin(), maxid.end()) : 0;
    *num_nodes = m ? (uint32_t)(mx + 1) : 0;
    *num_edges = m;
    *src = s;
    *dst = d;
    return DCG_OK;
}

int dcg_write_graph(const char *path, uint32_t num_nodes, const uint32_t *src, const uint32_t *dst,
                    uint64_t num_edges, int binary) {
    if (!path || (num_edges && (!src || !dst))) return DCG_ERR_ARG;
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return DCG_ERR_IO;
    bool ok = true;
    if (binary) {
        const uint32_t version = 1;
        ok = std::fwrite(kBinMagic, 1, 8, fp) == 8 && std::fwrite(&version, 4, 1, fp) == 1 &&
             std::fwrite(&num_nodes, 4, 1, fp) == 1 && std::fwrite(&num_edges, 8, 1, fp) == 1 &&
             std::fwrite(src, 4, num_edges, fp) == num_edges && std::fwrite(dst, 4, num_edges, fp) == num_edges;
    } else {
        std::vector<char> buf;
        buf.reserve(1 << 22);
        char line[32];
        for (uint64_t i = 0; i < num_edges && ok; ++i) {
            const int k = std::snprintf(line, sizeof line, "%u %u\n", src[i], dst[i]);
            buf.insert(buf.end(), line, line + k);
            if (buf.size() > (1u << 22) - 32) {
                ok = std::fwrite(buf.data(), 1, buf.size(), fp) == buf.size();
                buf.clear();
            }
        }
        if (ok && !buf.empty()) ok = std::fwrite(buf.data(), 1, buf.size(), fp) == buf.size();
    }
    ok = std::fclose(fp) == 0 && ok;
    return ok ? DCG_OK : DCG_ERR_IO;
}

int dcg_read_updates(const char *sizes_path, const char *updates_path, uint64_t **sizes, uint64_t *num_batches,
                     dcg_edge **adds, uint64_t *num_adds, dcg_edge **dels, uint64_t *num_dels, int threads) {
    if (!sizes_path || !updates_path || !sizes || !num_batches || !adds || !num_adds || !dels || !num_dels)
        return DCG_ERR_ARG;
    *sizes = nullptr;
    *adds = *dels = nullptr;
    std::vector<uint64_t> sz;
    {
        Mapped f;
        if (!f.open(sizes_path)) return DCG_ERR_IO;
        const char *p = f.p, *e = f.p + f.n;
        while (p < e) {
            const char *le = line_end(p, e);
            const char *q = p;
            uint64_t a, d;
            if (next_uint(q, le, a) && next_uint(q, le, d)) {
                sz.push_back(a);
                sz.push_back(d);
            }
            p = le + 1;
        }
    }
    Mapped f;
    if (!f.open(updates_path)) return DCG_ERR_IO;
    const int T = f.n < (1u << 20) ? 1 : thread_count(threads);
    const std::vector<size_t> cut = line_cuts(f.p, f.n, T);
    // per thread: additions and deletions in its range (count, then write)
    std::vector<uint64_t> ca(T + 1, 0), cd(T + 1, 0);
    auto pass = [&](int t, dcg_edge *oa, dcg_edge *od) {
        const char *p = f.p + cut[t], *e = f.p + cut[t + 1];
        uint64_t ka = 0, kd = 0;
        while (p < e) {
            const char *le = line_end(p, e);
            const char *q = p;
            while (q < le && (*q == ' ' || *q == '\t')) ++q;
            const char kind = q < le ? *q : 0;
            uint64_t u, v;
            if ((kind == 'a' || kind == 'd') && (++q, next_uint(q, le, u)) && next_uint(q, le, v)) {
                if (kind == 'a') {
                    if (oa) oa[ka] = dcg_edge{(uint32_t)u, (uint32_t)v};
                    ++ka;
                } else {
                    if (od) od[kd] = dcg_edge{(uint32_t)u, (uint32_t)v};
                    ++kd;
                }
            }
            p = le + 1;
        }
        if (!oa) {
            ca[t + 1] = ka;
            cd[t + 1] = kd;
        }
    };
    parallel(T, [&](int t) { pass(t, nullptr, nullptr); });
    for (int t = 0; t < T; ++t) {
        ca[t + 1] += ca[t];
        cd[t + 1] += cd[t];
    }
    dcg_edge *a = alloc_array<dcg_edge>(ca[T]), *d = alloc_array<dcg_edge>(cd[T]);
    uint64_t *s = alloc_array<uint64_t>(sz.size());
    if (!a || !d || !s) {
        std::free(a);
        std::free(d);
        std::free(s);
        return DCG_ERR_CAPACITY;
    }
    parallel(T, [&](int t) { pass(t, a + ca[t], d + cd[t]); });
    std::copy(sz.begin(), sz.end(), s);
    *sizes = s;
    *num_batches = sz.size() / 2;
    *adds = a;
    *num_adds = ca[T];
    *dels = d;
    *num_dels = cd[T];
    return DCG_OK;
}

int dcg_write_updates(const char *sizes_path, const char *updates_path, const uint64_t *sizes, uint64_t num_batches,
                      const dcg_edge *adds, const dcg_edge *dels) {
    if (!sizes_path || !updates_path || (num_batches && !sizes)) return DCG_ERR_ARG;
    uint64_t na = 0, nd = 0;
    for (uint64_t b = 0; b < num_batches; ++b) {
        na += sizes[2 * b];
        nd += sizes[2 * b + 1];
    }
    if ((na && !adds) || (nd && !dels)) return DCG_ERR_ARG;
    FILE *fs = std::fopen(sizes_path, "wb");
    if (!fs) return DCG_ERR_IO;
    bool ok = true;
    for (uint64_t b = 0; b < num_batches && ok; ++b)
        ok = std::fprintf(fs, "%llu %llu\n", (unsigned long long)sizes[2 * b], (unsigned long long)sizes[2 * b + 1]) > 0;
    ok = std::fclose(fs) == 0 && ok;
    FILE *fu = std::fopen(updates_path, "wb");
    if (!fu) return DCG_ERR_IO;
    // per batch: its deletions, then its additions; w = (u + v) % 128 + 1
    uint64_t ia = 0, id = 0;
    for (uint64_t b = 0; b < num_batches && ok; ++b) {
        for (uint64_t k = 0; k < sizes[2 * b + 1] && ok; ++k, ++id)
            ok = std::fprintf(fu, "d %u %u %u\n", dels[id].u, dels[id].v, (dels[id].u + dels[id].v) % 128 + 1) > 0;
        for (uint64_t k = 0; k < sizes[2 * b] && ok; ++k, ++ia)
            ok = std::fprintf(fu, "a %u %u %u\n", adds[ia].u, adds[ia].v, (adds[ia].u + adds[ia].v) % 128 + 1) > 0;
    }
    ok = std::fclose(fu) == 0 && ok;
    return ok ? DCG_OK : DCG_ERR_IO;
}

}  // extern "C"
