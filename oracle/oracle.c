/*
 * oracle.c -- CPU oracle for the incremental streaming-graph hot path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, never by the product library.
 *
 * What pins each part (tests/test_oracle_*.py):
 *   orc_store_*        reference KATs tests/source_local_chunk_store_test.cpp:35-56 and the
 *                      topology_replay fixture (CMakeLists.txt:121-131) -> golden strings.
 *   orc_sssp/orc_bfs   bfs_dynamic_smoke.py KAT graph (hand-derivable distances), brute-force
 *                      Bellman-Ford on tiny random graphs, closed forms (chains, cycles).
 *   orc_cc             cc_dynamic_smoke.py literal batches (labels derivable by hand) and
 *                      brute-force transitive closure on tiny graphs.
 *   orc_pagerank       closed forms (2-cycle, star, chain), dense linear solve (numpy) on
 *                      tiny graphs.
 *   orc_check_*        planted faults (wrong distance, missing witness, bad parent).
 *   orc_incr_*         from-scratch recompute after every phase (independent code path).
 *   orc_mlp_*          central finite differences.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

/* splitmix64 finalizer (Steele, Lea, Flood 2014). */
uint64_t orc_mix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

/* include/framework/variants/api.cuh:26-28 (reference mode); seeded mode is the
 * pure, stateless replacement for "uniform [1,W] weights from seed". */
uint32_t orc_weight(int mode, uint64_t seed, uint32_t wmax, uint32_t u, uint32_t v) {
    switch (mode) {
    case ORC_W_REFERENCE: return (uint32_t)(((uint64_t)u + v) % 128u + 1u);
    case ORC_W_SEEDED: return 1u + (uint32_t)(orc_mix64(seed ^ (((uint64_t)u << 32) | v)) % wmax);
    case ORC_W_UNIT: return 1u;
    default: return 0u;
    }
}

/* ------------------------------------------------------------------ store */
typedef struct { uint32_t *d; uint32_t deg, cap; } orc_row;
struct orc_store { uint32_t n; uint64_t m; orc_row *rows; };

static void row_push(orc_row *r, uint32_t v) {
    if (r->deg == r->cap) {
        uint32_t nc = r->cap ? r->cap * 2 : 4;
        r->d = (uint32_t *)realloc(r->d, (size_t)nc * sizeof(uint32_t));
        r->cap = nc;
    }
    r->d[r->deg++] = v;
}

/* TopologyReplayModel::ApplyDelete (include/framework/topology_replay.h:80-88):
 * erase the first occurrence, survivors keep their order. */
static int row_erase_first(orc_row *r, uint32_t v) {
    for (uint32_t i = 0; i < r->deg; ++i) {
        if (r->d[i] == v) {
            memmove(r->d + i, r->d + i + 1, (size_t)(r->deg - i - 1) * sizeof(uint32_t));
            r->deg--;
            return 1;
        }
    }
    return 0;
}

uint64_t orc_hash_sequence(const uint32_t *xs, uint64_t k) {
    /* include/framework/topology_replay.h:50-58 */
    uint64_t h = 1469598103934665603ull;
    for (uint64_t i = 0; i < k; ++i) { h ^= (uint64_t)xs[i]; h *= 1099511628211ull; }
    h ^= k; h *= 1099511628211ull;
    return h;
}

orc_store *orc_store_create(uint32_t n) {
    orc_store *s = (orc_store *)calloc(1, sizeof(orc_store));
    s->n = n;
    s->rows = (orc_row *)calloc(n ? n : 1, sizeof(orc_row));
    return s;
}

void orc_store_destroy(orc_store *s) {
    if (!s) return;
    for (uint32_t i = 0; i < s->n; ++i) free(s->rows[i].d);
    free(s->rows);
    free(s);
}

int orc_store_load(orc_store *s, const uint32_t *src, const uint32_t *dst, uint64_t m) {
    for (uint64_t i = 0; i < m; ++i) {
        if (src[i] >= s->n) return -1;   /* destinations are not range-checked by the model */
        row_push(&s->rows[src[i]], dst[i]);
        s->m++;
    }
    return 0;
}

static int cmp_u32(const void *a, const void *b) {
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? -1 : x > y;
}

/* Sorted unique sources touched by one phase (TopologyMutationBatch::TouchedSources). */
static uint32_t *touched_sources(const orc_edge *a, uint64_t na, const orc_edge *b, uint64_t nb,
                                 uint64_t *count) {
    uint32_t *t = (uint32_t *)malloc((size_t)(na + nb + 1) * sizeof(uint32_t));
    uint64_t k = 0;
    for (uint64_t i = 0; i < na; ++i) t[k++] = a[i].u;
    for (uint64_t i = 0; i < nb; ++i) t[k++] = b[i].u;
    qsort(t, (size_t)k, sizeof(uint32_t), cmp_u32);
    uint64_t w = 0;
    for (uint64_t i = 0; i < k; ++i) if (w == 0 || t[w - 1] != t[i]) t[w++] = t[i];
    *count = w;
    return t;
}

static uint64_t row_hash(const orc_store *s, uint32_t u) {
    if (u >= s->n) return 0;
    return orc_hash_sequence(s->rows[u].d, s->rows[u].deg);
}

/* src/topology_replay.cpp:152-195: all deletions first, then all additions;
 * a touched source counts as changed iff its row sequence changed. */
void orc_store_apply(orc_store *s, const orc_edge *dels, uint64_t nd,
                     const orc_edge *adds, uint64_t na, orc_batch_metrics *out) {
    orc_batch_metrics m;
    memset(&m, 0, sizeof(m));
    uint64_t nt = 0;
    uint32_t *t = touched_sources(dels, nd, adds, na, &nt);
    uint64_t *before = (uint64_t *)malloc((size_t)(nt + 1) * sizeof(uint64_t));
    uint32_t *bdeg = (uint32_t *)malloc((size_t)(nt + 1) * sizeof(uint32_t));
    for (uint64_t i = 0; i < nt; ++i) {
        before[i] = row_hash(s, t[i]);
        bdeg[i] = t[i] < s->n ? s->rows[t[i]].deg : 0;
    }
    for (uint64_t i = 0; i < nd; ++i) {
        if (dels[i].u >= s->n || !row_erase_first(&s->rows[dels[i].u], dels[i].v)) m.missing_deletes++;
        else s->m--;
    }
    for (uint64_t i = 0; i < na; ++i) {
        /* the model rejects only out-of-range sources (source_local_chunk_store.h:271-277) */
        if (adds[i].u >= s->n) { m.invalid_additions++; continue; }
        row_push(&s->rows[adds[i].u], adds[i].v);
        s->m++;
    }
    for (uint64_t i = 0; i < nt; ++i) {
        uint32_t d = t[i] < s->n ? s->rows[t[i]].deg : 0;
        if (d != bdeg[i] || row_hash(s, t[i]) != before[i]) m.changed_sources++;
    }
    m.touched_sources = nt;
    m.deletions = nd;
    m.additions = na;
    m.graph_edges = s->m;
    free(t); free(before); free(bdeg);
    if (out) *out = m;
}

void orc_store_apply_deletes(orc_store *s, const orc_edge *dels, uint64_t nd, orc_batch_metrics *out) {
    orc_store_apply(s, dels, nd, NULL, 0, out);
}

void orc_store_apply_adds(orc_store *s, const orc_edge *adds, uint64_t na, orc_batch_metrics *out) {
    orc_store_apply(s, NULL, 0, adds, na, out);
}

uint32_t orc_store_nodes(const orc_store *s) { return s->n; }
uint64_t orc_store_edges(const orc_store *s) { return s->m; }
uint32_t orc_store_degree(const orc_store *s, uint32_t u) { return u < s->n ? s->rows[u].deg : 0; }

uint32_t orc_store_row(const orc_store *s, uint32_t u, uint32_t *out, uint32_t cap) {
    if (u >= s->n) return 0;
    uint32_t d = s->rows[u].deg;
    memcpy(out, s->rows[u].d, (size_t)(d < cap ? d : cap) * sizeof(uint32_t));
    return d;
}

void orc_store_csr(const orc_store *s, uint64_t *rp, uint32_t *col) {
    uint64_t k = 0;
    for (uint32_t u = 0; u < s->n; ++u) {
        rp[u] = k;
        memcpy(col + k, s->rows[u].d, (size_t)s->rows[u].deg * sizeof(uint32_t));
        k += s->rows[u].deg;
    }
    rp[s->n] = k;
}

/* ------------------------------------------------------ from-scratch algorithms */
typedef struct { uint64_t key; uint32_t v; } heap_item;
typedef struct { heap_item *a; size_t n, cap; } heap_t;

static void heap_push(heap_t *h, uint64_t key, uint32_t v) {
    if (h->n == h->cap) {
        h->cap = h->cap ? h->cap * 2 : 1024;
        h->a = (heap_item *)realloc(h->a, h->cap * sizeof(heap_item));
    }
    size_t i = h->n++;
    while (i > 0) {
        size_t p = (i - 1) / 2;
        if (h->a[p].key <= key) break;
        h->a[i] = h->a[p];
        i = p;
    }
    h->a[i].key = key; h->a[i].v = v;
}

static heap_item heap_pop(heap_t *h) {
    heap_item top = h->a[0], last = h->a[--h->n];
    size_t i = 0;
    for (;;) {
        size_t c = 2 * i + 1;
        if (c >= h->n) break;
        if (c + 1 < h->n && h->a[c + 1].key < h->a[c].key) c++;
        if (h->a[c].key >= last.key) break;
        h->a[i] = h->a[c];
        i = c;
    }
    if (h->n) h->a[i] = last;
    return top;
}

/* Dijkstra with 64-bit candidate sums (tests/bfs_dynamic_smoke.py:16-40,
 * samples/hybrid_sssp/hybrid_sssp.cu:253-256). */
void orc_sssp(uint32_t n, const uint64_t *rp, const uint32_t *col, uint32_t source,
              int wmode, uint64_t seed, uint32_t wmax, uint32_t *dist) {
    for (uint32_t i = 0; i < n; ++i) dist[i] = ORC_INF;
    if (source >= n) return;
    heap_t h = {0, 0, 0};
    dist[source] = 0;
    heap_push(&h, 0, source);
    while (h.n) {
        heap_item it = heap_pop(&h);
        if (it.key != dist[it.v]) continue;
        for (uint64_t e = rp[it.v]; e < rp[it.v + 1]; ++e) {
            uint32_t d = col[e];
            uint64_t c = it.key + orc_weight(wmode, seed, wmax, it.v, d);
            if (c < dist[d]) { dist[d] = (uint32_t)c; heap_push(&h, c, d); }
        }
    }
    free(h.a);
}

/* FIFO BFS (tests/bfs_dynamic_smoke.py:31-38). */
void orc_bfs(uint32_t n, const uint64_t *rp, const uint32_t *col, uint32_t source, uint32_t *dist) {
    for (uint32_t i = 0; i < n; ++i) dist[i] = ORC_INF;
    if (source >= n) return;
    uint32_t *q = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
    size_t head = 0, tail = 0;
    dist[source] = 0;
    q[tail++] = source;
    while (head < tail) {
        uint32_t u = q[head++];
        for (uint64_t e = rp[u]; e < rp[u + 1]; ++e) {
            uint32_t v = col[e];
            if (dist[v] == ORC_INF) { dist[v] = dist[u] + 1; q[tail++] = v; }
        }
    }
    free(q);
}

/* Increasing-seed directed flood fill (samples/hybrid_cc/hybrid_cc_host.cu:8-32). */
void orc_cc(uint32_t n, const uint64_t *rp, const uint32_t *col, uint32_t *label) {
    for (uint32_t i = 0; i < n; ++i) label[i] = ORC_INF;
    uint32_t *stack = (uint32_t *)malloc((size_t)(n ? n : 1) * sizeof(uint32_t));
    for (uint32_t seed = 0; seed < n; ++seed) {
        if (label[seed] != ORC_INF) continue;
        size_t top = 0;
        label[seed] = seed;
        stack[top++] = seed;
        while (top) {
            uint32_t u = stack[--top];
            for (uint64_t e = rp[u]; e < rp[u + 1]; ++e) {
                uint32_t v = col[e];
                if (label[v] != ORC_INF) continue;
                label[v] = seed;
                stack[top++] = v;
            }
        }
    }
    free(stack);
}

/* FP64 Jacobi (samples/hybrid_pr/hybrid_pr_host.cu:15-33; base 0.15 there). */
static void pr_multiply(uint32_t n, const uint64_t *rp, const uint32_t *col, double base,
                        const double *x, double *next) {
    for (uint32_t v = 0; v < n; ++v) next[v] = base;
    for (uint32_t u = 0; u < n; ++u) {
        uint64_t deg = rp[u + 1] - rp[u];
        if (!deg) continue;
        double c = 0.85 * x[u] / (double)deg;
        for (uint64_t e = rp[u]; e < rp[u + 1]; ++e) next[col[e]] += c;
    }
}

int orc_pagerank(uint32_t n, const uint64_t *rp, const uint32_t *col, double base,
                 double tol_sum, int max_rounds, double *x) {
    double *next = (double *)malloc((size_t)(n ? n : 1) * sizeof(double));
    for (uint32_t v = 0; v < n; ++v) x[v] = 0.0;
    int rounds = -1;
    for (int r = 0; r < max_rounds; ++r) {
        pr_multiply(n, rp, col, base, x, next);
        double delta = 0;
        for (uint32_t v = 0; v < n; ++v) delta += fabs(next[v] - x[v]);
        memcpy(x, next, (size_t)n * sizeof(double));
        if (delta <= tol_sum) { rounds = r + 1; break; }
    }
    free(next);
    return rounds;
}

/* samples/hybrid_sssp/hybrid_sssp.cu:335-347 (also hybrid_cc.cu:204-216). */
uint64_t orc_value_checksum(const uint32_t *values, uint32_t n) {
    uint64_t c = 1469598103934665603ull;
    for (uint32_t v = 0; v < n; ++v) {
        if (values[v] == ORC_INF) continue;
        c ^= (uint64_t)v + 0x9e3779b97f4a7c15ull + ((uint64_t)values[v] << 6) + ((uint64_t)values[v] >> 2);
        c *= 1099511628211ull;
    }
    return c;
}

/* CheckBellmanRelaxed + CheckParentWitness (samples/hybrid_sssp/hybrid_sssp.cu:228-333). */
void orc_check_sssp(uint32_t n, const uint64_t *rp, const uint32_t *col, uint32_t source,
                    int wmode, uint64_t seed, uint32_t wmax, const uint32_t *dist,
                    const uint32_t *parent, orc_sssp_check *out) {
    orc_sssp_check r;
    memset(&r, 0, sizeof(r));
    uint8_t *tight = (uint8_t *)calloc(n ? n : 1, 1);
    if (source < n) tight[source] = 1;
    for (uint32_t u = 0; u < n; ++u) {
        if (dist[u] == ORC_INF) continue;
        for (uint64_t e = rp[u]; e < rp[u + 1]; ++e) {
            uint32_t v = col[e];
            if (v >= n) continue;
            uint64_t c = (uint64_t)dist[u] + orc_weight(wmode, seed, wmax, u, v);
            if (dist[v] != ORC_INF && c == dist[v]) tight[v] = 1;
            if (c < dist[v]) r.relaxable_edges++;
        }
    }
    for (uint32_t v = 0; v < n; ++v)
        if (dist[v] != ORC_INF && !tight[v]) r.missing_tight_witnesses++;
    if (parent) {
        for (uint32_t v = 0; v < n; ++v) {
            if (v == source || dist[v] == ORC_INF) continue;
            uint32_t p = parent[v];
            int valid = p < n && dist[p] != ORC_INF;
            if (valid) {
                int exists = 0;
                for (uint64_t e = rp[p]; e < rp[p + 1]; ++e) if (col[e] == v) { exists = 1; break; }
                int matches = (uint64_t)dist[p] + orc_weight(wmode, seed, wmax, p, v) == dist[v];
                if (!exists) r.missing_parent_edges++;
                else if (!matches) r.wrong_parent_distances++;
                valid = exists && matches;
            }
            if (!valid) r.invalid_parents++;
        }
    }
    r.source_ok = source < n && dist[source] == 0;
    free(tight);
    *out = r;
}

/* PageRankCheck (samples/hybrid_pr/hybrid_pr_host.cu:9-56), with the base term
 * and epsilon given in the caller's scale. */
void orc_check_pagerank(uint32_t n, const uint64_t *rp, const uint32_t *col, double base,
                        double eps, double max_abs, const float *rank, const float *residual,
                        orc_pr_check *out) {
    orc_pr_check r;
    memset(&r, 0, sizeof(r));
    double *ref = (double *)malloc((size_t)(n ? n : 1) * sizeof(double));
    double *act = (double *)calloc((size_t)(n ? n : 1), sizeof(double));
    double *next = (double *)malloc((size_t)(n ? n : 1) * sizeof(double));
    /* base*n*1e-12 == 1e-12*n at the reference's base of 0.15 per vertex, scaled. */
    double tol = 1e-12 * (base / 0.15) * (double)(n ? n : 1);
    r.converged = orc_pagerank(n, rp, col, base, tol, 10000, ref) > 0;
    for (uint32_t v = 0; v < n; ++v) act[v] = rank[v];
    pr_multiply(n, rp, col, base, act, next);
    r.finite = 1;
    for (uint32_t v = 0; v < n; ++v) {
        r.finite &= isfinite(rank[v]) && isfinite(residual[v]) && rank[v] >= -eps;
        double err = fabs((double)rank[v] - ref[v]);
        r.l1 += err;
        if (err > r.max_error) r.max_error = err;
        r.defect += fabs(next[v] - (double)rank[v] - (double)residual[v]);
        double ar = fabs((double)residual[v]);
        if (ar > r.max_residual) r.max_residual = ar;
    }
    double mass = 0;
    for (uint32_t v = 0; v < n; ++v) mass += fabs(ref[v]);
    double unit = base / 0.15; /* one "vertex" of mass in the caller's scale */
    r.roundoff = 64.0 * FLT_EPSILON * (mass > unit ? mass : unit);
    r.bound = ((double)n * eps + r.roundoff) / 0.15;
    /* the reference's acceptance (L1 bound, invariant, residual threshold), plus
     * a max-abs bound against the FP64 solve when the caller gives one */
    r.ok = r.finite && r.converged && r.l1 <= r.bound && r.defect <= r.roundoff &&
           r.max_residual <= (double)(float)eps && (max_abs <= 0 || r.max_error <= max_abs);
    free(ref); free(act); free(next);
    *out = r;
}

/* ------------------------------------------------ hotness ranking and admission */
/* comp_hotness_sssp (driver.cuh:495-517: score = hotness[0..3] summed, 0 where
 * the buffer is INF unless the program keeps signed residuals), the stable
 * descending sort of (score, id) (graph_datum.cuh:291-299: ties keep ascending
 * ids), extract_vtx_degree + ExclusiveSum (driver.cuh:548-560,
 * graph_datum.cuh:300-318) and CacheCandidateCount (framework.cuh:1321-1350),
 * written as the reference does it: all vertices when the last exclusive prefix
 * plus the last degree is below the cap, otherwise the binary search for the
 * first k whose prefix of k+1 vertices reaches the cap.  window holds the four
 * 8-bit counters of a vertex in one word; value (may be NULL) masks unreached
 * vertices.  Fills ids (n entries, the whole ranking) and returns the count. */
uint64_t orc_hot_candidates(uint32_t n, const uint32_t *window, const uint32_t *value, const uint32_t *degree,
                            uint64_t cap, uint32_t *ids) {
    if (n == 0) return 0;
    uint32_t *score = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
    uint64_t bucket[1021];
    memset(bucket, 0, sizeof(bucket));
    for (uint32_t v = 0; v < n; ++v) {
        const uint32_t w = window[v];
        uint32_t sc = (w & 0xFFu) + ((w >> 8) & 0xFFu) + ((w >> 16) & 0xFFu) + (w >> 24);
        if (value && value[v] == ORC_INF) sc = 0;
        score[v] = sc;
        bucket[sc]++;
    }
    /* counting sort, descending score; vertices of one score in ascending id */
    uint64_t start[1021], pos = 0;
    for (int sc = 1020; sc >= 0; --sc) { start[sc] = pos; pos += bucket[sc]; }
    for (uint32_t v = 0; v < n; ++v) ids[start[score[v]]++] = v;
    free(score);
    if (cap == 0) return 0;   /* num_of_cache == 0 */
    uint64_t *sum = (uint64_t *)malloc((size_t)n * sizeof(uint64_t));
    uint64_t run = 0;
    for (uint32_t i = 0; i < n; ++i) { sum[i] = run; run += degree[ids[i]]; }
    uint64_t k;
    if (sum[n - 1] + degree[ids[n - 1]] < cap) {
        k = n;
    } else {
        uint64_t low = 0, high = n - 1;
        while (low < high) {
            const uint64_t middle = low + (high - low) / 2;
            if (sum[middle + 1] < cap) low = middle + 1;
            else high = middle;
        }
        k = low;
    }
    free(sum);
    return k;
}

/* ------------------------------------------------------- CPU incremental path */
struct orc_incr {
    uint32_t n;
    int app, wmode;
    uint32_t source, wmax;
    uint64_t seed;
    orc_store *out, *in;
    uint32_t *value, *parent;
    uint8_t *mark;
    uint32_t *queue;
    /* PageRank */
    float *rank, *res;
    double pr_base, pr_eps;
    int pr_max_rounds;
    uint32_t *front, *next_front;
    uint64_t front_n;
    uint8_t *queued;
};

static uint32_t incr_w(const orc_incr *e, uint32_t u, uint32_t v) {
    if (e->app == ORC_APP_CC) return 0;
    if (e->app == ORC_APP_BFS) return 1;
    return orc_weight(e->wmode, e->seed, e->wmax, u, v);
}

orc_incr *orc_incr_create(uint32_t n, const uint64_t *rp, const uint32_t *col, int app,
                          uint32_t source, int wmode, uint64_t seed, uint32_t wmax,
                          double pr_base, double pr_eps, int pr_max_rounds) {
    orc_incr *e = (orc_incr *)calloc(1, sizeof(orc_incr));
    e->n = n; e->app = app; e->source = source; e->wmode = wmode; e->seed = seed; e->wmax = wmax;
    e->out = orc_store_create(n);
    e->in = orc_store_create(n);
    for (uint32_t u = 0; u < n; ++u)
        for (uint64_t k = rp[u]; k < rp[u + 1]; ++k) {
            row_push(&e->out->rows[u], col[k]);
            row_push(&e->in->rows[col[k]], u);
        }
    e->out->m = e->in->m = rp[n];
    size_t nn = n ? n : 1;
    e->value = (uint32_t *)malloc(nn * sizeof(uint32_t));
    e->parent = (uint32_t *)malloc(nn * sizeof(uint32_t));
    e->mark = (uint8_t *)calloc(nn, 1);
    e->queue = (uint32_t *)malloc(nn * sizeof(uint32_t));
    if (app == ORC_APP_PR) {
        e->rank = (float *)calloc(nn, sizeof(float));
        e->res = (float *)calloc(nn, sizeof(float));
        e->front = (uint32_t *)malloc(nn * sizeof(uint32_t));
        e->next_front = (uint32_t *)malloc(nn * sizeof(uint32_t));
        e->queued = (uint8_t *)calloc(nn, 1);
        e->pr_base = pr_base; e->pr_eps = pr_eps; e->pr_max_rounds = pr_max_rounds;
    }
    return e;
}

void orc_incr_destroy(orc_incr *e) {
    if (!e) return;
    orc_store_destroy(e->out); orc_store_destroy(e->in);
    free(e->value); free(e->parent); free(e->mark); free(e->queue);
    free(e->rank); free(e->res); free(e->front); free(e->next_front); free(e->queued);
    free(e);
}

void orc_incr_set_state(orc_incr *e, const uint32_t *value, const uint32_t *parent) {
    memcpy(e->value, value, (size_t)e->n * sizeof(uint32_t));
    memcpy(e->parent, parent, (size_t)e->n * sizeof(uint32_t));
}

static uint32_t init_value(const orc_incr *e, uint32_t v) {
    if (e->app == ORC_APP_CC) return v;
    return v == e->source ? 0u : ORC_INF;
}

/* FIFO push closure over e->queue[0..count); e->mark flags queued vertices.
 * The ring holds at most n entries because a vertex is queued at most once. */
static void incr_push_closure(orc_incr *e, size_t count) {
    uint32_t *q = e->queue;
    size_t n = e->n, head = 0, tail = count % (n ? n : 1);
    while (count) {
        uint32_t s = q[head];
        head = head + 1 == n ? 0 : head + 1;
        count--;
        e->mark[s] = 0;
        uint32_t vs = e->value[s];
        if (vs == ORC_INF) continue;
        orc_row *r = &e->out->rows[s];
        for (uint32_t k = 0; k < r->deg; ++k) {
            uint32_t d = r->d[k];
            uint64_t c = (uint64_t)vs + incr_w(e, s, d);
            if (c < e->value[d]) {
                e->value[d] = (uint32_t)c;
                e->parent[d] = s;
                if (!e->mark[d]) {
                    e->mark[d] = 1;
                    q[tail] = d;
                    tail = tail + 1 == n ? 0 : tail + 1;
                    count++;
                }
            }
        }
    }
}

void orc_incr_converge(orc_incr *e) {
    for (uint32_t v = 0; v < e->n; ++v) { e->value[v] = init_value(e, v); e->parent[v] = ORC_INF; }
    memset(e->mark, 0, e->n);
    size_t count = 0;
    if (e->app == ORC_APP_CC) {
        for (uint32_t v = 0; v < e->n; ++v) { e->queue[count++] = v; e->mark[v] = 1; }
    } else if (e->source < e->n) {
        e->queue[count++] = e->source;
        e->mark[e->source] = 1;
    }
    incr_push_closure(e, count);
}

/* Dependency predicate: api.cuh:43-49 (SSSP/BFS) and hybrid_cc.cu:60-64 (CC).
 * The source's own distance never depends on an edge. */
static int incr_dependency(const orc_incr *e, uint32_t s, uint32_t d) {
    if (e->app == ORC_APP_CC) return s != d && e->value[s] == e->value[d] && e->value[d] < d;
    if (d == e->source) return 0;
    uint32_t p = e->parent[d];
    int ptight = p != ORC_INF && e->value[p] != ORC_INF &&
                 (uint64_t)e->value[p] + incr_w(e, p, d) == (uint64_t)e->value[d];
    uint64_t cand = (uint64_t)e->value[s] + incr_w(e, s, d);
    return p == s || (e->value[s] != ORC_INF && cand == (uint64_t)e->value[d] && !ptight);
}

/* Delete stage: seed (driver.cuh:394-441), invalidation closure over the
 * pre-delete rows (framework.cuh:4884-4924, push_functor.h:1154-1180), reset
 * (driver.cuh:443-463), physical delete, then repair: one pull-min over the
 * post-delete in-rows (framework.cuh:189-263, ties to the smallest source)
 * followed by a push closure to the same fixed point. */
void orc_incr_delete_phase(orc_incr *e, const orc_edge *dels, uint64_t nd) {
    uint32_t n = e->n;
    uint32_t *aff = (uint32_t *)malloc((size_t)(n ? n : 1) * sizeof(uint32_t));
    uint8_t *is_aff = (uint8_t *)calloc(n ? n : 1, 1);
    size_t na = 0;
    for (uint64_t i = 0; i < nd; ++i) {
        uint32_t u = dels[i].u, v = dels[i].v;
        if (u >= n || v >= n || is_aff[v]) continue;
        if (incr_dependency(e, u, v)) { is_aff[v] = 1; aff[na++] = v; }
    }
    for (size_t k = 0; k < na; ++k) e->parent[aff[k]] = ORC_INF;
    for (size_t k = 0; k < na; ++k) {
        uint32_t s = aff[k];
        orc_row *r = &e->out->rows[s];
        for (uint32_t j = 0; j < r->deg; ++j) {
            uint32_t d = r->d[j];
            if (is_aff[d]) continue;
            if (incr_dependency(e, s, d)) {
                e->parent[d] = ORC_INF;
                is_aff[d] = 1;
                aff[na++] = d;
            }
        }
    }
    for (size_t k = 0; k < na; ++k) { e->value[aff[k]] = init_value(e, aff[k]); e->parent[aff[k]] = ORC_INF; }
    for (uint64_t i = 0; i < nd; ++i) {
        uint32_t u = dels[i].u, v = dels[i].v;
        if (u >= n || v >= n) continue;
        if (row_erase_first(&e->out->rows[u], v)) {
            e->out->m--;
            row_erase_first(&e->in->rows[v], u);
            e->in->m--;
        }
    }
    memset(e->mark, 0, n);
    size_t count = 0;
    for (size_t k = 0; k < na; ++k) {
        uint32_t v = aff[k];
        uint32_t best = e->value[v], bp = ORC_INF;
        orc_row *r = &e->in->rows[v];
        for (uint32_t j = 0; j < r->deg; ++j) {
            uint32_t x = r->d[j];
            if (e->value[x] == ORC_INF) continue;
            uint64_t c = (uint64_t)e->value[x] + incr_w(e, x, v);
            if (c < best || (c == best && bp != ORC_INF && x < bp)) { best = (uint32_t)c; bp = x; }
        }
        if (best < e->value[v]) { e->value[v] = best; e->parent[v] = bp; }
        if (e->value[v] != ORC_INF && !e->mark[v]) { e->mark[v] = 1; e->queue[count++] = v; }
    }
    incr_push_closure(e, count);
    free(aff);
    free(is_aff);
}

/* Add stage: append (source_local_chunk_store.h:809-815), seed from the added
 * edges (framework.cuh:297-370), then the closure to quiescence. */
void orc_incr_add_phase(orc_incr *e, const orc_edge *adds, uint64_t na) {
    uint32_t n = e->n;
    memset(e->mark, 0, n);
    size_t count = 0;
    for (uint64_t i = 0; i < na; ++i) {
        uint32_t u = adds[i].u, v = adds[i].v;
        if (u >= n || v >= n) continue;
        row_push(&e->out->rows[u], v); e->out->m++;
        row_push(&e->in->rows[v], u); e->in->m++;
    }
    for (uint64_t i = 0; i < na; ++i) {
        uint32_t u = adds[i].u, v = adds[i].v;
        if (u >= n || v >= n || e->value[u] == ORC_INF) continue;
        uint64_t c = (uint64_t)e->value[u] + incr_w(e, u, v);
        if (c < e->value[v]) {
            e->value[v] = (uint32_t)c;
            e->parent[v] = u;
            if (!e->mark[v]) { e->mark[v] = 1; e->queue[count++] = v; }
        }
    }
    incr_push_closure(e, count);
}

void orc_incr_values(const orc_incr *e, uint32_t *value, uint32_t *parent) {
    if (value) memcpy(value, e->value, (size_t)e->n * sizeof(uint32_t));
    if (parent) memcpy(parent, e->parent, (size_t)e->n * sizeof(uint32_t));
}

uint64_t orc_incr_edges(const orc_incr *e) { return e->out->m; }
void orc_incr_csr(const orc_incr *e, uint64_t *rp, uint32_t *col) { orc_store_csr(e->out, rp, col); }

/* ---- PageRank signed residual push (include/framework/pr_residual.cuh:21-188) ---- */
static void pr_add(orc_incr *e, uint32_t d, float c) {
    float old = e->res[d];
    float val = old + c;
    e->res[d] = val;
    float eps = (float)e->pr_eps;
    if (fabsf(old) <= eps && fabsf(val) > eps && !e->queued[d]) {
        e->queued[d] = 1;
        e->next_front[e->front_n++] = d;
    }
}

static void pr_scatter_row(orc_incr *e, uint32_t s, float delta_over_deg) {
    orc_row *r = &e->out->rows[s];
    for (uint32_t k = 0; k < r->deg; ++k) pr_add(e, r->d[k], delta_over_deg);
}

static void pr_converge(orc_incr *e, uint64_t count) {
    /* e->front holds `count` queued vertices. */
    float eps = (float)e->pr_eps;
    float *deltas = (float *)malloc((size_t)(count ? count : 1) * sizeof(float));
    for (int round = 0; count && round < e->pr_max_rounds; ++round) {
        if (count > e->n) count = e->n;
        deltas = (float *)realloc(deltas, (size_t)(count ? count : 1) * sizeof(float));
        for (uint64_t i = 0; i < count; ++i) {
            uint32_t v = e->front[i];
            float r = e->res[v];
            e->queued[v] = 0;
            deltas[i] = 0;
            if (fabsf(r) > eps) { e->rank[v] += r; e->res[v] = 0; deltas[i] = r; }
        }
        e->front_n = 0;
        for (uint64_t i = 0; i < count; ++i) {
            uint32_t s = e->front[i];
            uint32_t deg = e->out->rows[s].deg;
            if (deg && deltas[i] != 0) pr_scatter_row(e, s, 0.85f * deltas[i] / (float)deg);
        }
        uint32_t *t = e->front; e->front = e->next_front; e->next_front = t;
        count = e->front_n;
    }
    free(deltas);
    /* unfinished work carries over (Runtime::BeginBatch copies the queue) */
    e->front_n = count;
}

void orc_incr_pr_start(orc_incr *e) {
    for (uint32_t v = 0; v < e->n; ++v) {
        e->rank[v] = 0; e->res[v] = (float)e->pr_base; e->front[v] = v; e->queued[v] = 1;
    }
    pr_converge(e, e->n);
}

void orc_incr_pr_batch(orc_incr *e, const orc_edge *dels, uint64_t nd, const orc_edge *adds, uint64_t na) {
    uint64_t nt = 0;
    uint32_t *t = touched_sources(dels, nd, adds, na, &nt);
    uint64_t carried = e->front_n;
    memcpy(e->next_front, e->front, (size_t)carried * sizeof(uint32_t));
    e->front_n = carried;
    for (uint64_t i = 0; i < nt; ++i) {
        uint32_t s = t[i];
        if (s >= e->n) continue;
        uint32_t deg = e->out->rows[s].deg;
        if (deg && e->rank[s] != 0) pr_scatter_row(e, s, 0.85f * e->rank[s] * -1.0f / (float)deg);
    }
    for (uint64_t i = 0; i < nd; ++i) {
        uint32_t u = dels[i].u, v = dels[i].v;
        if (u >= e->n || v >= e->n) continue;
        if (row_erase_first(&e->out->rows[u], v)) { e->out->m--; row_erase_first(&e->in->rows[v], u); e->in->m--; }
    }
    for (uint64_t i = 0; i < na; ++i) {
        uint32_t u = adds[i].u, v = adds[i].v;
        if (u >= e->n || v >= e->n) continue;
        row_push(&e->out->rows[u], v); e->out->m++;
        row_push(&e->in->rows[v], u); e->in->m++;
    }
    for (uint64_t i = 0; i < nt; ++i) {
        uint32_t s = t[i];
        if (s >= e->n) continue;
        uint32_t deg = e->out->rows[s].deg;
        if (deg && e->rank[s] != 0) pr_scatter_row(e, s, 0.85f * e->rank[s] * 1.0f / (float)deg);
    }
    free(t);
    uint32_t *x = e->front; e->front = e->next_front; e->next_front = x;
    pr_converge(e, e->front_n);
}

void orc_incr_pr_values(const orc_incr *e, float *rank, float *residual) {
    if (rank) memcpy(rank, e->rank, (size_t)e->n * sizeof(float));
    if (residual) memcpy(residual, e->res, (size_t)e->n * sizeof(float));
}

/* --------------------------------------------------------------- MLP (FP64) */
void orc_mlp_forward_backward(int rows, const float *x, const float *y,
                              const float *w1, const float *b1, const float *w2,
                              const float *b2, const float *w3, const float *b3,
                              double *out, double *loss,
                              double *gw1, double *gb1, double *gw2, double *gb2,
                              double *gw3, double *gb3) {
    enum { I = 8, H = 64, O = 2 };
    double h1[H], h2[H], g2[H], g1[H];
    double L = 0;
    if (gw1) {
        memset(gw1, 0, sizeof(double) * I * H); memset(gb1, 0, sizeof(double) * H);
        memset(gw2, 0, sizeof(double) * H * H); memset(gb2, 0, sizeof(double) * H);
        memset(gw3, 0, sizeof(double) * H * O); memset(gb3, 0, sizeof(double) * O);
    }
    for (int r = 0; r < rows; ++r) {
        const float *xr = x + (size_t)r * I;
        for (int j = 0; j < H; ++j) {
            double a = b1[j];
            for (int i = 0; i < I; ++i) a += (double)xr[i] * w1[i * H + j];
            h1[j] = a > 0 ? a : 0;
        }
        for (int j = 0; j < H; ++j) {
            double a = b2[j];
            for (int i = 0; i < H; ++i) a += h1[i] * w2[i * H + j];
            h2[j] = a > 0 ? a : 0;
        }
        double o[O];
        for (int k = 0; k < O; ++k) {
            double a = b3[k];
            for (int i = 0; i < H; ++i) a += h2[i] * w3[i * O + k];
            o[k] = a;
            if (out) out[(size_t)r * O + k] = a;
        }
        if (!y) continue;
        double t0 = y[(size_t)r * O], t1 = y[(size_t)r * O + 1];
        /* stable BCE with logits */
        double z = o[0];
        L += (z > 0 ? z : 0) - z * t0 + log1p(exp(-fabs(z)));
        L += 0.5 * (o[1] - t1) * (o[1] - t1);
        if (!gw1) continue;
        double go[O];
        go[0] = (1.0 / (1.0 + exp(-z)) - t0) / rows;
        go[1] = (o[1] - t1) / rows;
        for (int k = 0; k < O; ++k) {
            gb3[k] += go[k];
            for (int i = 0; i < H; ++i) gw3[i * O + k] += h2[i] * go[k];
        }
        for (int i = 0; i < H; ++i) {
            double a = 0;
            for (int k = 0; k < O; ++k) a += w3[i * O + k] * go[k];
            g2[i] = h2[i] > 0 ? a : 0;
        }
        for (int j = 0; j < H; ++j) {
            gb2[j] += g2[j];
            for (int i = 0; i < H; ++i) gw2[i * H + j] += h1[i] * g2[j];
        }
        for (int i = 0; i < H; ++i) {
            double a = 0;
            for (int j = 0; j < H; ++j) a += w2[i * H + j] * g2[j];
            g1[i] = h1[i] > 0 ? a : 0;
        }
        for (int j = 0; j < H; ++j) {
            gb1[j] += g1[j];
            for (int i = 0; i < I; ++i) gw1[i * H + j] += (double)xr[i] * g1[j];
        }
    }
    if (loss) *loss = rows ? L / rows : 0;
}
