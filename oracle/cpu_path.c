/*
 * cpu_path.c -- the reference's CPU-capable incremental path on all host cores.
 *
 * TEST INFRASTRUCTURE ONLY: the timed CPU baseline of bench.py and a checker in
 * tests/ (its values are compared with the from-scratch oracle).  The product
 * library never links it.
 *
 * Semantics are the reference's (the same as orc_incr in oracle.c, which is
 * the single-threaded restatement): SSSP/BFS/CC queries over one multiset store
 * of out-rows and in-rows; per snapshot
 *   delete stage: seeds (driver.cuh:394-441, the dependency predicate of
 *                 api.cuh:43-49 and, for CC, hybrid_cc.cu:60-64), invalidation
 *                 closure over the pre-delete rows (framework.cuh:4884-4924),
 *                 reset (driver.cuh:443-463), physical delete of first
 *                 occurrences (source_local_chunk_store.h:685-821), pull-min
 *                 repair over the in-rows of the affected set
 *                 (framework.cuh:189-263), push closure to the fixed point;
 *   add stage:    ordered appends (source_local_chunk_store.h:809-815), seeds
 *                 from the added edges (framework.cuh:297-370), push closure
 *                 (framework.cuh:5303-5349 semantics).
 * Parallel form: mutation is grouped by row (a stable two-pass radix sort of
 * the records by row id) and each touched row is rewritten by one worker in
 * one pass (the reference's source-partitioned worker pool,
 * source_local_chunk_store.h:117-123, 219-406); closures are level-synchronous
 * with a 64-bit atomic min on (value << 32 | parent) per relaxation and a round
 * stamp per vertex for frontier deduplication; invalidation claims use an
 * atomic flag per vertex.  Rows start as power-of-two blocks (at least 4 slots,
 * source_local_chunk_store.h:133-145) carved from one slab per direction and
 * move to a block twice as large when they outgrow it.  Rows built here hold
 * each vertex's entries in an arbitrary order (parallel placement): the
 * multiset, and so every value, is the same.
 */
#include "oracle.h"

#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define HEAP_BIT 0x80000000u   /* the row's block was allocated on its own */
/* Work below these sizes runs on the calling thread: a parallel region costs
 * more than a few thousand relaxations or rows. */
#define PAR_FRONT 2048     /* frontier entries of a closure round */
#define PAR_RECS 8192      /* update records of a phase */

typedef struct {
    uint32_t *d;
    uint32_t deg;
    uint32_t cap;   /* capacity, | HEAP_BIT when the block is not in the slab */
} crow;

typedef struct {
    int app, wmode;
    uint32_t source, wmax;
    uint64_t seed;
    uint64_t *st;        /* value << 32 | parent */
    uint32_t *stamp;     /* round stamps (frontier dedup) */
    uint32_t round;
} cquery;

struct orc_cpu {
    uint32_t n;
    int threads;
    crow *out, *in;
    uint32_t *out_slab, *in_slab;
    uint64_t m;
    cquery *q;
    int nq;
    uint8_t *aff_flag;   /* invalidation claims (cleared after each query's stage) */
    uint32_t *aff;       /* affected list */
    uint32_t *front, *next;
    uint64_t *keys, *keys_tmp;   /* mutation: records by row */
    size_t keys_cap;
    /* PageRank (orc_cpu_pr_*): FP32 ranks, residuals and per-round shares */
    float *rank, *res, *share;
    float pr_eps;
    uint32_t pr_rounds;
};

static inline uint32_t cap_of(const crow *r) { return r->cap & ~HEAP_BIT; }

static uint32_t pow2_cap(uint32_t deg) {
    uint32_t c = deg < 4 ? 4 : deg;
    c--;
    c |= c >> 1; c |= c >> 2; c |= c >> 4; c |= c >> 8; c |= c >> 16;
    return c + 1;
}

static uint32_t qweight(const cquery *q, uint32_t u, uint32_t v) {
    if (q->app == ORC_APP_CC) return 0;
    if (q->app == ORC_APP_BFS) return 1;
    return orc_weight(q->wmode, q->seed, q->wmax, u, v);
}

static inline uint32_t val_of(uint64_t s) { return (uint32_t)(s >> 32); }
static inline uint32_t par_of(uint64_t s) { return (uint32_t)s; }
static inline uint64_t pack(uint32_t v, uint32_t p) { return ((uint64_t)v << 32) | p; }
static inline uint64_t load64(const uint64_t *p) { return __atomic_load_n(p, __ATOMIC_RELAXED); }

/* Lower *p to x; true if this call lowered it. */
static inline int atomic_min64(uint64_t *p, uint64_t x) {
    uint64_t cur = load64(p);
    while (x < cur) {
        if (__atomic_compare_exchange_n(p, &cur, x, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) return 1;
    }
    return 0;
}

/* ---- rows ------------------------------------------------------------------ */
static uint32_t *build_rows(uint32_t n, crow *rows, const uint32_t *key, const uint32_t *val, uint64_t m) {
    uint32_t *deg = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < m; ++i) __atomic_fetch_add(&deg[key[i]], 1u, __ATOMIC_RELAXED);
    uint64_t total = 0;
    for (uint32_t u = 0; u < n; ++u) total += deg[u] ? pow2_cap(deg[u]) : 0;
    uint32_t *slab = (uint32_t *)malloc((total ? total : 1) * sizeof(uint32_t));
    uint64_t off = 0;
    for (uint32_t u = 0; u < n; ++u) {
        uint32_t cap = deg[u] ? pow2_cap(deg[u]) : 0;
        rows[u].d = cap ? slab + off : NULL;
        rows[u].deg = 0;
        rows[u].cap = cap;
        off += cap;
    }
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < m; ++i) {
        crow *r = &rows[key[i]];
        uint32_t at = __atomic_fetch_add(&r->deg, 1u, __ATOMIC_RELAXED);
        r->d[at] = val[i];
    }
    free(deg);
    return slab;
}

/* Room for `more` entries: move the row to a power-of-two block that fits. */
static void row_reserve(crow *r, uint32_t more) {
    uint32_t need = r->deg + more, cap = cap_of(r);
    if (need <= cap) return;
    uint32_t nc = pow2_cap(need);
    uint32_t *d = (uint32_t *)malloc((size_t)nc * sizeof(uint32_t));
    if (r->deg) memcpy(d, r->d, (size_t)r->deg * sizeof(uint32_t));
    if (r->cap & HEAP_BIT) free(r->d);
    r->d = d;
    r->cap = nc | HEAP_BIT;
}

orc_cpu *orc_cpu_create(uint32_t n, const uint32_t *src, const uint32_t *dst, uint64_t m, int threads) {
    orc_cpu *c = (orc_cpu *)calloc(1, sizeof(orc_cpu));
    c->n = n;
    c->threads = threads > 0 ? threads : omp_get_max_threads();
    omp_set_num_threads(c->threads);
    size_t nn = n ? n : 1;
    c->out = (crow *)calloc(nn, sizeof(crow));
    c->in = (crow *)calloc(nn, sizeof(crow));
    c->out_slab = build_rows(n, c->out, src, dst, m);
    c->in_slab = build_rows(n, c->in, dst, src, m);
    c->m = m;
    c->aff_flag = (uint8_t *)calloc(nn, 1);
    c->aff = (uint32_t *)malloc(nn * sizeof(uint32_t));
    c->front = (uint32_t *)malloc(nn * sizeof(uint32_t));
    c->next = (uint32_t *)malloc(nn * sizeof(uint32_t));
    return c;
}

int orc_cpu_threads(const orc_cpu *c) { return c->threads; }
uint64_t orc_cpu_edges(const orc_cpu *c) { return c->m; }

int orc_cpu_add_query(orc_cpu *c, int app, uint32_t source, int wmode, uint64_t seed, uint32_t wmax) {
    cquery *q = (cquery *)realloc(c->q, (size_t)(c->nq + 1) * sizeof(cquery));
    if (!q) return -1;
    c->q = q;
    cquery *x = &c->q[c->nq];
    memset(x, 0, sizeof(*x));
    x->app = app; x->wmode = wmode; x->seed = seed; x->wmax = wmax ? wmax : 1;
    x->source = source < c->n ? source : (c->n ? c->n - 1 : 0);
    x->st = (uint64_t *)malloc((size_t)(c->n ? c->n : 1) * sizeof(uint64_t));
    x->stamp = (uint32_t *)calloc(c->n ? c->n : 1, sizeof(uint32_t));
    return c->nq++;
}

void orc_cpu_destroy(orc_cpu *c) {
    if (!c) return;
    for (uint32_t u = 0; u < c->n; ++u) {
        if (c->out[u].cap & HEAP_BIT) free(c->out[u].d);
        if (c->in[u].cap & HEAP_BIT) free(c->in[u].d);
    }
    for (int k = 0; k < c->nq; ++k) { free(c->q[k].st); free(c->q[k].stamp); }
    free(c->q); free(c->out_slab); free(c->in_slab); free(c->out); free(c->in);
    free(c->aff_flag); free(c->aff); free(c->front); free(c->next); free(c->keys); free(c->keys_tmp);
    free(c->rank); free(c->res); free(c->share);
    free(c);
}

static uint32_t init_value(const cquery *q, uint32_t v) {
    if (q->app == ORC_APP_CC) return v;
    return v == q->source ? 0u : ORC_INF;
}

void orc_cpu_set_state(orc_cpu *c, int k, const uint32_t *value, const uint32_t *parent) {
    cquery *q = &c->q[k];
#pragma omp parallel for schedule(static)
    for (uint32_t v = 0; v < c->n; ++v) q->st[v] = pack(value[v], parent ? parent[v] : ORC_INF);
}

void orc_cpu_values(const orc_cpu *c, int k, uint32_t *value, uint32_t *parent) {
    const cquery *q = &c->q[k];
    for (uint32_t v = 0; v < c->n; ++v) {
        if (value) value[v] = val_of(q->st[v]);
        if (parent) parent[v] = par_of(q->st[v]);
    }
}

void orc_cpu_csr(const orc_cpu *c, uint64_t *rp, uint32_t *col) {
    rp[0] = 0;
    for (uint32_t u = 0; u < c->n; ++u) rp[u + 1] = rp[u] + c->out[u].deg;
    if (col)
        for (uint32_t u = 0; u < c->n; ++u)
            if (c->out[u].deg) memcpy(col + rp[u], c->out[u].d, (size_t)c->out[u].deg * sizeof(uint32_t));
}

/* ---- thread-local append buffers ----------------------------------------- */
typedef struct {
    uint32_t *d;
    size_t n, cap;
} vbuf;

static inline void vb_push(vbuf *b, uint32_t x) {
    if (b->n == b->cap) {
        b->cap = b->cap ? 2 * b->cap : 1024;
        b->d = (uint32_t *)realloc(b->d, b->cap * sizeof(uint32_t));
    }
    b->d[b->n++] = x;
}

/* Append every thread's buffer to out[*count ...] (called inside a parallel
 * region by every thread, after a barrier). */
static void vb_flush(vbuf *b, uint32_t *out, size_t *count) {
    size_t at = __atomic_fetch_add(count, b->n, __ATOMIC_RELAXED);
    if (b->n) memcpy(out + at, b->d, b->n * sizeof(uint32_t));
    b->n = 0;
}

/* ---- push closure ------------------------------------------------------------ */
/* Level-synchronous relaxation from front[0..cnt) to the fixed point. */
static void push_closure(orc_cpu *c, cquery *q, size_t cnt) {
    uint32_t *front = c->front, *next = c->next;
    while (cnt) {
        const uint32_t round = ++q->round;
        size_t ncount = 0;
#pragma omp parallel if (cnt > PAR_FRONT)
        {
            vbuf b = {0, 0, 0};
#pragma omp for schedule(dynamic, 64)
            for (size_t i = 0; i < cnt; ++i) {
                const uint32_t s = front[i];
                const uint32_t vs = val_of(load64(&q->st[s]));
                if (vs == ORC_INF) continue;
                const crow *r = &c->out[s];
                for (uint32_t k = 0; k < r->deg; ++k) {
                    const uint32_t d = r->d[k];
                    const uint64_t cand = (uint64_t)vs + qweight(q, s, d);
                    if (cand >= ORC_INF) continue;
                    if (atomic_min64(&q->st[d], pack((uint32_t)cand, s)) &&
                        __atomic_exchange_n(&q->stamp[d], round, __ATOMIC_RELAXED) != round)
                        vb_push(&b, d);
                }
            }
            vb_flush(&b, next, &ncount);
            free(b.d);
        }
        uint32_t *t = front; front = next; next = t;
        cnt = ncount;
    }
    c->front = front;
    c->next = next;
}

void orc_cpu_converge(orc_cpu *c, int k) {
    cquery *q = &c->q[k];
    size_t cnt = 0;
#pragma omp parallel for schedule(static)
    for (uint32_t v = 0; v < c->n; ++v) q->st[v] = pack(init_value(q, v), ORC_INF);
    if (q->app == ORC_APP_CC) {
        for (uint32_t v = 0; v < c->n; ++v) c->front[cnt++] = v;
    } else if (c->n) {
        c->front[cnt++] = q->source;
    }
    push_closure(c, q, cnt);
}

/* ---- delete stage: invalidation ---------------------------------------------- */
/* Dependency predicate (api.cuh:43-49; CC: hybrid_cc.cu:60-64) on the current
 * (pre-reset) values. */
static int dependency(const cquery *q, uint32_t s, uint32_t d) {
    const uint64_t ss = load64(&q->st[s]), sd = load64(&q->st[d]);
    const uint32_t vs = val_of(ss), vd = val_of(sd);
    if (q->app == ORC_APP_CC) return s != d && vs == vd && vd < d;
    if (d == q->source) return 0;
    const uint32_t p = par_of(sd);
    const uint64_t cand = (uint64_t)vs + qweight(q, s, d);
    if (p == s) return 1;
    if (vs == ORC_INF || cand != (uint64_t)vd) return 0;
    int ptight = 0;
    if (p != ORC_INF) {
        const uint32_t vp = val_of(load64(&q->st[p]));
        ptight = vp != ORC_INF && (uint64_t)vp + qweight(q, p, d) == (uint64_t)vd;
    }
    return !ptight;
}

/* Seeds, closure over the current rows and reset; returns the affected count
 * (list in c->aff; claims left in c->aff_flag for the repair). */
static size_t invalidate(orc_cpu *c, cquery *q, const orc_edge *dels, uint64_t nd) {
    size_t na = 0;
    const uint32_t n = c->n;
#pragma omp parallel if (nd > PAR_RECS)
    {
        vbuf b = {0, 0, 0};
#pragma omp for schedule(static)
        for (uint64_t i = 0; i < nd; ++i) {
            const uint32_t u = dels[i].u, v = dels[i].v;
            if (u >= n || v >= n || c->aff_flag[v]) continue;
            if (dependency(q, u, v) && __atomic_exchange_n(&c->aff_flag[v], 1, __ATOMIC_RELAXED) == 0) vb_push(&b, v);
        }
        vb_flush(&b, c->aff, &na);
        free(b.d);
    }
    /* claimed vertices lose their parent before the closure (the seed loop of
     * the reference marks them first) */
#pragma omp parallel for schedule(static) if (na > PAR_RECS)
    for (size_t k = 0; k < na; ++k) q->st[c->aff[k]] = pack(val_of(q->st[c->aff[k]]), ORC_INF);
    size_t lo = 0;
    while (lo < na) {
        const size_t hi = na;
#pragma omp parallel if (hi - lo > PAR_FRONT)
        {
            vbuf b = {0, 0, 0};
#pragma omp for schedule(dynamic, 64)
            for (size_t k = lo; k < hi; ++k) {
                const uint32_t s = c->aff[k];
                const crow *r = &c->out[s];
                for (uint32_t j = 0; j < r->deg; ++j) {
                    const uint32_t d = r->d[j];
                    if (c->aff_flag[d]) continue;
                    if (dependency(q, s, d) && __atomic_exchange_n(&c->aff_flag[d], 1, __ATOMIC_RELAXED) == 0) {
                        __atomic_store_n(&q->st[d], pack(val_of(load64(&q->st[d])), ORC_INF), __ATOMIC_RELAXED);
                        vb_push(&b, d);
                    }
                }
            }
#pragma omp barrier
            vb_flush(&b, c->aff, &na);
            free(b.d);
        }
        lo = hi;
    }
#pragma omp parallel for schedule(static) if (na > PAR_RECS)
    for (size_t k = 0; k < na; ++k) q->st[c->aff[k]] = pack(init_value(q, c->aff[k]), ORC_INF);
    return na;
}

/* Pull-min over the post-delete in-rows of the affected set, then the push
 * closure from every affected vertex with a finite value. */
static void repair(orc_cpu *c, cquery *q, size_t na) {
#pragma omp parallel for schedule(dynamic, 64) if (na > PAR_FRONT)
    for (size_t k = 0; k < na; ++k) {
        const uint32_t v = c->aff[k];
        uint64_t best = load64(&q->st[v]);
        const crow *r = &c->in[v];
        for (uint32_t j = 0; j < r->deg; ++j) {
            const uint32_t x = r->d[j];
            const uint32_t vx = val_of(load64(&q->st[x]));
            if (vx == ORC_INF) continue;
            const uint64_t cand = (uint64_t)vx + qweight(q, x, v);
            if (cand >= ORC_INF) continue;
            const uint64_t p = pack((uint32_t)cand, x);
            if (p < best) best = p;
        }
        atomic_min64(&q->st[v], best);
    }
    size_t cnt = 0;
    const uint32_t round = ++q->round;
    for (size_t k = 0; k < na; ++k) {
        const uint32_t v = c->aff[k];
        if (val_of(q->st[v]) != ORC_INF && q->stamp[v] != round) {
            q->stamp[v] = round;
            c->front[cnt++] = v;
        }
    }
    push_closure(c, q, cnt);
}

/* ---- mutation ------------------------------------------------------------------ */
/* Stable sort of 64-bit keys by their high 32 bits: four 8-bit LSD passes on
 * the calling thread for small inputs, else two 16-bit passes on all threads. */
static void sort_by_row(uint64_t *keys, uint64_t *tmp, size_t n, int threads) {
    if (n <= PAR_RECS) {
        size_t h[256];
        for (int pass = 0; pass < 4; ++pass) {
            const int shift = 32 + 8 * pass;
            memset(h, 0, sizeof(h));
            for (size_t i = 0; i < n; ++i) h[(keys[i] >> shift) & 0xFF]++;
            size_t sum = 0;
            for (int b = 0; b < 256; ++b) {
                const size_t x = h[b];
                h[b] = sum;
                sum += x;
            }
            for (size_t i = 0; i < n; ++i) tmp[h[(keys[i] >> shift) & 0xFF]++] = keys[i];
            memcpy(keys, tmp, n * sizeof(uint64_t));
        }
        return;
    }
    const int B = 1 << 16;
    size_t *hist = (size_t *)malloc((size_t)threads * B * sizeof(size_t));
    for (int pass = 0; pass < 2; ++pass) {
        const int shift = 32 + 16 * pass;
        memset(hist, 0, (size_t)threads * B * sizeof(size_t));
#pragma omp parallel num_threads(threads)
        {
            const int t = omp_get_thread_num();
            const size_t lo = n * (size_t)t / (size_t)threads, hi = n * (size_t)(t + 1) / (size_t)threads;
            size_t *h = hist + (size_t)t * B;
            for (size_t i = lo; i < hi; ++i) h[(keys[i] >> shift) & 0xFFFF]++;
#pragma omp barrier
#pragma omp single
            {
                size_t sum = 0;
                for (int b = 0; b < B; ++b)
                    for (int tt = 0; tt < threads; ++tt) {
                        const size_t x = hist[(size_t)tt * B + b];
                        hist[(size_t)tt * B + b] = sum;
                        sum += x;
                    }
            }
            for (size_t i = lo; i < hi; ++i) tmp[h[(keys[i] >> shift) & 0xFFFF]++] = keys[i];
        }
        memcpy(keys, tmp, n * sizeof(uint64_t));
    }
    free(hist);
}

static void ensure_keys(orc_cpu *c, size_t n) {
    if (n <= c->keys_cap) return;
    free(c->keys); free(c->keys_tmp);
    c->keys = (uint64_t *)malloc(n * sizeof(uint64_t));
    c->keys_tmp = (uint64_t *)malloc(n * sizeof(uint64_t));
    c->keys_cap = n;
}

static int cmp_u32(const void *a, const void *b) {
    const uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? -1 : x > y;
}

/* Delete the first occurrence of every (row, target) record, grouped by row:
 * one order-preserving pass per touched row.  Returns the missing count. */
static uint64_t delete_rows(crow *rows, uint64_t *keys, size_t nk) {
    uint64_t missing = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : missing) if (nk > PAR_RECS)
    for (size_t i = 0; i < nk; ++i) {
        if (i && (keys[i] >> 32) == (keys[i - 1] >> 32)) continue;   /* not a group head */
        size_t j = i + 1;
        while (j < nk && (keys[j] >> 32) == (keys[i] >> 32)) ++j;
        crow *r = &rows[keys[i] >> 32];
        const size_t k = j - i;
        uint32_t small_t[16], small_c[16];
        uint32_t *tgt = k <= 16 ? small_t : (uint32_t *)malloc(k * sizeof(uint32_t));
        uint32_t *cnt = k <= 16 ? small_c : (uint32_t *)malloc(k * sizeof(uint32_t));
        for (size_t t = 0; t < k; ++t) tgt[t] = (uint32_t)keys[i + t];
        qsort(tgt, k, sizeof(uint32_t), cmp_u32);
        /* distinct targets with the number of first occurrences to drop */
        size_t u = 0;
        for (size_t t = 0; t < k; ++t) {
            if (u && tgt[u - 1] == tgt[t]) {
                cnt[u - 1]++;
            } else {
                tgt[u] = tgt[t];
                cnt[u++] = 1;
            }
        }
        uint32_t left = (uint32_t)k, w = 0;
        for (uint32_t x = 0; x < r->deg; ++x) {
            const uint32_t e = r->d[x];
            if (left) {
                size_t lo = 0, hi = u;   /* first distinct target >= e */
                while (lo < hi) {
                    const size_t mid = (lo + hi) / 2;
                    if (tgt[mid] < e) lo = mid + 1;
                    else hi = mid;
                }
                if (lo < u && tgt[lo] == e && cnt[lo]) {
                    cnt[lo]--;
                    left--;
                    continue;   /* dropped */
                }
            }
            r->d[w++] = e;
        }
        r->deg = w;
        missing += left;
        if (tgt != small_t) { free(tgt); free(cnt); }
    }
    return missing;
}

/* Append the records of each row in record order (the sort is stable). */
static void append_rows(crow *rows, uint64_t *keys, size_t nk) {
#pragma omp parallel for schedule(dynamic, 16) if (nk > PAR_RECS)
    for (size_t i = 0; i < nk; ++i) {
        if (i && (keys[i] >> 32) == (keys[i - 1] >> 32)) continue;
        size_t j = i + 1;
        while (j < nk && (keys[j] >> 32) == (keys[i] >> 32)) ++j;
        crow *r = &rows[keys[i] >> 32];
        row_reserve(r, (uint32_t)(j - i));
        for (size_t t = i; t < j; ++t) r->d[r->deg++] = (uint32_t)keys[t];
    }
}

/* Both directions of one phase's records.  Returns missing deletes (delete)
 * or the valid record count (add). */
static uint64_t mutate(orc_cpu *c, const orc_edge *recs, uint64_t nr, int del) {
    ensure_keys(c, nr ? nr : 1);
    uint64_t result = 0;
    for (int dir = 0; dir < 2; ++dir) {
        size_t nk = 0;
        for (uint64_t i = 0; i < nr; ++i) {
            const uint32_t u = recs[i].u, v = recs[i].v;
            if (u >= c->n || v >= c->n) continue;
            c->keys[nk++] = dir == 0 ? (((uint64_t)u << 32) | v) : (((uint64_t)v << 32) | u);
        }
        sort_by_row(c->keys, c->keys_tmp, nk, c->threads);
        if (del) {
            const uint64_t miss = delete_rows(dir == 0 ? c->out : c->in, c->keys, nk);
            if (dir == 0) result = miss + (nr - nk);
        } else {
            append_rows(dir == 0 ? c->out : c->in, c->keys, nk);
            if (dir == 0) result = nk;
        }
    }
    return result;
}

uint64_t orc_cpu_delete_phase(orc_cpu *c, const orc_edge *dels, uint64_t nd) {
    size_t *na = (size_t *)malloc((size_t)(c->nq ? c->nq : 1) * sizeof(size_t));
    uint32_t **lists = (uint32_t **)malloc((size_t)(c->nq ? c->nq : 1) * sizeof(uint32_t *));
    for (int k = 0; k < c->nq; ++k) {
        na[k] = invalidate(c, &c->q[k], dels, nd);
        lists[k] = (uint32_t *)malloc((na[k] ? na[k] : 1) * sizeof(uint32_t));
        memcpy(lists[k], c->aff, na[k] * sizeof(uint32_t));
#pragma omp parallel for schedule(static) if (na[k] > PAR_RECS)
        for (size_t j = 0; j < na[k]; ++j) c->aff_flag[lists[k][j]] = 0;
    }
    const uint64_t missing = mutate(c, dels, nd, 1);
    c->m -= nd - missing;
    for (int k = 0; k < c->nq; ++k) {
        memcpy(c->aff, lists[k], na[k] * sizeof(uint32_t));
        repair(c, &c->q[k], na[k]);
        free(lists[k]);
    }
    free(lists);
    free(na);
    return missing;
}

/* The mutations of one snapshot (deletes, then additions) without any repair:
 * brings the store to a later point of the stream whose values the caller then
 * installs with orc_cpu_set_state. */
void orc_cpu_mutate_only(orc_cpu *c, const orc_edge *dels, uint64_t nd, const orc_edge *adds, uint64_t na) {
    const uint64_t missing = mutate(c, dels, nd, 1);
    c->m -= nd - missing;
    c->m += mutate(c, adds, na, 0);
}

uint64_t orc_cpu_add_phase(orc_cpu *c, const orc_edge *adds, uint64_t na) {
    const uint64_t valid = mutate(c, adds, na, 0);
    c->m += valid;
    for (int k = 0; k < c->nq; ++k) {
        cquery *q = &c->q[k];
        const uint32_t round = ++q->round;
        size_t cnt = 0;
#pragma omp parallel if (na > PAR_RECS)
        {
            vbuf b = {0, 0, 0};
#pragma omp for schedule(static)
            for (uint64_t i = 0; i < na; ++i) {
                const uint32_t u = adds[i].u, v = adds[i].v;
                if (u >= c->n || v >= c->n) continue;
                const uint32_t vu = val_of(load64(&q->st[u]));
                if (vu == ORC_INF) continue;
                const uint64_t cand = (uint64_t)vu + qweight(q, u, v);
                if (cand >= ORC_INF) continue;
                if (atomic_min64(&q->st[v], pack((uint32_t)cand, u)) &&
                    __atomic_exchange_n(&q->stamp[v], round, __ATOMIC_RELAXED) != round)
                    vb_push(&b, v);
            }
            vb_flush(&b, c->front, &cnt);
            free(b.d);
        }
        push_closure(c, q, cnt);
    }
    return valid;
}

/* ------------------------------------------------------------------ PageRank
 * The engine's large-batch PageRank path (bench.py's CPU baseline for the
 * PageRank workload), on all cores: the signed-residual amend of
 * pr_residual.cuh around the batch's mutation -- Amend(-1) over every touched
 * source's old out-row, the mutation, Amend(+1) over its new out-row (each a
 * share 0.85 * rank / out-degree, added with an atomic per entry) -- then dense
 * Jacobi rounds: every residual is consumed into its rank (pr_residual.cuh:
 * 53-69) and each vertex pulls its new residual from its in-row, until the
 * largest |residual| is at most eps or max_rounds rounds ran.  FP32 ranks and
 * residuals, as on the GPU. */
void orc_cpu_pr_init(orc_cpu *c, double base, double eps, uint32_t max_rounds) {
    const size_t nn = c->n ? c->n : 1;
    free(c->rank); free(c->res); free(c->share);
    c->rank = (float *)calloc(nn, sizeof(float));
    c->res = (float *)malloc(nn * sizeof(float));
    c->share = (float *)calloc(nn, sizeof(float));
    for (uint32_t v = 0; v < c->n; ++v) c->res[v] = (float)base;   /* rank 0, residual base: the static start */
    c->pr_eps = (float)eps;
    c->pr_rounds = max_rounds;
}

void orc_cpu_pr_set_state(orc_cpu *c, const float *rank, const float *res) {
    memcpy(c->rank, rank, (size_t)c->n * sizeof(float));
    memcpy(c->res, res, (size_t)c->n * sizeof(float));
}

void orc_cpu_pr_values(const orc_cpu *c, float *rank, float *res) {
    if (rank) memcpy(rank, c->rank, (size_t)c->n * sizeof(float));
    if (res) memcpy(res, c->res, (size_t)c->n * sizeof(float));
}

static void pr_amend(orc_cpu *c, size_t nt, float sign) {
#pragma omp parallel for schedule(dynamic, 64) if (nt > 64)
    for (size_t i = 0; i < nt; ++i) {
        const uint32_t u = c->aff[i];
        const crow *r = &c->out[u];
        if (!r->deg || c->rank[u] == 0.f) continue;
        const float x = 0.85f * c->rank[u] * sign / (float)r->deg;
        for (uint32_t j = 0; j < r->deg; ++j) {
#pragma omp atomic
            c->res[r->d[j]] += x;
        }
    }
}

/* One batch; returns the rounds run (the static solve: a batch without records). */
uint32_t orc_cpu_pr_batch(orc_cpu *c, const orc_edge *dels, uint64_t nd, const orc_edge *adds, uint64_t na) {
    const uint32_t n = c->n;
    /* touched sources, each once */
    size_t nt = 0;
    for (int k = 0; k < 2; ++k) {
        const orc_edge *e = k ? adds : dels;
        const uint64_t ne = k ? na : nd;
        for (uint64_t i = 0; i < ne; ++i) {
            const uint32_t u = e[i].u;
            if (u >= n || e[i].v >= n || c->aff_flag[u]) continue;
            c->aff_flag[u] = 1;
            c->aff[nt++] = u;
        }
    }
    for (size_t i = 0; i < nt; ++i) c->aff_flag[c->aff[i]] = 0;
    pr_amend(c, nt, -1.f);
    const uint64_t missing = mutate(c, dels, nd, 1);
    c->m -= nd - missing;
    c->m += mutate(c, adds, na, 0);
    pr_amend(c, nt, 1.f);
    uint32_t r = 0;
    float rmax = 0.f;
    for (; r < c->pr_rounds; ++r) {
        if (r > 0 && !(rmax > c->pr_eps)) break;
#pragma omp parallel for schedule(static)
        for (uint32_t v = 0; v < n; ++v) {
            const float x = c->res[v];
            c->rank[v] += x;
            c->res[v] = 0.f;
            const uint32_t d = c->out[v].deg;
            c->share[v] = d ? 0.85f * x / (float)d : 0.f;
        }
        rmax = 0.f;
#pragma omp parallel for schedule(dynamic, 1024) reduction(max : rmax)
        for (uint32_t v = 0; v < n; ++v) {
            const crow *row = &c->in[v];
            float s = 0.f;
            for (uint32_t j = 0; j < row->deg; ++j) s += c->share[row->d[j]];
            c->res[v] = s;
            const float a = s < 0.f ? -s : s;
            if (a > rmax) rmax = a;
        }
    }
    return r;
}

