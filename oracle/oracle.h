/*
 * oracle.h -- CPU oracle for the incremental streaming-graph hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline).  The product library never links it.
 *
 * Every function restates the reference's semantics in plain C; each cites the
 * reference file:line it follows.  How each one is pinned (golden vectors,
 * closed forms, brute force) is listed in the header comment of oracle.c.
 */
#ifndef DCG_ORACLE_H
#define DCG_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_INF 0xFFFFFFFFu

typedef struct { uint32_t u, v; } orc_edge;

/* Edge-weight policies (weights are pure functions of (u, v), never stored). */
enum {
    ORC_W_REFERENCE = 0, /* (u+v)%128+1          include/framework/variants/api.cuh:26-28 */
    ORC_W_SEEDED = 1,    /* 1 + mix64(seed ^ (u<<32|v)) % wmax                            */
    ORC_W_UNIT = 2,      /* BFS hop count        samples/hybrid_bfs/hybrid_bfs.cu:56-62     */
    ORC_W_ZERO = 3       /* CC label forwarding  samples/hybrid_cc/hybrid_cc.cu:66-72       */
};
uint64_t orc_mix64(uint64_t x);
uint32_t orc_weight(int mode, uint64_t seed, uint32_t wmax, uint32_t u, uint32_t v);

/* ---- multiset topology model (TopologyReplayModel) ---- */
typedef struct orc_store orc_store;
typedef struct {
    uint64_t deletions, additions;
    uint64_t touched_sources, changed_sources;
    uint64_t missing_deletes, invalid_additions;
    uint64_t graph_edges;
} orc_batch_metrics;

orc_store *orc_store_create(uint32_t n);
void orc_store_destroy(orc_store *s);
/* Rows are filled in file order (src/utils/parser.cpp:379-400). */
int orc_store_load(orc_store *s, const uint32_t *src, const uint32_t *dst, uint64_t m);
void orc_store_apply(orc_store *s, const orc_edge *dels, uint64_t nd,
                     const orc_edge *adds, uint64_t na, orc_batch_metrics *out);
/* Single-phase variants (the delete stage and add stage of one snapshot). */
void orc_store_apply_deletes(orc_store *s, const orc_edge *dels, uint64_t nd, orc_batch_metrics *out);
void orc_store_apply_adds(orc_store *s, const orc_edge *adds, uint64_t na, orc_batch_metrics *out);
uint32_t orc_store_nodes(const orc_store *s);
uint64_t orc_store_edges(const orc_store *s);
uint32_t orc_store_degree(const orc_store *s, uint32_t u);
uint32_t orc_store_row(const orc_store *s, uint32_t u, uint32_t *out, uint32_t cap);
/* CSR export: row_ptr has n+1 entries, col has orc_store_edges() entries. */
void orc_store_csr(const orc_store *s, uint64_t *row_ptr, uint32_t *col);
uint64_t orc_hash_sequence(const uint32_t *xs, uint64_t k);

/* ---- from-scratch algorithms over a CSR ---- */
void orc_sssp(uint32_t n, const uint64_t *rp, const uint32_t *col, uint32_t source,
              int wmode, uint64_t seed, uint32_t wmax, uint32_t *dist);
void orc_bfs(uint32_t n, const uint64_t *rp, const uint32_t *col, uint32_t source, uint32_t *dist);
void orc_cc(uint32_t n, const uint64_t *rp, const uint32_t *col, uint32_t *label);
/* FP64 Jacobi: x = base + 0.85 P^T x, dangling mass dropped; stops when
 * sum|x_{k+1}-x_k| <= tol_sum.  Returns rounds, or -1 if not converged. */
int orc_pagerank(uint32_t n, const uint64_t *rp, const uint32_t *col, double base,
                 double tol_sum, int max_rounds, double *x);
uint64_t orc_value_checksum(const uint32_t *values, uint32_t n);

/* ---- the reference's own checkers ---- */
typedef struct {
    uint64_t relaxable_edges, missing_tight_witnesses;
    uint64_t invalid_parents, missing_parent_edges, wrong_parent_distances;
    int source_ok;
} orc_sssp_check;
void orc_check_sssp(uint32_t n, const uint64_t *rp, const uint32_t *col, uint32_t source,
                    int wmode, uint64_t seed, uint32_t wmax, const uint32_t *dist,
                    const uint32_t *parent, orc_sssp_check *out);
typedef struct {
    double l1, bound, max_error, defect, max_residual, roundoff;
    int converged, finite, ok;
} orc_pr_check;
void orc_check_pagerank(uint32_t n, const uint64_t *rp, const uint32_t *col, double base,
                        double eps, double max_abs, const float *rank, const float *residual,
                        orc_pr_check *out);

/* ---- hotness ranking and cache admission (driver.cuh:486-579,
 * framework.cuh:1321-1350); window: four 8-bit counters per vertex ---- */
uint64_t orc_hot_candidates(uint32_t n, const uint32_t *window, const uint32_t *value, const uint32_t *degree,
                            uint64_t cap, uint32_t *ids);

/* ---- CPU incremental path (port of the reference's CPU-capable path) ---- */
enum { ORC_APP_SSSP = 0, ORC_APP_BFS = 1, ORC_APP_CC = 2, ORC_APP_PR = 3 };
typedef struct orc_incr orc_incr;
orc_incr *orc_incr_create(uint32_t n, const uint64_t *rp, const uint32_t *col, int app,
                          uint32_t source, int wmode, uint64_t seed, uint32_t wmax,
                          double pr_base, double pr_eps, int pr_max_rounds);
void orc_incr_destroy(orc_incr *e);
/* Seed the vertex state (e.g. with a from-scratch result) instead of recomputing. */
void orc_incr_set_state(orc_incr *e, const uint32_t *value, const uint32_t *parent);
void orc_incr_converge(orc_incr *e);
void orc_incr_pr_start(orc_incr *e);
void orc_incr_delete_phase(orc_incr *e, const orc_edge *dels, uint64_t nd);
void orc_incr_add_phase(orc_incr *e, const orc_edge *adds, uint64_t na);
void orc_incr_pr_batch(orc_incr *e, const orc_edge *dels, uint64_t nd,
                       const orc_edge *adds, uint64_t na);
void orc_incr_values(const orc_incr *e, uint32_t *value, uint32_t *parent);
void orc_incr_pr_values(const orc_incr *e, float *rank, float *residual);
uint64_t orc_incr_edges(const orc_incr *e);
void orc_incr_csr(const orc_incr *e, uint64_t *rp, uint32_t *col);

/* ---- the same path on all host cores (cpu_path.c): several SSSP/BFS/CC
 * queries over one store, the timed CPU baseline.  Rows are built from an edge
 * list in any order; `threads` <= 0 uses every core OpenMP sees. ---- */
typedef struct orc_cpu orc_cpu;
orc_cpu *orc_cpu_create(uint32_t n, const uint32_t *src, const uint32_t *dst, uint64_t m, int threads);
void orc_cpu_destroy(orc_cpu *c);
int orc_cpu_threads(const orc_cpu *c);
uint64_t orc_cpu_edges(const orc_cpu *c);
/* Returns the query index. */
int orc_cpu_add_query(orc_cpu *c, int app, uint32_t source, int wmode, uint64_t seed, uint32_t wmax);
void orc_cpu_converge(orc_cpu *c, int query);
void orc_cpu_set_state(orc_cpu *c, int query, const uint32_t *value, const uint32_t *parent);
/* Every query; return the missing deletes / the valid additions. */
uint64_t orc_cpu_delete_phase(orc_cpu *c, const orc_edge *dels, uint64_t nd);
uint64_t orc_cpu_add_phase(orc_cpu *c, const orc_edge *adds, uint64_t na);
void orc_cpu_mutate_only(orc_cpu *c, const orc_edge *dels, uint64_t nd, const orc_edge *adds, uint64_t na);
void orc_cpu_pr_init(orc_cpu *c, double base, double eps, uint32_t max_rounds);
void orc_cpu_pr_set_state(orc_cpu *c, const float *rank, const float *res);
void orc_cpu_pr_values(const orc_cpu *c, float *rank, float *res);
uint32_t orc_cpu_pr_batch(orc_cpu *c, const orc_edge *dels, uint64_t nd, const orc_edge *adds, uint64_t na);
void orc_cpu_values(const orc_cpu *c, int query, uint32_t *value, uint32_t *parent);
void orc_cpu_csr(const orc_cpu *c, uint64_t *rp, uint32_t *col);

/* ---- MLP 8->64->64->2 reference (FP64) for gradient checks ---- */
/* Weights are given as float (the device keeps FP16 copies); the forward pass,
 * loss and gradients are evaluated in double.  Loss = mean over rows of
 * BCE(sigmoid(o0), y0) + 0.5*(o1 - y1)^2.  Gradients are d(loss)/d(param). */
void orc_mlp_forward_backward(int rows, const float *x, const float *y,
                              const float *w1, const float *b1, const float *w2,
                              const float *b2, const float *w3, const float *b3,
                              double *out, double *loss,
                              double *gw1, double *gb1, double *gw2, double *gb2,
                              double *gw3, double *gb3);

#ifdef __cplusplus
}
#endif
#endif
