/*
 * dcg.h -- C ABI of the MI355X incremental streaming-graph engine.
 *
 * Plain C types only (pointers, sizes, integers).  Every entry point returns 0 on
 * success or a negative dcg_status; dcg_last_error() gives the message.  An
 * engine handle is single-threaded and not re-entrant; distinct handles are
 * independent.  The engine owns all device memory, the pinned host chunk store
 * and its mirrors; input arrays are borrowed for the duration of a call.
 *
 * Each entry point cites the reference interface it replaces
 * (sepgraph::engine::Engine in include/framework/framework.cuh and the sample
 * drivers under samples/hybrid_*).
 */
#ifndef DCG_H
#define DCG_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DCG_INF 0xFFFFFFFFu
#define DCG_ABI_VERSION 1

typedef struct dcg_engine dcg_engine;

/* One update record: an edge u->v (weights are pure functions of (u,v)).
 * Replaces WEdge{u,v,w} of include/framework/Loader.h:21-22 (the w column of
 * the update file is ignored there too, framework.cuh:2611-2616). */
typedef struct { uint32_t u, v; } dcg_edge;

typedef enum {
    DCG_OK = 0,
    DCG_ERR_ARG = -1,       /* bad argument (null pointer, out-of-range id, bad size) */
    DCG_ERR_HIP = -2,       /* a HIP runtime call failed                               */
    DCG_ERR_STATE = -3,     /* call out of order (e.g. add before load)               */
    DCG_ERR_CAPACITY = -4,  /* chunk store or frontier capacity exhausted             */
    DCG_ERR_NUMERIC = -5,   /* PageRank produced a non-finite rank/residual           */
    DCG_ERR_UNSUPPORTED = -6,
    DCG_ERR_IO = -7         /* a trace file could not be written                       */
} dcg_status;

typedef enum { DCG_APP_SSSP = 0, DCG_APP_BFS = 1, DCG_APP_CC = 2, DCG_APP_PR = 3 } dcg_app;

/* Edge weights used by SSSP (BFS uses 1, CC uses 0).  REFERENCE is
 * (u+v)%128+1 (include/framework/variants/api.cuh:26-28). SEEDED is
 * 1 + splitmix64(seed ^ (u<<32|v)) % weight_max. */
typedef enum { DCG_W_REFERENCE = 0, DCG_W_SEEDED = 1 } dcg_weight_mode;

/* Where the chunk store lives.  DEVICE keeps every chunk in HBM; HOST keeps the
 * authoritative chunks in pinned host memory read zero-copy, with a device
 * chunk cache capped at cache_bytes (the reference's -cache flag,
 * include/framework/graph_datum.cuh:164-186, graph-wide instead of per-row). */
typedef enum { DCG_PLACE_DEVICE = 0, DCG_PLACE_HOST = 1 } dcg_placement;

/* Chunk-cache admission policy.  HOTNESS is the reference's 4-snapshot
 * activity window with an LRU tie-break (driver.cuh:486-579,
 * framework.cuh:1264-1567); LEARNED is the online-trained MLP with HOTNESS as
 * fallback; NONE never admits (all-zero-copy). */
typedef enum { DCG_CACHE_NONE = 0, DCG_CACHE_HOTNESS = 1, DCG_CACHE_LEARNED = 2 } dcg_cache_policy;

typedef struct {
    uint32_t app;             /* dcg_app                                            */
    uint32_t num_queries;     /* concurrent sources (SSSP/BFS); CC and PR use 1     */
    const uint32_t *sources;  /* num_queries source ids (clamped to [0,n-1] at init) */
    uint32_t weight_mode;     /* dcg_weight_mode                                    */
    uint32_t weight_max;      /* W for DCG_W_SEEDED                                 */
    uint64_t weight_seed;
    uint32_t placement;       /* dcg_placement                                      */
    uint32_t cache_policy;    /* dcg_cache_policy                                   */
    uint64_t cache_bytes;     /* device chunk-cache cap for DCG_PLACE_HOST          */
    double store_slack;       /* spare slot fraction for relocated rows (>= 0.05)   */
    double pr_base;           /* PR base term per vertex (reference: 0.15)          */
    double pr_epsilon;        /* PR residual threshold (reference --error)          */
    uint32_t pr_max_rounds;   /* 1..100 (reference --pr_max_rounds)                 */
    uint32_t block_size;      /* threads per workgroup for traversal kernels (256)  */
    int32_t device;           /* HIP device ordinal                                 */
    /* Optional per-query program (DCG_APP_SSSP/BFS/CC, num_queries entries): one
     * engine then repairs SSSP, BFS and CC queries over one chunk store in the
     * same launches.  NULL: every query runs `app`.  Ignored for PR. */
    const uint32_t *query_apps;
    uint32_t reserved[4];
} dcg_config;

/* Per-phase counters (the reference's [P0-DELETE-ATTR]/[C3-*] log fields). */
typedef struct {
    uint64_t requests;            /* update records in this phase                      */
    uint64_t touched_sources;     /* distinct valid sources touched                    */
    uint64_t changed_sources;     /* sources whose row changed                         */
    uint64_t missing_deletes;     /* deletions with no remaining occurrence            */
    uint64_t invalid_additions;   /* additions with an endpoint id >= n                */
    uint64_t relocated_rows;      /* rows moved to a larger power-of-two block         */
    uint64_t affected;            /* vertices invalidated by the delete phase          */
    uint64_t rounds;              /* closure rounds (device-side)                      */
    uint64_t edges_scanned;       /* adjacency entries read by repair kernels          */
    uint64_t cached_edge_reads;   /* of which served from the device chunk cache       */
    uint64_t host_edge_reads;     /* of which served zero-copy from host memory        */
    double   ms;                  /* wall time of the phase                            */
    uint64_t bucket_advances;     /* bucket advances of the ordered push closure       */
    uint64_t far_entries;         /* far-pile entries those advances examined          */
    uint64_t coexec_rounds;       /* GPU closures of a co-executed add stage           */
    uint64_t cpu_relaxations;     /* edges the host threads relaxed                    */
    uint64_t proposals_to_cpu;    /* GPU -> CPU proposals (minimum per vertex)         */
    uint64_t proposals_to_gpu;    /* CPU -> GPU proposals (minimum per vertex)         */
} dcg_phase_stats;

/* Cache maintenance counters.  With the learned policy, learned_decisions and
 * fallback_decisions count the admission decisions taken so far from the
 * network's ranking and from the hotness ranking, and mlp_capture /
 * hot_capture are the judge's last comparison: the share of the last
 * snapshot's edge reads each ranking's top `frames` chunks would have held.
 * Stores reach both a resident chunk's frame and its host copy (write-through),
 * so evictions copy nothing back: written_back_chunks and d2h_bytes stay 0. */
typedef struct {
    uint64_t frames, resident_chunks, total_chunks;
    uint64_t loaded_chunks, evicted_chunks, written_back_chunks;
    uint64_t h2d_bytes, d2h_bytes;
    double   ms;
    uint64_t learned_decisions, fallback_decisions;
    double   mlp_capture, hot_capture;
} dcg_cache_stats;

/* A set of vertex programs: the engine's repair kernels instantiated for up to
 * 8 min-lattice programs (include/dcg_program.hpp) as device-function handles
 * that the engine launches.  The built-in set (SSSP, BFS, CC: indices 0, 1, 2 =
 * dcg_app) is compiled into the library; a plugin built from
 * include/dcg_program_plugin.hpp exports a table for its own programs.  Replaces
 * the compile-time TApp parameter of sepgraph::engine::Engine
 * (framework.cuh:780-783): the reference instantiates the engine for one app
 * struct; here the table selects the instantiation at run time. */
#define DCG_PROGRAM_ABI 1
#define DCG_PROGRAM_MAX 8
enum { DCG_PROGRAM_ROOTED = 1, DCG_PROGRAM_VERTEX_SEEDS = 2 };
typedef enum {
    DCG_PK_PUSH = 0,        /* persistent closures: (TravArgs) */
    DCG_PK_PULL = 1,
    DCG_PK_INVALIDATE = 2,
    DCG_PK_STATIC = 3,      /* static convergence (push) */
    DCG_PK_SEED_DELS = 4,   /* (TravArgs) */
    DCG_PK_SEED_ADDS = 5,   /* (TravArgs, uint32_t *frontier) */
    DCG_PK_RESET = 6,       /* (TravArgs) */
    DCG_PK_FIX_PARENTS = 7, /* (TravArgs, const uint32_t *list, const void *list_ctl, uint32_t base) */
    DCG_PK_INIT = 8,        /* (TravArgs) */
    DCG_PK_FRONTIER = 9,    /* (TravArgs, uint32_t *frontier, const uint64_t *slot0) */
    DCG_PK_NUM = 10
} dcg_program_kernel;
typedef struct {
    uint32_t abi;                          /* DCG_PROGRAM_ABI */
    uint32_t num_programs;                 /* 1..DCG_PROGRAM_MAX */
    uint32_t flags[DCG_PROGRAM_MAX];       /* DCG_PROGRAM_ROOTED | DCG_PROGRAM_VERTEX_SEEDS per program */
    const char *name;
    const void *kernels[DCG_PK_NUM];       /* host handles of the device kernels */
    /* bucket width of program p's ordered closure for the width setting
     * `quarters` and the weight cap; 0xFFFFFFFF: never ordered */
    uint32_t (*bucket_width)(uint32_t program, uint32_t quarters, uint32_t weight_max);
    uint32_t tag;                          /* checked against the library's TravArgs layout */
} dcg_program;

/* Engine construction.  Replaces Engine::Engine + Engine::SetOptions
 * (framework.cuh:1158-1176, 1245-1262). */
int dcg_config_init(dcg_config *cfg);
int dcg_create(const dcg_config *cfg, dcg_engine **out);
void dcg_destroy(dcg_engine *eng);
const char *dcg_last_error(const dcg_engine *eng);
const char *dcg_status_string(int status);

/* Run the engine's queries with a program set (call after dcg_create, before a
 * graph load; NULL restores the built-in set).  dcg_config.query_apps then hold
 * program indices below prog->num_programs (NULL: every query runs program 0);
 * the engine's app must not be PR.  The table must stay valid while the engine
 * lives.  DCG_ERR_ARG for a table of another ABI or TravArgs layout. */
int dcg_set_program(dcg_engine *eng, const dcg_program *prog);

/* Graph load.  Replaces Engine::LoadGraph (framework.cuh:1569-1798): rows keep
 * file order (src/utils/parser.cpp:379-400), n = num_nodes.  src/dst may be
 * host pointers.  dcg_load_csr takes a host CSR (row_ptr has n+1 entries). */
int dcg_load_edges(dcg_engine *eng, uint32_t num_nodes, const uint32_t *src,
                   const uint32_t *dst, uint64_t num_edges);
int dcg_load_csr(dcg_engine *eng, uint32_t num_nodes, const uint64_t *row_ptr,
                 const uint32_t *col);
/* Same, but src/dst already in device memory (sorted or not). */
int dcg_load_edges_device(dcg_engine *eng, uint32_t num_nodes, const uint32_t *d_src,
                          const uint32_t *d_dst, uint64_t num_edges);

/* Static convergence.  Replaces Engine::InitGraph + Engine::Start
 * (framework.cuh:1803-1853, 1963-2005) and Engine::StartPageRank
 * (framework.cuh:2681-2690). */
int dcg_start(dcg_engine *eng);

/* Update stream.  Replaces Loader + Engine::get_update_file
 * (include/framework/Loader.h:30-101, framework.cuh:2387-2418): the whole
 * stream is copied into engine-owned pinned host memory; each stage then
 * copies its slice [begin, begin+count) to the device inside the stage, as the
 * reference's del_edge/add_edge fill their mapped WEdge buffers per batch
 * (framework.cuh:2595-2677).  A second load replaces the stream. */
int dcg_load_updates(dcg_engine *eng, const dcg_edge *adds, uint64_t num_adds,
                     const dcg_edge *dels, uint64_t num_dels);
/* Copy the slices of a coming snapshot to the device on the engine's copy
 * stream, overlapping the snapshot that runs now (double buffering: two device
 * slots per direction).  The stage that asks for exactly that slice waits for
 * the copy instead of copying.  No reference counterpart. */
int dcg_prefetch_updates(dcg_engine *eng, uint64_t add_begin, uint64_t add_count,
                         uint64_t del_begin, uint64_t del_count);

/* Delete stage of one snapshot: mutation, invalidation, repair.  Replaces
 * Engine::del_edge / update_tree_del (framework.cuh:2595-2618, 4811-4957).
 * [begin, begin+count) indexes the staged deletions; a batch holds fewer than
 * 2^30 records (DCG_ERR_ARG otherwise; the same bound applies to additions). */
int dcg_del_edge(dcg_engine *eng, uint64_t begin, uint64_t count, dcg_phase_stats *stats);
/* Add stage: mutation, seeding, closure.  Replaces Engine::add_edge /
 * update_tree_add (framework.cuh:3525-3756). */
int dcg_add_edge(dcg_engine *eng, uint64_t begin, uint64_t count, dcg_phase_stats *stats);
/* One snapshot step: the delete stage of [del_begin, del_begin+del_count) and
 * the add stage of [add_begin, add_begin+add_count), enqueued back to back with
 * one host wait at the end.  Same results and statistics as dcg_del_edge then
 * dcg_add_edge (the pair the reference's sample loop calls per snapshot,
 * samples/hybrid_sssp/hybrid_sssp.cu:446 and :494); an error of either stage
 * is reported after both ran.  Either stats pointer may be null. */
int dcg_snapshot(dcg_engine *eng, uint64_t del_begin, uint64_t del_count, uint64_t add_begin, uint64_t add_count,
                 dcg_phase_stats *del_stats, dcg_phase_stats *add_stats);
/* PageRank batch (delete+add mutation between the two amend passes, then the
 * residual solve).  Replaces Engine::UpdatePageRank (framework.cuh:2692-2736). */
int dcg_update_pagerank(dcg_engine *eng, uint64_t add_begin, uint64_t add_count,
                        uint64_t del_begin, uint64_t del_count, dcg_phase_stats *stats);

/* Cache maintenance.  Replaces compute_hot_vertices_* + confirm_candidate_batch
 * + CacheRefreshRequired + evication_cache + compact_cache + LoadCache
 * (hybrid_sssp.cu:494-521, framework.cuh:1264-1567, 2978-3061). */
int dcg_cache_maintain(dcg_engine *eng, dcg_cache_stats *stats);
/* Prefetch for the staged batch about to be applied (learned policy only; a
 * no-op for HOTNESS/NONE).  No reference counterpart: the reference admits
 * only after a batch. */
int dcg_cache_prefetch(dcg_engine *eng, uint64_t add_begin, uint64_t add_count,
                       uint64_t del_begin, uint64_t del_count, dcg_cache_stats *stats);
int dcg_cache_policy_set(dcg_engine *eng, uint32_t policy);

/* The reference's vertex hotness ranking and cache admission
 * (compute_hot_vertices_sssp + CacheCandidateCount, driver.cuh:486-579,
 * graph_datum.cuh:291-318, framework.cuh:1321-1350), computed on the GPU from
 * the engine's per-vertex window of out-row expansions in the last four
 * snapshots (four saturating 8-bit counters per vertex; the current snapshot's
 * in the low byte): score = the window's sum, 0 for vertices `query` has not
 * reached (PageRank: no mask); vertices ranked by descending score, ties by
 * ascending id; the candidates are the longest ranked prefix whose out-degree
 * sum stays below cap_slots (all when the total fits, none when cap_slots is
 * 0).  Writes the first min(count, ids_len) ranked ids, sets *count, and then
 * shifts the window one snapshot (reset_hotness).  The engine's own chunk
 * cache ranks chunks, not vertices; this is the reference's vertex-granular
 * ranking for callers that cache rows. */
int dcg_hot_candidates(dcg_engine *eng, uint32_t query, uint64_t cap_slots, uint32_t *ids, uint64_t ids_len,
                       uint64_t *count);
/* The window words (n entries; byte k = expansions k snapshots ago, saturated at 255). */
int dcg_hot_window(dcg_engine *eng, uint32_t *out, size_t n);

/* Diagnostic traces, in the reference's file formats.
 *
 * Repair snapshot (I16, include/framework/i16_repair_snapshot.h:9-52, written
 * by framework.cuh:4189-4212 and 4668-4677): the first delete stage after
 * this call writes `path` (through path.tmp, renamed when complete): magic
 * 0x4931365354415445, version 1, infinity 0xFFFFFFFF, weight rule 1, batch 0,
 * the query's source, n; then the affected vertices of `query` after the
 * invalidation (u64 count + u32 ids, ascending), their live in-rows as u64
 * offsets (count + 1) and u32 sources, and the states {vertex, value,
 * buffer, parent} (u64 count + 16 B each) of the sorted unique union of the
 * sources, the affected vertices and the query's source, once before the
 * repair (affected values reset) and once after it.  buffer = value (this
 * engine keeps one state word per vertex).  Min-lattice programs only; the
 * stage runs its repair without the captured graph.  An empty path disarms. */
int dcg_set_repair_snapshot(dcg_engine *eng, const char *path, uint32_t query);

/* Cache-candidate trace (F1, TraceCacheCandidates, framework.cuh:1414-1507;
 * reader cache_patch_trace.h:15-60): dcg_set_cache_trace truncates `path` and
 * writes the magic 0x4631434143484531 (an empty path closes the trace);
 * dcg_trace_cache_candidates appends one record for the last
 * dcg_hot_candidates ranking: u32 batch, u64 capacity (its cap_slots), u64
 * count, then count x {u32 vertex, u32 live out-degree} for the admitted
 * prefix in rank order; count is 0 when the admitted list equals the previous
 * record's. */
typedef struct {
    uint64_t admitted_vertices, admitted_edges, changed_records, capacity_edges, bytes;
} dcg_trace_stats;
int dcg_set_cache_trace(dcg_engine *eng, const char *path);
int dcg_trace_cache_candidates(dcg_engine *eng, uint32_t batch, dcg_trace_stats *stats);
/* Stable GPU radix sort of caller keys by bits [begin_bit, end_bit), values
 * (optional) carried along; host arrays sorted in place.  The sort the engine
 * uses for its rankings, exposed for checks.  No reference counterpart. */
int dcg_radix_sort(int32_t device, uint64_t *keys, uint32_t *vals, uint64_t n, int begin_bit, int end_bit);

/* Result gather.  Replaces Engine::GatherValue/GatherParent/GatherBuffer
 * (framework.cuh:4959-4984, graph_datum.cuh:346-398).  values: u32 (SSSP/BFS
 * distances, CC labels) or f32 PR ranks; buffers: PR residuals; parents: u32.
 * n must equal the vertex count. */
int dcg_gather_values(dcg_engine *eng, uint32_t query, void *out, size_t n);
int dcg_gather_parents(dcg_engine *eng, uint32_t query, uint32_t *out, size_t n);
int dcg_gather_buffers(dcg_engine *eng, uint32_t query, void *out, size_t n);

/* Topology export (for the in-binary checkers; replaces ChunkStore()
 * Descriptor/SlabData reads, hybrid_sssp.cu:228-333).  Pass row_ptr == NULL
 * to query the edge count. */
int dcg_get_csr(dcg_engine *eng, uint64_t *row_ptr, uint32_t *col, uint64_t *num_edges);
int dcg_num_nodes(const dcg_engine *eng, uint32_t *n);

/* Counters of the whole run so far.  device_bytes = everything the engine
 * holds in device memory: store_device_bytes (cache frames, or the whole slot
 * space with device placement; row descriptors, sorted-prefix lengths, fences,
 * free pool, read counters, hotness window, cache-policy arrays) +
 * state_device_bytes (per-query state, parents, stamps, frontier queues, far
 * piles, PageRank arrays) + batch scratch (grouping keys, sort scratch, staged
 * records, the learned policy's buffers). */
typedef struct {
    uint64_t snapshots, store_slots, store_used_slots, chunks, frames;
    uint64_t host_bytes, device_bytes;
    uint64_t cached_edge_reads, host_edge_reads;
    uint64_t free_blocks, reused_blocks;   /* relocation blocks in the free pool / reused from it */
    uint64_t store_device_bytes, state_device_bytes, frame_bytes;
} dcg_engine_info;
int dcg_get_info(dcg_engine *eng, dcg_engine_info *info);

/* Per-kernel measurement (the role of the reference's communication meter,
 * include/utils/communication_meter.h, for device work): launches, summed
 * launch time from HIP events on the engine stream, and algorithmic bytes
 * (24 B per expanded row: 16 B row descriptor + 8 B owner state; 12 B per
 * scanned adjacency entry: 4 B id + 8 B destination state; mutation: 4 B per
 * slot read or written + 16 B per touched row descriptor). */
typedef enum {
    DCG_K_PUSH = 0, DCG_K_PULL = 1, DCG_K_INVALIDATE = 2, DCG_K_PR_SOLVE = 3, DCG_K_PR_AMEND = 4,
    DCG_K_DELETE_ROWS = 5,   /* delete mutation, wave path (rows of at most 1,024 slots) */
    DCG_K_APPEND_ROWS = 6,   /* append / relocation, both paths */
    DCG_K_DELETE_BIG = 7,    /* delete mutation, workgroup path (larger rows) */
    DCG_K_DELETE_PASS = 8,   /* both directions' delete launches together (the in-row pass runs
                                beside the out-row pass): wall time from the first launch to the
                                end of the last, with the bytes and host lines of all four */
    DCG_K_REPAIR_DELETE = 9, /* the delete stage's repair launches together (seeds, invalidation,
                                reset, pull, push closure, CC parent check): one record per stage,
                                priced by its closures' rows and adjacency entries */
    DCG_K_REPAIR_ADD = 10,   /* the add stage's repair launches together (seeds, push closure) */
    DCG_K_INGEST = 11,       /* host-to-device copies of update records (8 B per record) */
    DCG_K_STATIC = 12,       /* static convergence closures (dcg_start, the autotuner) */
    DCG_K_NUM = 13
} dcg_kernel_id;
typedef struct {
    uint64_t launches;
    double   ms;
    double   algo_bytes;
    uint64_t edges, expansions;
    uint64_t host_lines;   /* delete wave path: 64-byte lines read from pinned host memory */
} dcg_kernel_stats;
int dcg_get_kernel_stats(dcg_engine *eng, uint32_t kernel, dcg_kernel_stats *out);
int dcg_reset_kernel_stats(dcg_engine *eng);

/* The learned policy's predictor (MLP 8->64->64->2, FP16 weights, FP32
 * accumulation) evaluated on caller rows: x is rows*8, y (optional) rows*2
 * targets, params the 4,868 FP32 parameters (W1 8x64, b1, W2 64x64, b2, W3 64x2,
 * b3; row-major, input index first), out rows*2 outputs, grad (with y) the
 * gradient of mean(BCE(sigmoid(o0), y0) + 0.5*(o1-y1)^2).  No reference
 * counterpart: exposed so the device passes can be checked against an FP64
 * reference. */
int dcg_mlp_eval(int32_t device, uint32_t rows, const float *x, const float *y, const float *params,
                 float *out, float *grad);
/* The same evaluation on caller DEVICE buffers (x rows*8, y rows*2 or NULL,
 * params 4,868, out rows*2, grad 4,868 or NULL; grad needs y), allocating
 * nothing: every intermediate (the per-workgroup gradient rows, the FP16
 * parameters) lives in a static workspace of the library, 19.5 MiB on each GPU
 * it runs on, held until the process ends.  rows <= 2^24 (DCG_ERR_CAPACITY
 * beyond).  Runs on the null stream of `device` and returns when done; calls
 * that overlap on one GPU share the workspace and race.  Results equal
 * dcg_mlp_eval's bit for bit.  No reference counterpart: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops). */
int dcg_mlp_eval_device(int32_t device, uint32_t rows, const float *x, const float *y,
                        const float *params, float *out, float *grad);

/* Autotuner (no reference counterpart; the reference fixes block 512 and its
 * CTA/warp thresholds, include/utils/graphs/traversal.h:80-90, and its ordered
 * repair's bucket width 128, include/framework/ordered_gpu_repair.cuh:116-160):
 * times the persistent closure on the loaded graph for candidate workgroups per
 * CU, large-row thresholds, piece sizes and bucket widths (coordinate descent)
 * and keeps the fastest.  Call after dcg_start; leaves the converged state
 * unchanged. */
typedef struct {
    uint32_t blocks_per_cu, big_row, piece, evaluations;
    double best_ms;
    uint32_t bucket_quarters;
    uint32_t solo_max;        /* frontiers this small run on one workgroup */
    double first_ms;          /* the parameters in use before tuning */
    uint32_t grid_shift;      /* closure grids span (workgroups per CU * CUs) >> grid_shift */
} dcg_tune_result;
int dcg_autotune(dcg_engine *eng, dcg_tune_result *out);
/* Autotune on the incremental workload itself: each of the next `stages`
 * repair stages (dcg_del_edge / dcg_add_edge; not PageRank) runs its repair
 * once per candidate, every run from the same post-mutation state, and keeps
 * the fastest parameters (coordinate descent over bucket width, workgroups per
 * CU, grid size, solo-frontier threshold, large-row threshold and piece size; each run
 * timed with HIP events).  Values are those of any single run.  The stage's
 * timing records include the extra runs, so tune outside a measured window.
 * Each stage keeps its own choice of grid, solo threshold and row split (the
 * bucket width is shared).  dcg_autotune_last reports the last tuned stage,
 * dcg_autotune_stage_last the last tuning of stage 0 (delete) or 1 (add);
 * evaluations 0 if none (also after dcg_set_tuning / dcg_set_closure_grid /
 * dcg_autotune, which set the parameters of both stages).  No reference
 * counterpart: the reference's launch parameters are fixed
 * (traversal.h:80-90). */
int dcg_autotune_stages(dcg_engine *eng, uint32_t stages);
int dcg_autotune_last(dcg_engine *eng, dcg_tune_result *out);
int dcg_autotune_stage_last(dcg_engine *eng, uint32_t stage, dcg_tune_result *out);
/* CPU/GPU co-execution of the insertion closure (the reference's owner runtime:
 * --sssp_cpu_partition_capacity / --sssp_cpu_domain_map, framework.cuh:923-1155,
 * 5198-5517, dual_domain_event_runtime.h:81-268).  cpu_owned[v] != 0 gives vertex
 * v to host threads: in every add stage their state of each SSSP/BFS query lives
 * on the host, GPU relaxations into them become proposals, the host relaxes their
 * out-rows and proposes back for GPU-owned vertices, bulk synchronously, until
 * neither side has proposals; values are those of the all-GPU closure.  Built-in
 * programs only; CC queries and delete stages run on the GPU alone.  n must be
 * the vertex count; a null mask (or one with no vertex) turns it off. */
int dcg_set_cpu_domain(dcg_engine *eng, const uint8_t *cpu_owned, size_t n);
/* How long a workgroup of a persistent closure waits at the grid barrier before the
 * launch gives up (default 10 s; no reference counterpart: grid.sync() waits
 * forever).  A launch that gives up ends every round at once and the stage fails
 * with DCG_ERR_HIP ("a grid barrier timed out"); the engine's state is then
 * undefined and it should be destroyed.  0 gives up at the first wait. */
int dcg_set_barrier_timeout(dcg_engine *eng, uint64_t microseconds);
/* Batch maintenance (the reference's CG_BATCH_MAINTENANCE, effective_update_batch.h:
 * 34-45): 0 (default) groups batches of up to 4,096 records in one workgroup's
 * LDS sort and larger ones with the device-wide radix sort; 1 ("large") sends
 * every batch through the device-wide sort.  Values and rows never depend on it. */
int dcg_set_batch_maintenance(dcg_engine *eng, int large);
/* Closure round trace (a diagnostic with no reference counterpart): with
 * capacity > 0, workgroup 0 of every persistent closure launch appends one
 * record at the start of each round (and one per bucket advance and one when it
 * leaves the launch) until `capacity` records are held; 0 turns it off.  tag =
 * control-block index | mode << 8 (0 push, 1 pull, 2 invalidate, 3 PageRank
 * push, 4 PageRank amend) | DCG_ROUND_* flags; frontier = entries queued for the
 * round; t = device wall clock (100 MHz). */
typedef struct {
    uint32_t tag, round, frontier, reserved;
    uint64_t t;
} dcg_round_rec;
#define DCG_ROUND_STATIC  (1u << 12) /* static convergence (dcg_start, dcg_autotune) */
#define DCG_ROUND_ADVANCE (1u << 13) /* a bucket advance refilled the frontier */
#define DCG_ROUND_END     (1u << 14) /* workgroup 0 left the launch */
#define DCG_ROUND_SOLO    (1u << 15) /* the round ran on workgroup 0 alone */
int dcg_set_round_trace(dcg_engine *eng, uint32_t capacity);
/* Copies up to `cap` records (oldest first) into out, sets *written to the count
 * appended since the last call (more than the capacity when some were
 * dropped) and clears the trace. */
int dcg_round_trace(dcg_engine *eng, dcg_round_rec *out, size_t cap, size_t *written);
/* Set the tuned parameters directly (0 keeps a value). */
int dcg_set_tuning(dcg_engine *eng, uint32_t blocks_per_cu, uint32_t big_row, uint32_t piece);
/* The closure grid the stage autotuner also picks (dcg_tune_result.grid_shift,
 * .solo_max): persistent closures span (workgroups per CU * CUs) >> grid_shift
 * workgroups (at least one; grid_shift <= 16), and frontiers of at most
 * solo_max entries run on one workgroup alone.  Values never depend on it.
 * Replaces the reference's fixed launch sizing (KernelSizing,
 * include/framework/traversal.h:80-90) as dcg_set_tuning does. */
int dcg_set_closure_grid(dcg_engine *eng, uint32_t grid_shift, uint32_t solo_max);
/* Bucket width of the push closure (ordered repair, ordered_gpu_repair.cuh:116-160,
 * used by every SSSP/BFS closure here): SSSP queries w_max*quarters/4, BFS
 * queries `quarters` levels; 0 turns the buckets off (plain frontier order);
 * 0xFFFFFFFF (the default) picks 64 on graphs whose largest out-degree is at
 * most 64 (grids, road networks) and 0 otherwise.  Values never depend on it. */
int dcg_set_buckets(dcg_engine *eng, uint32_t quarters);
/* Replay each stage's repair launches (seeds, invalidation, reset, pull, push
 * closures, CC parent check) as one captured hipGraph (default on; no reference
 * counterpart).  Per-snapshot arguments (the batch's device records, its
 * count, round stamps) live in device memory, so the graph captured at the
 * first snapshot serves all later ones.  Each closure inside a graph is timed by
 * an event pair captured with it.  Values never depend on it. */
int dcg_set_graphs(dcg_engine *eng, int on);

/* Link microbenchmarks on `device`: HBM copy rate (read+write bytes / time of a
 * `bytes`-sized copy kernel) and zero-copy read rate of pinned host memory. */
int dcg_calibrate(int32_t device, uint64_t bytes, double *hbm_copy_gbs, double *host_read_gbs);
/* Random 64-byte line reads per second (16 lanes read one line, as the delete
 * search reads a prefix window) over `bytes` of pinned host memory and of HBM:
 * the host link's random-read ceiling that bounds the delete mutation of a
 * host-placed store.  No reference counterpart (measurement utility). */
int dcg_calibrate_lines(int32_t device, uint64_t bytes, double *host_lines_per_s, double *hbm_lines_per_s);

/* Wait for every stream of the engine: the engine stream (which also runs the
 * learned policy's training), the second mutation stream and the batch copy
 * stream (timing boundaries). */
int dcg_sync(dcg_engine *eng);

/* Graph and update files.  market_big text (src/utils/parser.cpp:296-406):
 * "src dst [w]" per line, n = largest id + 1, edges in file order; lines that
 * do not start with two integers are skipped.  A file that starts with the
 * 8 bytes "DCGEDGE1" is this library's binary edge list instead (u32 version
 * 1, u32 num_nodes, u64 num_edges, u32 sources, u32 destinations), read
 * without parsing.  threads <= 0: all host cores.  The arrays are allocated
 * here and released with dcg_free.  No device is used. */
int dcg_read_graph(const char *path, uint32_t *num_nodes, uint32_t **src, uint32_t **dst, uint64_t *num_edges,
                   int threads);
/* binary != 0 writes the binary edge list, else market_big lines "src dst". */
int dcg_write_graph(const char *path, uint32_t num_nodes, const uint32_t *src, const uint32_t *dst,
                    uint64_t num_edges, int binary);
/* Update stream files (include/framework/Loader.h:30-101): the size file holds
 * "<adds> <dels>" per batch (sizes[2b], sizes[2b+1]); the update file "a u v w"
 * and "d u v w" lines, additions and deletions each in file order (w is not
 * kept: the engine applies its program's weights). */
int dcg_read_updates(const char *sizes_path, const char *updates_path, uint64_t **sizes, uint64_t *num_batches,
                     dcg_edge **adds, uint64_t *num_adds, dcg_edge **dels, uint64_t *num_dels, int threads);
/* Writes each batch's deletions then its additions, w = (u + v) % 128 + 1. */
int dcg_write_updates(const char *sizes_path, const char *updates_path, const uint64_t *sizes, uint64_t num_batches,
                      const dcg_edge *adds, const dcg_edge *dels);
void dcg_free(void *p);

/* Seeded input generators (no reference counterpart: the reference reads
 * files produced by its scripts/ tools).  RMAT: the first num_edges distinct
 * non-self-loop pairs in draw order of a counter-based RMAT(a,b,c) stream,
 * returned sorted by (src,dst).  Buffers are caller-allocated host memory. */
int dcg_gen_rmat(uint32_t scale, uint64_t num_edges, double a, double b, double c,
                 uint64_t seed, uint32_t *src, uint32_t *dst);
/* Same edge set as dcg_gen_rmat, generated on the GPU (stable radix sort of the
 * draws plus a first-occurrence scan); src/dst are host buffers. */
int dcg_gen_rmat_gpu(int32_t device, uint32_t scale, uint64_t num_edges, double a, double b, double c,
                     uint64_t seed, uint32_t *src, uint32_t *dst);
/* Road-like grid (side x side, ids r*side+c): 4-neighbour arcs both ways plus a
 * seeded diag_frac of cells with a diagonal to (r+1,c+1), both ways; sorted by
 * (src,dst).  With src == NULL only *num_edges is set. */
int dcg_gen_grid(uint32_t side, double diag_frac, uint64_t seed, uint32_t *src, uint32_t *dst,
                 uint64_t *num_edges);
/* `count` batch sizes lognormal(mu, sigma) clipped to [lo, hi]. */
int dcg_gen_lognormal_sizes(uint32_t count, double mu, double sigma, uint64_t lo, uint64_t hi, uint64_t seed,
                            uint64_t *out);
/* Update stream: per snapshot `dels` deletions uniform over the edges present
 * at the snapshot start (without replacement) and `adds` insertions uniform
 * over absent non-loop pairs (distinct within the snapshot). The graph must be
 * given sorted by (src,dst) without duplicates. */
int dcg_gen_stream(uint32_t num_nodes, const uint32_t *src, const uint32_t *dst,
                   uint64_t num_edges, uint32_t snapshots, const uint64_t *adds_per,
                   const uint64_t *dels_per, uint64_t seed, dcg_edge *adds, dcg_edge *dels);

#ifdef __cplusplus
}
#endif
#endif
