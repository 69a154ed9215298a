"""Host-side mirror of the reference engine surface over the C ABI (include/dcg.h).

Method names follow sepgraph::engine::Engine (include/framework/framework.cuh):
LoadGraph -> load_graph, InitGraph+Start -> start, get_update_file -> load_updates,
del_edge / add_edge / UpdatePageRank, Gather{Value,Parent,Buffer}, and the cache
maintenance sequence of samples/hybrid_sssp/hybrid_sssp.cu:494-521.

There is no fallback: if libdcg.so is missing or a HIP call fails, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdcg.so")

APP_SSSP, APP_BFS, APP_CC, APP_PR = 0, 1, 2, 3
W_REFERENCE, W_SEEDED = 0, 1
PLACE_DEVICE, PLACE_HOST = 0, 1
CACHE_NONE, CACHE_HOTNESS, CACHE_LEARNED = 0, 1, 2
INF = 0xFFFFFFFF
APPS = {"sssp": APP_SSSP, "bfs": APP_BFS, "cc": APP_CC, "pr": APP_PR}


class DcgError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[{status}] {message}")
        self.status = status


class Edge(C.Structure):
    _fields_ = [("u", C.c_uint32), ("v", C.c_uint32)]


class Config(C.Structure):
    _fields_ = [
        ("app", C.c_uint32), ("num_queries", C.c_uint32), ("sources", C.POINTER(C.c_uint32)),
        ("weight_mode", C.c_uint32), ("weight_max", C.c_uint32), ("weight_seed", C.c_uint64),
        ("placement", C.c_uint32), ("cache_policy", C.c_uint32), ("cache_bytes", C.c_uint64),
        ("store_slack", C.c_double), ("pr_base", C.c_double), ("pr_epsilon", C.c_double),
        ("pr_max_rounds", C.c_uint32), ("block_size", C.c_uint32), ("device", C.c_int32),
        ("query_apps", C.POINTER(C.c_uint32)), ("reserved", C.c_uint32 * 4),
    ]


class PhaseStats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in (
        "requests", "touched_sources", "changed_sources", "missing_deletes", "invalid_additions",
        "relocated_rows", "affected", "rounds", "edges_scanned", "cached_edge_reads",
        "host_edge_reads")] + [("ms", C.c_double)] + [(k, C.c_uint64) for k in (
        "bucket_advances", "far_entries", "coexec_rounds", "cpu_relaxations", "proposals_to_cpu",
        "proposals_to_gpu")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class CacheStats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in (
        "frames", "resident_chunks", "total_chunks", "loaded_chunks", "evicted_chunks",
        "written_back_chunks", "h2d_bytes", "d2h_bytes")] + [("ms", C.c_double)] + [
        ("learned_decisions", C.c_uint64), ("fallback_decisions", C.c_uint64), ("mlp_capture", C.c_double),
        ("hot_capture", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class KernelStats(C.Structure):
    _fields_ = [("launches", C.c_uint64), ("ms", C.c_double), ("algo_bytes", C.c_double),
                ("edges", C.c_uint64), ("expansions", C.c_uint64), ("host_lines", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


KERNELS = {"push": 0, "pull": 1, "invalidate": 2, "pr_solve": 3, "pr_amend": 4, "delete_rows": 5,
           "append_rows": 6, "delete_big": 7, "delete_pass": 8, "repair_delete": 9, "repair_add": 10,
           "ingest": 11, "static": 12}
# records spanning several kernels (not a single launch)
MULTI_KERNEL = ("delete_pass", "repair_delete", "repair_add")


class TuneResult(C.Structure):
    _fields_ = [("blocks_per_cu", C.c_uint32), ("big_row", C.c_uint32), ("piece", C.c_uint32),
                ("evaluations", C.c_uint32), ("best_ms", C.c_double), ("bucket_quarters", C.c_uint32),
                ("solo_max", C.c_uint32), ("first_ms", C.c_double), ("grid_shift", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


ROUND_REC = np.dtype([("tag", "<u4"), ("round", "<u4"), ("frontier", "<u4"), ("reserved", "<u4"), ("t", "<u8")])
ROUND_STATIC, ROUND_ADVANCE, ROUND_END, ROUND_SOLO = 1 << 12, 1 << 13, 1 << 14, 1 << 15


class TraceStats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("admitted_vertices", "admitted_edges", "changed_records",
                                          "capacity_edges", "bytes")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class EngineInfo(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in (
        "snapshots", "store_slots", "store_used_slots", "chunks", "frames", "host_bytes",
        "device_bytes", "cached_edge_reads", "host_edge_reads", "free_blocks", "reused_blocks",
        "store_device_bytes", "state_device_bytes", "frame_bytes")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib() -> C.CDLL:
    """Load libdcg.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `make -C dycgraph_amd`")
    L = C.CDLL(LIB_PATH)
    P, U32, U64, I = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
    sig = {
        "dcg_config_init": (I, [C.POINTER(Config)]),
        "dcg_create": (I, [C.POINTER(Config), C.POINTER(P)]),
        "dcg_destroy": (None, [P]),
        "dcg_last_error": (C.c_char_p, [P]),
        "dcg_status_string": (C.c_char_p, [I]),
        "dcg_load_edges": (I, [P, U32, P, P, U64]),
        "dcg_load_edges_device": (I, [P, U32, P, P, U64]),
        "dcg_load_csr": (I, [P, U32, P, P]),
        "dcg_start": (I, [P]),
        "dcg_set_program": (I, [P, P]),
        "dcg_load_updates": (I, [P, P, U64, P, U64]),
        "dcg_prefetch_updates": (I, [P, U64, U64, U64, U64]),
        "dcg_del_edge": (I, [P, U64, U64, C.POINTER(PhaseStats)]),
        "dcg_add_edge": (I, [P, U64, U64, C.POINTER(PhaseStats)]),
        "dcg_snapshot": (I, [P, U64, U64, U64, U64, C.POINTER(PhaseStats), C.POINTER(PhaseStats)]),
        "dcg_update_pagerank": (I, [P, U64, U64, U64, U64, C.POINTER(PhaseStats)]),
        "dcg_cache_maintain": (I, [P, C.POINTER(CacheStats)]),
        "dcg_cache_prefetch": (I, [P, U64, U64, U64, U64, C.POINTER(CacheStats)]),
        "dcg_cache_policy_set": (I, [P, U32]),
        "dcg_hot_candidates": (I, [P, U32, U64, P, U64, C.POINTER(U64)]),
        "dcg_hot_window": (I, [P, P, C.c_size_t]),
        "dcg_set_repair_snapshot": (I, [P, C.c_char_p, U32]),
        "dcg_set_cache_trace": (I, [P, C.c_char_p]),
        "dcg_trace_cache_candidates": (I, [P, U32, C.POINTER(TraceStats)]),
        "dcg_radix_sort": (I, [C.c_int32, P, P, U64, I, I]),
        "dcg_gather_values": (I, [P, U32, P, C.c_size_t]),
        "dcg_gather_parents": (I, [P, U32, P, C.c_size_t]),
        "dcg_gather_buffers": (I, [P, U32, P, C.c_size_t]),
        "dcg_get_csr": (I, [P, P, P, C.POINTER(U64)]),
        "dcg_num_nodes": (I, [P, C.POINTER(U32)]),
        "dcg_get_info": (I, [P, C.POINTER(EngineInfo)]),
        "dcg_sync": (I, [P]),
        "dcg_get_kernel_stats": (I, [P, U32, C.POINTER(KernelStats)]),
        "dcg_autotune": (I, [P, C.POINTER(TuneResult)]),
        "dcg_set_tuning": (I, [P, U32, U32, U32]),
        "dcg_set_closure_grid": (I, [P, U32, U32]),
        "dcg_autotune_stages": (I, [P, U32]),
        "dcg_set_batch_maintenance": (I, [P, I]),
        "dcg_set_cpu_domain": (I, [P, P, C.c_size_t]),
        "dcg_set_barrier_timeout": (I, [P, U64]),
        "dcg_set_round_trace": (I, [P, U32]),
        "dcg_round_trace": (I, [P, P, C.c_size_t, C.POINTER(C.c_size_t)]),
        "dcg_autotune_last": (I, [P, C.POINTER(TuneResult)]),
        "dcg_autotune_stage_last": (I, [P, U32, C.POINTER(TuneResult)]),
        "dcg_set_buckets": (I, [P, U32]),
        "dcg_set_graphs": (I, [P, I]),
        "dcg_reset_kernel_stats": (I, [P]),
        "dcg_mlp_eval": (I, [C.c_int32, U32, P, P, P, P, P]),
        "dcg_mlp_eval_device": (I, [C.c_int32, U32, P, P, P, P, P]),
        "dcg_gen_rmat": (I, [U32, U64, C.c_double, C.c_double, C.c_double, U64, P, P]),
        "dcg_gen_rmat_gpu": (I, [C.c_int32, U32, U64, C.c_double, C.c_double, C.c_double, U64, P, P]),
        "dcg_gen_stream": (I, [U32, P, P, U64, U32, P, P, U64, P, P]),
        "dcg_calibrate": (I, [C.c_int32, U64, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "dcg_calibrate_lines": (I, [C.c_int32, U64, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "dcg_gen_grid": (I, [U32, C.c_double, U64, P, P, C.POINTER(U64)]),
        "dcg_gen_lognormal_sizes": (I, [U32, C.c_double, C.c_double, U64, U64, U64, P]),
        "dcg_read_graph": (I, [C.c_char_p, C.POINTER(U32), C.POINTER(P), C.POINTER(P), C.POINTER(U64), I]),
        "dcg_write_graph": (I, [C.c_char_p, U32, P, P, U64, I]),
        "dcg_read_updates": (I, [C.c_char_p, C.c_char_p, C.POINTER(P), C.POINTER(U64), C.POINTER(P),
                                 C.POINTER(U64), C.POINTER(P), C.POINTER(U64), I]),
        "dcg_write_updates": (I, [C.c_char_p, C.c_char_p, P, U64, P, P]),
        "dcg_free": (None, [P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def edges_array(pairs) -> np.ndarray:
    """(k,2) uint32 array of (u, v) update records (the dcg_edge layout)."""
    a = np.ascontiguousarray(np.asarray(pairs, dtype=np.uint32).reshape(-1, 2))
    return a


def gen_rmat(scale: int, num_edges: int, a=0.57, b=0.19, c=0.19, seed=1, gpu=None, device=0):
    """Seeded RMAT edge set sorted by (src, dst).  gpu=None picks the GPU path for
    more than 2^22 edges; both paths return the same edges."""
    src = np.empty(num_edges, np.uint32)
    dst = np.empty(num_edges, np.uint32)
    if gpu is None:
        gpu = num_edges > (1 << 22)
    if gpu:
        st = lib().dcg_gen_rmat_gpu(device, scale, num_edges, a, b, c, seed, _ptr(src), _ptr(dst))
    else:
        st = lib().dcg_gen_rmat(scale, num_edges, a, b, c, seed, _ptr(src), _ptr(dst))
    if st:
        raise DcgError(st, "dcg_gen_rmat failed")
    return src, dst


def calibrate(nbytes=1 << 30, device=0):
    """(HBM copy GB/s, pinned-host zero-copy read GB/s) measured on the GPU."""
    h, z = C.c_double(), C.c_double()
    st = lib().dcg_calibrate(device, nbytes, C.byref(h), C.byref(z))
    if st:
        raise DcgError(st, "dcg_calibrate failed")
    return h.value, z.value


def calibrate_lines(nbytes=1 << 30, device=0):
    """(pinned-host, HBM) random 64-byte line reads per second measured on the GPU."""
    h, d = C.c_double(), C.c_double()
    st = lib().dcg_calibrate_lines(device, nbytes, C.byref(h), C.byref(d))
    if st:
        raise DcgError(st, "dcg_calibrate_lines failed")
    return h.value, d.value


def gen_grid(side, diag_frac=0.05, seed=1):
    m = C.c_uint64()
    st = lib().dcg_gen_grid(side, diag_frac, seed, None, None, C.byref(m))
    if st:
        raise DcgError(st, "dcg_gen_grid failed")
    src = np.empty(m.value, np.uint32)
    dst = np.empty(m.value, np.uint32)
    st = lib().dcg_gen_grid(side, diag_frac, seed, _ptr(src), _ptr(dst), C.byref(m))
    if st:
        raise DcgError(st, "dcg_gen_grid failed")
    return src, dst


def gen_lognormal_sizes(count, mu=8.0, sigma=1.5, lo=100, hi=100000, seed=1):
    out = np.empty(count, np.uint64)
    st = lib().dcg_gen_lognormal_sizes(count, mu, sigma, lo, hi, seed, _ptr(out))
    if st:
        raise DcgError(st, "dcg_gen_lognormal_sizes failed")
    return out


def gen_stream(n, src, dst, adds_per, dels_per, seed=7):
    adds_per = np.ascontiguousarray(adds_per, dtype=np.uint64)
    dels_per = np.ascontiguousarray(dels_per, dtype=np.uint64)
    adds = np.empty((int(adds_per.sum()), 2), np.uint32)
    dels = np.empty((int(dels_per.sum()), 2), np.uint32)
    src = np.ascontiguousarray(src, dtype=np.uint32)
    dst = np.ascontiguousarray(dst, dtype=np.uint32)
    st = lib().dcg_gen_stream(n, _ptr(src), _ptr(dst), len(src), len(adds_per), _ptr(adds_per),
                              _ptr(dels_per), seed, _ptr(adds), _ptr(dels))
    if st:
        raise DcgError(st, "dcg_gen_stream failed")
    return adds, dels


MLP_SHAPES = [(8, 64), (64,), (64, 64), (64,), (64, 2), (2,)]


_plugins = {}


def load_program_plugin(path, symbol):
    """Load a vertex-program plugin (a shared library built from
    include/dcg_program_plugin.hpp) and return its dcg_program table pointer;
    the library stays loaded for the life of the process."""
    lib()   # the engine's HIP runtime first
    if path not in _plugins:
        _plugins[path] = C.CDLL(os.path.abspath(path))
    fn = getattr(_plugins[path], symbol)
    fn.restype = C.c_void_p
    fn.argtypes = []
    ptr = fn()
    if not ptr:
        raise DcgError(-1, f"{symbol} returned no table")
    return C.c_void_p(ptr)


def _take(ptr, count, dtype):
    """Copy a library-allocated array into numpy and release it."""
    out = np.empty(count, dtype)
    if count:
        C.memmove(out.ctypes.data, ptr, out.nbytes)
    lib().dcg_free(ptr)
    return out


def read_graph(path, threads=0):
    """(n, src, dst) from a market_big text file or a binary edge list."""
    n, m = C.c_uint32(), C.c_uint64()
    s, d = C.c_void_p(), C.c_void_p()
    st = lib().dcg_read_graph(os.fsencode(path), C.byref(n), C.byref(s), C.byref(d), C.byref(m), threads)
    if st != 0:
        raise DcgError(st, f"dcg_read_graph failed: {path}")
    return n.value, _take(s, m.value, np.uint32), _take(d, m.value, np.uint32)


def write_graph(path, n, src, dst, binary=False):
    s = np.ascontiguousarray(src, np.uint32)
    d = np.ascontiguousarray(dst, np.uint32)
    st = lib().dcg_write_graph(os.fsencode(path), n, _ptr(s), _ptr(d), len(s), 1 if binary else 0)
    if st != 0:
        raise DcgError(st, f"dcg_write_graph failed: {path}")


def read_updates(sizes_path, updates_path, threads=0):
    """(sizes (batches x [adds, dels]), adds (k x 2), dels (k x 2)) from the reference's update files."""
    sz, a, d = C.c_void_p(), C.c_void_p(), C.c_void_p()
    nb, na, nd = C.c_uint64(), C.c_uint64(), C.c_uint64()
    st = lib().dcg_read_updates(os.fsencode(sizes_path), os.fsencode(updates_path), C.byref(sz), C.byref(nb),
                                C.byref(a), C.byref(na), C.byref(d), C.byref(nd), threads)
    if st != 0:
        raise DcgError(st, "dcg_read_updates failed")
    return (_take(sz, 2 * nb.value, np.uint64).reshape(-1, 2), _take(a, 2 * na.value, np.uint32).reshape(-1, 2),
            _take(d, 2 * nd.value, np.uint32).reshape(-1, 2))


def write_updates(sizes_path, updates_path, sizes, adds, dels):
    sz = np.ascontiguousarray(np.asarray(sizes, np.uint64).reshape(-1, 2))
    a = np.ascontiguousarray(np.asarray(adds, np.uint32).reshape(-1, 2))
    d = np.ascontiguousarray(np.asarray(dels, np.uint32).reshape(-1, 2))
    st = lib().dcg_write_updates(os.fsencode(sizes_path), os.fsencode(updates_path), _ptr(sz), len(sz), _ptr(a),
                                 _ptr(d))
    if st != 0:
        raise DcgError(st, "dcg_write_updates failed")


def radix_sort(keys, vals=None, begin_bit=0, end_bit=64, device=0):
    """Stable GPU radix sort of uint64 keys (and uint32 values) by bits
    [begin_bit, end_bit); returns sorted copies."""
    k = np.array(keys, dtype=np.uint64, copy=True)
    v = None if vals is None else np.array(vals, dtype=np.uint32, copy=True)
    if v is not None and len(v) != len(k):
        raise ValueError("keys and values differ in length")
    st = lib().dcg_radix_sort(device, _ptr(k), None if v is None else _ptr(v), len(k), begin_bit, end_bit)
    if st != 0:
        raise DcgError(st, "dcg_radix_sort failed")
    return k if v is None else (k, v)


def mlp_eval(x, params, y=None, device=0):
    """Run the policy's predictor on the GPU: returns (out, grads or None)."""
    x = np.ascontiguousarray(x, np.float32)
    flat = np.ascontiguousarray(np.concatenate([np.asarray(p, np.float32).reshape(-1) for p in params]))
    out = np.empty((x.shape[0], 2), np.float32)
    grad = np.empty(flat.size, np.float32)
    yy = None if y is None else np.ascontiguousarray(y, np.float32)
    st = lib().dcg_mlp_eval(device, x.shape[0], _ptr(x), None if yy is None else _ptr(yy), _ptr(flat), _ptr(out),
                            _ptr(grad))
    if st:
        raise DcgError(st, "dcg_mlp_eval failed")
    if y is None:
        return out, None
    gs, k = [], 0
    for shp in MLP_SHAPES:
        sz = int(np.prod(shp))
        gs.append(grad[k:k + sz].reshape(shp))
        k += sz
    return out, gs


@dataclass
class EngineConfig:
    app: str = "sssp"
    sources: tuple = (0,)
    weight_mode: int = W_REFERENCE
    weight_max: int = 100
    weight_seed: int = 1
    placement: int = PLACE_DEVICE
    cache_policy: int = CACHE_HOTNESS
    cache_bytes: int = 4_000_000_000
    store_slack: float = 0.125
    pr_base: float = 0.15
    pr_epsilon: float = 1e-6
    pr_max_rounds: int = 100
    device: int = 0
    query_apps: tuple = ()      # per-query program ("sssp"/"bfs"/"cc"); empty: all `app`


class Engine:
    """One engine handle (not re-entrant), owning device memory and the chunk store."""

    def __init__(self, config: EngineConfig | None = None, **kw):
        cfg = config or EngineConfig(**kw)
        self.cfg = cfg
        L = lib()
        c = Config()
        L.dcg_config_init(C.byref(c))
        c.app = APPS[cfg.app] if isinstance(cfg.app, str) else int(cfg.app)
        self.app = c.app
        srcs = np.ascontiguousarray(np.asarray(cfg.sources, dtype=np.uint32).reshape(-1))
        self._srcs = srcs
        c.num_queries = len(srcs)
        c.sources = srcs.ctypes.data_as(C.POINTER(C.c_uint32))
        c.weight_mode, c.weight_max, c.weight_seed = cfg.weight_mode, cfg.weight_max, cfg.weight_seed
        c.placement, c.cache_policy, c.cache_bytes = cfg.placement, cfg.cache_policy, cfg.cache_bytes
        c.store_slack = cfg.store_slack
        c.pr_base, c.pr_epsilon, c.pr_max_rounds = cfg.pr_base, cfg.pr_epsilon, cfg.pr_max_rounds
        c.device = cfg.device
        if cfg.query_apps:
            qa = np.ascontiguousarray([APPS[a] if isinstance(a, str) else int(a) for a in cfg.query_apps],
                                      dtype=np.uint32)
            assert len(qa) == len(srcs), "query_apps needs one entry per source"
            self._qapps = qa
            c.query_apps = qa.ctypes.data_as(C.POINTER(C.c_uint32))
        h = C.c_void_p()
        self._h = None
        self._check(L.dcg_create(C.byref(c), C.byref(h)), "dcg_create")
        self._h = h
        self.n = 0
        self.queries = 1 if self.app in (APP_CC, APP_PR) else len(srcs)

    def _check(self, st, what):
        if st:
            msg = lib().dcg_last_error(self._h).decode() if self._h else lib().dcg_status_string(st).decode()
            raise DcgError(st, f"{what}: {msg}")

    def close(self):
        if self._h:
            lib().dcg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_program(self, table):
        """Install a program table (a dcg_program pointer from load_program_plugin,
        or None for the built-in SSSP/BFS/CC set) before load_graph."""
        self._program = table
        self._check(lib().dcg_set_program(self._h, table), "dcg_set_program")

    # --- LoadGraph / InitGraph / Start
    def load_graph(self, n, src, dst):
        src = np.ascontiguousarray(src, dtype=np.uint32)
        dst = np.ascontiguousarray(dst, dtype=np.uint32)
        self._check(lib().dcg_load_edges(self._h, n, _ptr(src), _ptr(dst), len(src)), "dcg_load_edges")
        self.n = n

    def start(self):
        self._check(lib().dcg_start(self._h), "dcg_start")

    # --- update stream
    def load_updates(self, adds, dels):
        adds = edges_array(adds if len(adds) else np.zeros((0, 2)))
        dels = edges_array(dels if len(dels) else np.zeros((0, 2)))
        self._adds, self._dels = adds, dels
        self._check(lib().dcg_load_updates(self._h, _ptr(adds), len(adds), _ptr(dels), len(dels)),
                    "dcg_load_updates")

    def prefetch_updates(self, add_begin, add_count, del_begin, del_count):
        """Copy a coming snapshot's records to the device on the copy stream."""
        self._check(lib().dcg_prefetch_updates(self._h, add_begin, add_count, del_begin, del_count),
                    "dcg_prefetch_updates")

    def del_edge(self, begin, count):
        st = PhaseStats()
        self._check(lib().dcg_del_edge(self._h, begin, count, C.byref(st)), "dcg_del_edge")
        return st.as_dict()

    def add_edge(self, begin, count):
        st = PhaseStats()
        self._check(lib().dcg_add_edge(self._h, begin, count, C.byref(st)), "dcg_add_edge")
        return st.as_dict()

    def snapshot(self, del_begin, del_count, add_begin, add_count):
        """Delete stage then add stage of one snapshot with one host wait; the two
        stages' statistics (the step's wall time is on the delete record)."""
        ds, ads = PhaseStats(), PhaseStats()
        self._check(lib().dcg_snapshot(self._h, del_begin, del_count, add_begin, add_count, C.byref(ds),
                                       C.byref(ads)), "dcg_snapshot")
        return ds.as_dict(), ads.as_dict()

    def update_pagerank(self, add_begin, add_count, del_begin, del_count):
        st = PhaseStats()
        self._check(lib().dcg_update_pagerank(self._h, add_begin, add_count, del_begin, del_count,
                                              C.byref(st)), "dcg_update_pagerank")
        return st.as_dict()

    def cache_maintain(self):
        st = CacheStats()
        self._check(lib().dcg_cache_maintain(self._h, C.byref(st)), "dcg_cache_maintain")
        return st.as_dict()

    def cache_prefetch(self, add_begin, add_count, del_begin, del_count):
        st = CacheStats()
        self._check(lib().dcg_cache_prefetch(self._h, add_begin, add_count, del_begin, del_count,
                                             C.byref(st)), "dcg_cache_prefetch")
        return st.as_dict()

    def set_cache_policy(self, policy):
        self._check(lib().dcg_cache_policy_set(self._h, policy), "dcg_cache_policy_set")

    # --- Gather*
    def hot_candidates(self, cap_slots, query=0, limit=None):
        """The reference's hotness ranking and cache admission: (count, ids),
        ids = the first min(count, limit) ranked vertices; shifts the window."""
        limit = self.n if limit is None else limit
        ids = np.empty(max(limit, 1), np.uint32)
        cnt = C.c_uint64()
        self._check(lib().dcg_hot_candidates(self._h, query, cap_slots, _ptr(ids), limit, C.byref(cnt)),
                    "dcg_hot_candidates")
        return cnt.value, ids[:min(cnt.value, limit)]

    def set_repair_snapshot(self, path, query=0):
        """Write the next delete stage's repair snapshot to `path` ('' disarms)."""
        self._check(lib().dcg_set_repair_snapshot(self._h, os.fsencode(path), query), "dcg_set_repair_snapshot")

    def set_cache_trace(self, path):
        """Open (truncate) a cache-candidate trace at `path` ('' closes it)."""
        self._check(lib().dcg_set_cache_trace(self._h, os.fsencode(path)), "dcg_set_cache_trace")

    def trace_cache_candidates(self, batch):
        """Append the last hot_candidates ranking's admitted prefix as one record."""
        st = TraceStats()
        self._check(lib().dcg_trace_cache_candidates(self._h, batch, C.byref(st)), "dcg_trace_cache_candidates")
        return st.as_dict()

    def hot_window(self):
        out = np.empty(self.n, np.uint32)
        self._check(lib().dcg_hot_window(self._h, _ptr(out), self.n), "dcg_hot_window")
        return out

    def gather_value(self, query=0):
        out = np.empty(self.n, np.float32 if self.app == APP_PR else np.uint32)
        self._check(lib().dcg_gather_values(self._h, query, _ptr(out), self.n), "dcg_gather_values")
        return out

    def gather_parent(self, query=0):
        out = np.empty(self.n, np.uint32)
        self._check(lib().dcg_gather_parents(self._h, query, _ptr(out), self.n), "dcg_gather_parents")
        return out

    def gather_buffer(self, query=0):
        out = np.empty(self.n, np.float32 if self.app == APP_PR else np.uint32)
        self._check(lib().dcg_gather_buffers(self._h, query, _ptr(out), self.n), "dcg_gather_buffers")
        return out

    def csr(self):
        m = C.c_uint64()
        self._check(lib().dcg_get_csr(self._h, None, None, C.byref(m)), "dcg_get_csr")
        rp = np.empty(self.n + 1, np.uint64)
        col = np.empty(max(m.value, 1), np.uint32)
        self._check(lib().dcg_get_csr(self._h, _ptr(rp), _ptr(col), C.byref(m)), "dcg_get_csr")
        return rp, col[: m.value]

    def info(self):
        st = EngineInfo()
        self._check(lib().dcg_get_info(self._h, C.byref(st)), "dcg_get_info")
        return st.as_dict()

    def kernel_stats(self):
        out = {}
        for name, kid in KERNELS.items():
            st = KernelStats()
            self._check(lib().dcg_get_kernel_stats(self._h, kid, C.byref(st)), "dcg_get_kernel_stats")
            out[name] = st.as_dict()
        return out

    def autotune(self):
        st = TuneResult()
        self._check(lib().dcg_autotune(self._h, C.byref(st)), "dcg_autotune")
        return st.as_dict()

    def set_cpu_domain(self, cpu_owned):
        """Give the vertices with cpu_owned[v] != 0 to host threads in every add
        stage (co-execution); None turns it off."""
        if cpu_owned is None:
            self._check(lib().dcg_set_cpu_domain(self._h, None, 0), "dcg_set_cpu_domain")
            return
        m = np.ascontiguousarray(cpu_owned, np.uint8)
        self._check(lib().dcg_set_cpu_domain(self._h, _ptr(m), len(m)), "dcg_set_cpu_domain")

    def set_barrier_timeout(self, microseconds):
        """Grid-barrier patience of the persistent closures (0: give up at the first wait)."""
        self._check(lib().dcg_set_barrier_timeout(self._h, microseconds), "dcg_set_barrier_timeout")

    def set_round_trace(self, capacity):
        """Record one entry per round of every closure launch (capacity records; 0 off)."""
        self._check(lib().dcg_set_round_trace(self._h, capacity), "dcg_set_round_trace")
        self._rt_cap = capacity

    def round_trace(self, cap=1 << 20):
        """The records since the last call (numpy structured array: tag, round,
        frontier, t) and the count appended (above len() when some were dropped);
        clears the trace."""
        out = np.zeros(cap, ROUND_REC)
        written = C.c_size_t(0)
        self._check(lib().dcg_round_trace(self._h, out.ctypes.data, cap, C.byref(written)), "dcg_round_trace")
        return out[:min(cap, written.value, getattr(self, "_rt_cap", 0))], written.value

    def set_batch_maintenance(self, large):
        """True: every batch through the device-wide sort path (CG_BATCH_MAINTENANCE=large)."""
        self._check(lib().dcg_set_batch_maintenance(self._h, 1 if large else 0), "dcg_set_batch_maintenance")

    def autotune_stages(self, stages=1):
        """Tune the next `stages` repair stages on their own workload (each stage's
        repair runs once per candidate from the same post-mutation state)."""
        self._check(lib().dcg_autotune_stages(self._h, stages), "dcg_autotune_stages")

    def last_tune(self):
        st = TuneResult()
        self._check(lib().dcg_autotune_last(self._h, C.byref(st)), "dcg_autotune_last")
        return st.as_dict()

    def stage_tune(self, stage):
        """The last tuning of stage 0 (delete) or 1 (add)."""
        st = TuneResult()
        self._check(lib().dcg_autotune_stage_last(self._h, stage, C.byref(st)), "dcg_autotune_stage_last")
        return st.as_dict()

    def set_tuning(self, blocks_per_cu=0, big_row=0, piece=0, bucket_quarters=None, graphs=None):
        self._check(lib().dcg_set_tuning(self._h, blocks_per_cu, big_row, piece), "dcg_set_tuning")
        if bucket_quarters is not None:
            self._check(lib().dcg_set_buckets(self._h, bucket_quarters), "dcg_set_buckets")
        if graphs is not None:
            self._check(lib().dcg_set_graphs(self._h, 1 if graphs else 0), "dcg_set_graphs")

    def set_closure_grid(self, grid_shift, solo_max):
        """Closures span (workgroups per CU * CUs) >> grid_shift workgroups; frontiers
        of at most solo_max entries run on one workgroup alone."""
        self._check(lib().dcg_set_closure_grid(self._h, grid_shift, solo_max), "dcg_set_closure_grid")

    def reset_kernel_stats(self):
        self._check(lib().dcg_reset_kernel_stats(self._h), "dcg_reset_kernel_stats")

    def sync(self):
        self._check(lib().dcg_sync(self._h), "dcg_sync")
