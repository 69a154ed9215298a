"""CPU oracle (TEST INFRASTRUCTURE ONLY).

ctypes wrapper over oracle/liboracle.so (oracle.c).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may use it, and only as
the checker or as the timed CPU baseline; the product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
INF = 0xFFFFFFFF
W_REFERENCE, W_SEEDED, W_UNIT, W_ZERO = 0, 1, 2, 3
APP_SSSP, APP_BFS, APP_CC, APP_PR = 0, 1, 2, 3

_L = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


class BatchMetrics(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("deletions", "additions", "touched_sources", "changed_sources",
                                         "missing_deletes", "invalid_additions", "graph_edges")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SsspCheck(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("relaxable_edges", "missing_tight_witnesses", "invalid_parents",
                                         "missing_parent_edges", "wrong_parent_distances")] + [("source_ok", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PrCheck(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("l1", "bound", "max_error", "defect", "max_residual", "roundoff")] + \
               [(k, C.c_int) for k in ("converged", "finite", "ok")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P, U32, U64, I, D = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int, C.c_double
        sig = {
            "orc_mix64": (U64, [U64]),
            "orc_weight": (U32, [I, U64, U32, U32, U32]),
            "orc_store_create": (P, [U32]),
            "orc_store_destroy": (None, [P]),
            "orc_store_load": (I, [P, P, P, U64]),
            "orc_store_apply": (None, [P, P, U64, P, U64, C.POINTER(BatchMetrics)]),
            "orc_store_edges": (U64, [P]),
            "orc_store_degree": (U32, [P, U32]),
            "orc_store_row": (U32, [P, U32, P, U32]),
            "orc_store_csr": (None, [P, P, P]),
            "orc_hash_sequence": (U64, [P, U64]),
            "orc_sssp": (None, [U32, P, P, U32, I, U64, U32, P]),
            "orc_bfs": (None, [U32, P, P, U32, P]),
            "orc_cc": (None, [U32, P, P, P]),
            "orc_pagerank": (I, [U32, P, P, D, D, I, P]),
            "orc_value_checksum": (U64, [P, U32]),
            "orc_check_sssp": (None, [U32, P, P, U32, I, U64, U32, P, P, C.POINTER(SsspCheck)]),
            "orc_check_pagerank": (None, [U32, P, P, D, D, D, P, P, C.POINTER(PrCheck)]),
            "orc_incr_create": (P, [U32, P, P, I, U32, I, U64, U32, D, D, I]),
            "orc_incr_destroy": (None, [P]),
            "orc_incr_set_state": (None, [P, P, P]),
            "orc_incr_converge": (None, [P]),
            "orc_incr_pr_start": (None, [P]),
            "orc_incr_delete_phase": (None, [P, P, U64]),
            "orc_incr_add_phase": (None, [P, P, U64]),
            "orc_incr_pr_batch": (None, [P, P, U64, P, U64]),
            "orc_incr_values": (None, [P, P, P]),
            "orc_incr_pr_values": (None, [P, P, P]),
            "orc_incr_edges": (U64, [P]),
            "orc_incr_csr": (None, [P, P, P]),
            "orc_mlp_forward_backward": (None, [I, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P]),
            "orc_hot_candidates": (U64, [U32, P, P, P, U64, P]),
            "orc_cpu_create": (P, [U32, P, P, U64, I]),
            "orc_cpu_destroy": (None, [P]),
            "orc_cpu_threads": (I, [P]),
            "orc_cpu_edges": (U64, [P]),
            "orc_cpu_add_query": (I, [P, I, U32, I, U64, U32]),
            "orc_cpu_converge": (None, [P, I]),
            "orc_cpu_set_state": (None, [P, I, P, P]),
            "orc_cpu_delete_phase": (U64, [P, P, U64]),
            "orc_cpu_add_phase": (U64, [P, P, U64]),
            "orc_cpu_mutate_only": (None, [P, P, U64, P, U64]),
            "orc_cpu_pr_init": (None, [P, C.c_double, C.c_double, U32]),
            "orc_cpu_pr_set_state": (None, [P, P, P]),
            "orc_cpu_pr_values": (None, [P, P, P]),
            "orc_cpu_pr_batch": (U32, [P, P, U64, P, U64]),
            "orc_cpu_values": (None, [P, I, P, P]),
            "orc_cpu_csr": (None, [P, P, P]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _L = L
    return _L


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _edges(pairs):
    return np.ascontiguousarray(np.asarray(pairs, dtype=np.uint32).reshape(-1, 2))


def csr_from_edges(n, src, dst):
    """CSR whose rows keep the given edge order (file order)."""
    src = np.asarray(src, dtype=np.uint32)
    dst = np.asarray(dst, dtype=np.uint32)
    order = np.argsort(src, kind="stable")
    counts = np.bincount(src, minlength=n).astype(np.uint64)
    rp = np.zeros(n + 1, np.uint64)
    np.cumsum(counts, out=rp[1:])
    return rp, np.ascontiguousarray(dst[order])


class Store:
    """Multiset topology model (TopologyReplayModel)."""

    def __init__(self, n, src=(), dst=()):
        self.h = lib().orc_store_create(n)
        self.n = n
        src = np.ascontiguousarray(src, dtype=np.uint32)
        dst = np.ascontiguousarray(dst, dtype=np.uint32)
        if lib().orc_store_load(self.h, _p(src), _p(dst), len(src)):
            raise ValueError("edge id out of range")

    def apply(self, dels=(), adds=()):
        d, a = _edges(dels), _edges(adds)
        m = BatchMetrics()
        lib().orc_store_apply(self.h, _p(d), len(d), _p(a), len(a), C.byref(m))
        return m.as_dict()

    def row(self, u):
        deg = lib().orc_store_degree(self.h, u)
        out = np.empty(max(deg, 1), np.uint32)
        lib().orc_store_row(self.h, u, _p(out), deg)
        return out[:deg].tolist()

    def edges(self):
        return lib().orc_store_edges(self.h)

    def csr(self):
        rp = np.empty(self.n + 1, np.uint64)
        col = np.empty(max(self.edges(), 1), np.uint32)
        lib().orc_store_csr(self.h, _p(rp), _p(col))
        return rp, col[: self.edges()]

    def __del__(self):
        try:
            lib().orc_store_destroy(self.h)
        except Exception:
            pass


def hash_sequence(xs):
    a = np.ascontiguousarray(xs, dtype=np.uint32)
    return lib().orc_hash_sequence(_p(a), len(a))


def sssp(n, rp, col, source, wmode=W_REFERENCE, seed=1, wmax=100):
    out = np.empty(n, np.uint32)
    lib().orc_sssp(n, _p(np.ascontiguousarray(rp, np.uint64)), _p(np.ascontiguousarray(col, np.uint32)),
                   source, wmode, seed, wmax, _p(out))
    return out


def bfs(n, rp, col, source):
    out = np.empty(n, np.uint32)
    lib().orc_bfs(n, _p(np.ascontiguousarray(rp, np.uint64)), _p(np.ascontiguousarray(col, np.uint32)), source,
                  _p(out))
    return out


def cc(n, rp, col):
    out = np.empty(n, np.uint32)
    lib().orc_cc(n, _p(np.ascontiguousarray(rp, np.uint64)), _p(np.ascontiguousarray(col, np.uint32)), _p(out))
    return out


def pagerank(n, rp, col, base=0.15, tol_sum=None, max_rounds=10000):
    x = np.empty(n, np.float64)
    if tol_sum is None:
        tol_sum = 1e-12 * (base / 0.15) * max(n, 1)
    r = lib().orc_pagerank(n, _p(np.ascontiguousarray(rp, np.uint64)), _p(np.ascontiguousarray(col, np.uint32)),
                           base, tol_sum, max_rounds, _p(x))
    return x, r


def checksum(values):
    a = np.ascontiguousarray(values, dtype=np.uint32)
    return lib().orc_value_checksum(_p(a), len(a))


def weight(mode, seed, wmax, u, v):
    return lib().orc_weight(mode, seed, wmax, u, v)


def check_sssp(n, rp, col, source, dist, parent=None, wmode=W_REFERENCE, seed=1, wmax=100):
    r = SsspCheck()
    lib().orc_check_sssp(n, _p(np.ascontiguousarray(rp, np.uint64)), _p(np.ascontiguousarray(col, np.uint32)),
                         source, wmode, seed, wmax, _p(np.ascontiguousarray(dist, np.uint32)),
                         None if parent is None else _p(np.ascontiguousarray(parent, np.uint32)), C.byref(r))
    return r.as_dict()


def check_pagerank(n, rp, col, rank, residual, base=0.15, eps=1e-6, max_abs=None):
    """The reference's acceptance (hybrid_pr_host.cu:9-56) against an FP64 solve,
    plus max |rank - FP64| <= max_abs when given."""
    r = PrCheck()
    lib().orc_check_pagerank(n, _p(np.ascontiguousarray(rp, np.uint64)), _p(np.ascontiguousarray(col, np.uint32)),
                             base, eps, 0.0 if max_abs is None else float(max_abs),
                             _p(np.ascontiguousarray(rank, np.float32)),
                             _p(np.ascontiguousarray(residual, np.float32)), C.byref(r))
    return r.as_dict()


def hot_candidates(window, degree, cap, value=None):
    """The reference's hotness ranking and admission: (count, full ranking)."""
    w = np.ascontiguousarray(window, np.uint32)
    d = np.ascontiguousarray(degree, np.uint32)
    v = None if value is None else np.ascontiguousarray(value, np.uint32)
    ids = np.empty(len(w), np.uint32)
    k = lib().orc_hot_candidates(len(w), _p(w), None if v is None else _p(v), _p(d), int(cap), _p(ids))
    return k, ids


class CpuPath:
    """The reference's CPU path on all host cores (cpu_path.c): SSSP/BFS/CC
    queries over one store; the timed CPU baseline."""

    def __init__(self, n, src, dst, threads=0):
        self.n = n
        self._src = np.ascontiguousarray(src, np.uint32)
        self._dst = np.ascontiguousarray(dst, np.uint32)
        self._h = lib().orc_cpu_create(n, _p(self._src), _p(self._dst), len(self._src), threads)
        self._src = self._dst = None
        self.threads = lib().orc_cpu_threads(self._h)
        self.queries = 0

    def add_query(self, app, source=0, wmode=W_REFERENCE, seed=1, wmax=100):
        self.queries += 1
        return lib().orc_cpu_add_query(self._h, app, source, wmode, seed, wmax)

    def converge(self, q):
        lib().orc_cpu_converge(self._h, q)

    def set_state(self, q, value, parent=None):
        v = np.ascontiguousarray(value, np.uint32)
        p = None if parent is None else np.ascontiguousarray(parent, np.uint32)
        lib().orc_cpu_set_state(self._h, q, _p(v), None if p is None else _p(p))

    def delete_phase(self, dels):
        e = _edges(dels)
        return lib().orc_cpu_delete_phase(self._h, _p(e), len(e))

    def add_phase(self, adds):
        e = _edges(adds)
        return lib().orc_cpu_add_phase(self._h, _p(e), len(e))

    def mutate_only(self, dels, adds):
        """One snapshot's mutations without repair (values go stale until set_state)."""
        d, a = _edges(dels), _edges(adds)
        lib().orc_cpu_mutate_only(self._h, _p(d), len(d), _p(a), len(a))

    def pr_init(self, base, eps, max_rounds):
        """PageRank state: rank 0 and residual `base` everywhere (pr_batch of no
        records then runs the static solve)."""
        lib().orc_cpu_pr_init(self._h, base, eps, max_rounds)

    def pr_set_state(self, rank, res):
        lib().orc_cpu_pr_set_state(self._h, _p(np.ascontiguousarray(rank, np.float32)),
                                   _p(np.ascontiguousarray(res, np.float32)))

    def pr_batch(self, dels, adds):
        """Amend(-1), mutation, Amend(+1), dense rounds; returns the rounds run."""
        d, a = _edges(dels), _edges(adds)
        return lib().orc_cpu_pr_batch(self._h, _p(d), len(d), _p(a), len(a))

    def pr_values(self):
        r = np.empty(self.n, np.float32)
        s = np.empty(self.n, np.float32)
        lib().orc_cpu_pr_values(self._h, _p(r), _p(s))
        return r, s

    def values(self, q):
        v = np.empty(self.n, np.uint32)
        p = np.empty(self.n, np.uint32)
        lib().orc_cpu_values(self._h, q, _p(v), _p(p))
        return v, p

    def csr(self):
        m = lib().orc_cpu_edges(self._h)
        rp = np.empty(self.n + 1, np.uint64)
        col = np.empty(max(m, 1), np.uint32)
        lib().orc_cpu_csr(self._h, _p(rp), _p(col))
        return rp, col[:m]

    def close(self):
        if self._h:
            lib().orc_cpu_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class Incr:
    """CPU incremental path (port of the reference's CPU-capable path)."""

    def __init__(self, n, rp, col, app, source=0, wmode=W_REFERENCE, seed=1, wmax=100, pr_base=0.15,
                 pr_eps=1e-6, pr_max_rounds=100):
        self.n, self.app = n, app
        self.h = lib().orc_incr_create(n, _p(np.ascontiguousarray(rp, np.uint64)),
                                       _p(np.ascontiguousarray(col, np.uint32)), app, source, wmode, seed, wmax,
                                       pr_base, pr_eps, pr_max_rounds)

    def converge(self):
        if self.app == APP_PR:
            lib().orc_incr_pr_start(self.h)
        else:
            lib().orc_incr_converge(self.h)

    def set_state(self, value, parent):
        lib().orc_incr_set_state(self.h, _p(np.ascontiguousarray(value, np.uint32)),
                                 _p(np.ascontiguousarray(parent, np.uint32)))

    def delete_phase(self, dels):
        d = _edges(dels)
        lib().orc_incr_delete_phase(self.h, _p(d), len(d))

    def add_phase(self, adds):
        a = _edges(adds)
        lib().orc_incr_add_phase(self.h, _p(a), len(a))

    def pr_batch(self, dels, adds):
        d, a = _edges(dels), _edges(adds)
        lib().orc_incr_pr_batch(self.h, _p(d), len(d), _p(a), len(a))

    def values(self):
        v = np.empty(self.n, np.uint32)
        p = np.empty(self.n, np.uint32)
        lib().orc_incr_values(self.h, _p(v), _p(p))
        return v, p

    def pr_values(self):
        r = np.empty(self.n, np.float32)
        s = np.empty(self.n, np.float32)
        lib().orc_incr_pr_values(self.h, _p(r), _p(s))
        return r, s

    def csr(self):
        m = lib().orc_incr_edges(self.h)
        rp = np.empty(self.n + 1, np.uint64)
        col = np.empty(max(m, 1), np.uint32)
        lib().orc_incr_csr(self.h, _p(rp), _p(col))
        return rp, col[:m]

    def __del__(self):
        try:
            lib().orc_incr_destroy(self.h)
        except Exception:
            pass


def mlp_forward_backward(x, y, w1, b1, w2, b2, w3, b3, grads=True):
    rows = x.shape[0]
    f = lambda a: np.ascontiguousarray(a, np.float32)
    out = np.empty((rows, 2), np.float64)
    loss = C.c_double()
    g = [np.empty(s, np.float64) for s in ((8, 64), (64,), (64, 64), (64,), (64, 2), (2,))] if grads else [None] * 6
    lib().orc_mlp_forward_backward(rows, _p(f(x)), None if y is None else _p(f(y)), _p(f(w1)), _p(f(b1)), _p(f(w2)),
                                   _p(f(b2)), _p(f(w3)), _p(f(b3)), _p(out), C.cast(C.pointer(loss), C.c_void_p),
                                   *[_p(a) for a in g])
    return out, loss.value, g
