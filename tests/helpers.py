"""Shared helpers for the parity tests (graph builders, per-phase oracle runs)."""
import numpy as np

import oracle


def rmat_graph(scale, edges, seed=1):
    import dycgraph_amd as d
    src, dst = d.gen_rmat(scale, edges, seed=seed)
    return 1 << scale, src, dst


def oracle_values(app, n, rp, col, source=0, wmode=oracle.W_REFERENCE, seed=1, wmax=100):
    if app == "sssp":
        return oracle.sssp(n, rp, col, source, wmode, seed, wmax)
    if app == "bfs":
        return oracle.bfs(n, rp, col, source)
    if app == "cc":
        return oracle.cc(n, rp, col)
    raise ValueError(app)


def same_rows(eng_rp, eng_col, store):
    """Row-by-row sequence equality of the engine's out-rows and the model."""
    rp, col = store.csr()
    return np.array_equal(np.asarray(eng_rp, np.uint64), rp) and np.array_equal(np.asarray(eng_col), col)


def cc_parents_ok(n, rp, col, labels, par):
    """CC parents are rooted witnesses: a vertex without a parent is its own
    label's root; otherwise its parent carries the same label and the edge
    parent -> vertex exists in the current rows."""
    INF = 0xFFFFFFFF
    labels = np.asarray(labels, np.uint64)
    par = np.asarray(par, np.uint64)
    v = np.arange(n, dtype=np.uint64)
    roots = par == INF
    if not np.array_equal(labels[roots], v[roots]):
        return False
    src = np.repeat(v, np.diff(np.asarray(rp, np.int64)))
    have = np.unique((src << np.uint64(32)) | np.asarray(col, np.uint64))
    kids = v[~roots]
    ps = par[~roots]
    if not np.array_equal(labels[ps], labels[kids]):
        return False
    return bool(np.isin((ps << np.uint64(32)) | kids, have).all())
