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
