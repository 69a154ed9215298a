"""Vertex programs defined outside the engine (tests/programs/widest_path.hip,
built as a plugin from include/dcg_program_plugin.hpp) run through the engine's
repair kernels: the widest (maximum-bottleneck) path and hop-limited BFS, two
queries each, in one engine.  After the static convergence and every delete and
add stage, every query equals a brute-force recompute on the multiset model's
rows: a vectorised Bellman-Ford fixpoint of the program's own recurrence
(width[v] = max over in-edges of min(width[u], capacity); level[v] = min over
in-edges of level[u] + 1, capped), and every widest-path parent is a witness."""
import os

import numpy as np
import pytest

import oracle
from tests.conftest import announce
from tests.helpers import rmat_graph, same_rows

pytestmark = pytest.mark.gpu

PLUGIN = os.path.join(os.path.dirname(__file__), "programs", "libwidest_path.so")
INF = 0xFFFFFFFF
MAX_HOPS = 5
WMAX, WSEED = 1000, 7


def _capacities(src, dst):
    return np.array([oracle.weight(oracle.W_SEEDED, WSEED, WMAX, int(u), int(v)) for u, v in zip(src, dst)],
                    np.int64)


def brute_widest(n, rp, col, source):
    """Widths by fixpoint iteration: width[source] = +inf, width[v] = max over
    in-edges (u, v) of min(width[u], capacity(u, v)); 0 = unreached."""
    src = np.repeat(np.arange(n, dtype=np.int64), np.diff(np.asarray(rp, np.int64)))
    dst = np.asarray(col, np.int64)
    cap = _capacities(src, dst)
    big = np.int64(1) << 40
    w = np.zeros(n, np.int64)
    w[source] = big
    while True:
        offer = np.minimum(w[src], cap)
        nw = w.copy()
        np.maximum.at(nw, dst, offer)
        nw[source] = big
        if np.array_equal(nw, w):
            return w, (src, dst, cap)
        w = nw


def brute_hops(n, rp, col, source):
    src = np.repeat(np.arange(n, dtype=np.int64), np.diff(np.asarray(rp, np.int64)))
    dst = np.asarray(col, np.int64)
    lvl = np.full(n, INF, np.int64)
    lvl[source] = 0
    while True:
        offer = np.where(lvl[src] < MAX_HOPS, lvl[src] + 1, INF)
        nl = lvl.copy()
        np.minimum.at(nl, dst, offer)
        if np.array_equal(nl, lvl):
            return lvl
        lvl = nl


def _as_values(width):
    """Engine values of the widest-path program: ~width, IDENTITY when unreached."""
    out = np.full(width.size, INF, np.uint64)
    reach = width > 0
    out[reach] = np.where(width[reach] >= (1 << 40), 0, (~width[reach].astype(np.uint64)) & 0xFFFFFFFF)
    return out.astype(np.uint32)


def _check(eng, store, n, sources, stage):
    rp, col = store.csr()
    for q, s in enumerate(sources[:2]):
        width, (esrc, edst, cap) = brute_widest(n, rp, col, s)
        got = eng.gather_value(q)
        want = _as_values(width)
        assert np.array_equal(got, want), (stage, "widest", q, int(np.sum(got != want)))
        par = eng.gather_parent(q).astype(np.int64)
        kids = np.flatnonzero((width > 0) & (np.arange(n) != s))
        missing = kids[par[kids] == INF]
        if missing.size:
            rows = []
            for v in missing[:4].tolist():
                ins = np.flatnonzero(edst == v)
                rows.append((v, int(width[v]), [(int(esrc[i]), int(cap[i]), int(width[esrc[i]]),
                                                 int(par[esrc[i]])) for i in ins[:6]]))
            raise AssertionError((stage, "widest parents missing", int(missing.size), q, s, rows))
        # each parent edge exists and carries the vertex's width
        key = (esrc << 32) | edst
        capmap = dict(zip(key.tolist(), cap.tolist()))
        for v in kids[:: max(1, len(kids) // 500)]:
            p = int(par[v])
            c = capmap.get((p << 32) | int(v))
            assert c is not None and min(int(width[p]), c) == int(width[v]), (stage, "witness", int(v), p)
    for q, s in enumerate(sources[2:], start=2):
        want = brute_hops(n, rp, col, s).astype(np.uint32)
        got = eng.gather_value(q)
        assert np.array_equal(got, want), (stage, "hops", q, int(np.sum(got != want)))
    erp, ecol = eng.csr()
    assert same_rows(erp, ecol, store), f"rows differ after {stage}"


@pytest.mark.parametrize("placement,graphs", [(0, True), (1, True), (0, False)])
def test_plugin_programs_widest_path_and_hop_bfs(placement, graphs):
    import dycgraph_amd as d
    announce(f"plugin programs widest path + hop-limited BFS, placement={placement}, graphs={graphs}")
    assert os.path.exists(PLUGIN), f"{PLUGIN} is not built (make -C tests/programs)"
    table = d.load_program_plugin(PLUGIN, "dcg_widest_path_programs")
    n, src, dst = rmat_graph(11, 12000, seed=5)
    per = [400, 400, 400]
    adds, dels = d.gen_stream(n, src, dst, per, per, seed=21)
    sources = (0, int(src[777]), 0, int(src[4321]))
    eng = d.Engine(d.EngineConfig(app="sssp", sources=sources, query_apps=(0, 0, 1, 1), weight_mode=d.W_SEEDED,
                                  weight_max=WMAX, weight_seed=WSEED, placement=placement, cache_bytes=1 << 18))
    eng.set_program(table)
    eng.set_tuning(graphs=graphs)
    eng.load_graph(n, src, dst)
    eng.start()
    eng.load_updates(adds, dels)
    store = oracle.Store(n, src, dst)
    _check(eng, store, n, sources, "static")
    b = 0
    for s, k in enumerate(per):
        eng.del_edge(b, k)
        store.apply(dels[b:b + k], [])
        _check(eng, store, n, sources, f"delete {s}")
        eng.add_edge(b, k)
        store.apply([], adds[b:b + k])
        _check(eng, store, n, sources, f"add {s}")
        eng.cache_maintain()
        b += k
    eng.close()


def test_program_table_checks():
    """Tables are refused after a load, for PageRank engines, and with a
    query_apps entry beyond the table's programs."""
    import dycgraph_amd as d
    announce("program table checks")
    table = d.load_program_plugin(PLUGIN, "dcg_widest_path_programs")
    n, src, dst = rmat_graph(8, 1000, seed=1)
    eng = d.Engine(d.EngineConfig(app="sssp", sources=(0, 1), query_apps=(0, 2)))
    eng.set_program(table)
    with pytest.raises(d.DcgError):
        eng.load_graph(n, src, dst)          # program 2 does not exist in a 2-program table
    eng.close()
    eng = d.Engine(d.EngineConfig(app="sssp", sources=(0,)))
    eng.load_graph(n, src, dst)
    with pytest.raises(d.DcgError):
        eng.set_program(table)               # after a load
    eng.close()
    eng = d.Engine(d.EngineConfig(app="pr"))
    with pytest.raises(d.DcgError):
        eng.set_program(table)
    eng.close()
