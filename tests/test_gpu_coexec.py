"""CPU/GPU co-execution of the insertion closure against the oracle.

Host threads own a vertex set during every add stage; the GPU closure and the
host closure exchange proposals until neither has any.  The values must equal
the oracle's from-scratch Dijkstra / BFS / flood fill after every stage (CC
runs on the GPU alone beside the co-executed queries), and both sides must have
done work: host relaxations and proposals in both directions.  The grid, with
the CPU owning a band of rows that every shortest path crosses, makes the
exchange go back and forth many times.
"""
import numpy as np
import pytest

import oracle
from tests.conftest import announce
from tests.helpers import cc_parents_ok, oracle_values, rmat_graph

pytestmark = pytest.mark.gpu


def _check(eng, store, n, apps, sources, wmax, stage):
    rp, col = store.csr()
    for q, (app, s) in enumerate(zip(apps, sources)):
        want = oracle_values(app, n, rp, col, s, oracle.W_SEEDED, 5, wmax)
        got = eng.gather_value(q)
        assert np.array_equal(got, want), (stage, q, app, int(np.sum(got != want)))
        if app == "cc":
            assert cc_parents_ok(n, rp, col, got, eng.gather_parent(q)), (stage, q)
        else:
            chk = oracle.check_sssp(n, rp, col, s, got, eng.gather_parent(q),
                                    oracle.W_SEEDED if app == "sssp" else oracle.W_UNIT, 5, wmax)
            assert chk["invalid_parents"] == 0 and chk["relaxable_edges"] == 0, (stage, q, chk)


@pytest.mark.parametrize("graph,placement", [("rmat", 0), ("rmat", 1), ("grid", 0)])
def test_coexec_insertion_closure_exact(graph, placement):
    import dycgraph_amd as d
    announce(f"co-execution {graph} placement {placement}")
    if graph == "grid":
        side = 64
        src, dst = d.gen_grid(side, 0.05, seed=2)
        n = side * side
        owned = np.zeros(n, np.uint8)
        owned[20 * side:30 * side] = 1          # a band of rows every path from row 0 crosses
        apps, sources, wmax, per = ("sssp", "bfs", "sssp", "cc"), (0, 5, 4000, 0), 1000, [600, 900, 400]
    else:
        n, src, dst = rmat_graph(13, 1 << 17, seed=6)
        owned = (np.random.default_rng(3).random(n) < 0.3).astype(np.uint8)
        apps, sources, wmax, per = ("sssp", "bfs", "cc"), (0, 1, 0), 100, [1500, 700, 1200]
    adds, dels = d.gen_stream(n, src, dst, per, per, seed=8)
    eng = d.Engine(d.EngineConfig(app="sssp", sources=sources, query_apps=apps, weight_mode=d.W_SEEDED,
                                  weight_max=wmax, weight_seed=5, placement=placement, cache_bytes=2 << 18))
    eng.load_graph(n, src, dst)
    eng.set_cpu_domain(owned)
    eng.start()
    eng.load_updates(adds, dels)
    store = oracle.Store(n, src, dst)
    _check(eng, store, n, apps, sources, wmax, "static")
    off, relax, to_cpu, to_gpu, rounds = 0, 0, 0, 0, 0
    for s, k in enumerate(per):
        eng.del_edge(off, k)
        store.apply(dels[off:off + k], [])
        _check(eng, store, n, apps, sources, wmax, f"delete {s}")
        st = eng.add_edge(off, k)
        store.apply([], adds[off:off + k])
        _check(eng, store, n, apps, sources, wmax, f"add {s}")
        relax += st["cpu_relaxations"]
        to_cpu += st["proposals_to_cpu"]
        to_gpu += st["proposals_to_gpu"]
        rounds = max(rounds, st["coexec_rounds"])
        off += k
    assert relax > 0 and to_cpu > 0 and to_gpu > 0, (relax, to_cpu, to_gpu)
    if graph == "grid":
        assert rounds >= 3   # the exchange went back and forth
    # turning it off again keeps the engine exact
    eng.set_cpu_domain(None)
    eng.del_edge(0, 0)
    eng.close()


def test_coexec_rejects_plugins_and_pagerank():
    import dycgraph_amd as d
    n, src, dst = rmat_graph(10, 1 << 13, seed=1)
    eng = d.Engine(d.EngineConfig(app="pr"))
    eng.load_graph(n, src, dst)
    with pytest.raises(d.DcgError):
        eng.set_cpu_domain(np.ones(n, np.uint8))
    eng.close()
    eng = d.Engine(d.EngineConfig(app="sssp"))
    eng.load_graph(n, src, dst)
    with pytest.raises(d.DcgError):
        eng.set_cpu_domain(np.ones(n - 1, np.uint8))
    eng.close()
