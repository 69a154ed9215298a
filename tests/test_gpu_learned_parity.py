"""The learned-cache engine against the oracle.

The benchmark's headline form is one engine repairing SSSP, BFS and CC queries
over a host-placed chunk store with a capped device cache, the learned policy
deciding admissions (with its prefetch before every snapshot).  The policy only
moves chunks, so every query must equal a from-scratch Dijkstra / FIFO BFS /
flood fill of the multiset model after every delete and add stage, and the
engine's rows (written through cached frames, then evicted) must equal the
model's rows.  Reference checks after every stage: hybrid_sssp.cu:452-489.
"""
import numpy as np
import pytest

import oracle
from tests.conftest import announce
from tests.helpers import cc_parents_ok, oracle_values, rmat_graph, same_rows

pytestmark = pytest.mark.gpu


def _check_all(eng, store, n, apps, sources, seed, wmax, stage):
    rp, col = store.csr()
    for q, (app, s) in enumerate(zip(apps, sources)):
        want = oracle_values(app, n, rp, col, s, oracle.W_SEEDED, seed, wmax)
        got = eng.gather_value(q)
        assert np.array_equal(got, want), (stage, q, app, int(np.sum(got != want)))
        if app == "cc":
            assert cc_parents_ok(n, rp, col, got, eng.gather_parent(q)), (stage, q)
        else:
            chk = oracle.check_sssp(n, rp, col, s, got, eng.gather_parent(q),
                                    oracle.W_SEEDED if app == "sssp" else oracle.W_UNIT, seed, wmax)
            assert chk["invalid_parents"] == 0 and chk["relaxable_edges"] == 0, (stage, q, chk)
    erp, ecol = eng.csr()
    assert same_rows(erp, ecol, store), f"rows differ after {stage}"


@pytest.mark.parametrize("scale,edges,per,frames", [(16, 1 << 20, 3000, 4), (18, 1 << 22, 10000, 16)])
def test_learned_cache_sssp_bfs_cc(scale, edges, per, frames):
    """RMAT graph, host placement, a cache of a few 256 KiB frames (a small share of
    the chunks), SSSP + BFS + CC in one engine under the learned policy, with the
    next snapshot's records prefetched; 4 snapshots of `per` deletes and inserts."""
    import dycgraph_amd as d
    announce(f"learned cache sssp+bfs+cc rmat{scale} {edges} edges, {per}+{per} x 4, {frames} frames")
    n, src, dst = rmat_graph(scale, edges, seed=2)
    snaps = 4
    adds, dels = d.gen_stream(n, src, dst, [per] * snaps, [per] * snaps, seed=17)
    rng = np.random.default_rng(23)
    picks = [int(src[i]) for i in rng.integers(0, len(src), 2)]
    apps, sources = ("sssp", "bfs", "cc"), (picks[0], picks[1], 0)
    eng = d.Engine(d.EngineConfig(app="sssp", sources=sources, query_apps=apps, weight_mode=d.W_SEEDED,
                                  weight_max=100, weight_seed=11, placement=d.PLACE_HOST,
                                  cache_policy=d.CACHE_LEARNED, cache_bytes=frames * (1 << 18)))
    eng.load_graph(n, src, dst)
    eng.start()
    eng.load_updates(adds, dels)
    store = oracle.Store(n, src, dst)
    _check_all(eng, store, n, apps, sources, 11, 100, "static")
    decisions = 0
    for s in range(snaps):
        b = s * per
        if s + 1 < snaps:
            eng.prefetch_updates(b + per, per, b + per, per)
        cst = eng.cache_prefetch(b, per, b, per)
        eng.del_edge(b, per)
        store.apply(dels[b:b + per], [])
        _check_all(eng, store, n, apps, sources, 11, 100, f"delete {s}")
        eng.add_edge(b, per)
        store.apply([], adds[b:b + per])
        _check_all(eng, store, n, apps, sources, 11, 100, f"add {s}")
        cst = eng.cache_maintain()
        decisions = cst["learned_decisions"] + cst["fallback_decisions"]
        assert cst["resident_chunks"] <= cst["frames"]
    assert decisions >= snaps   # every snapshot took an admission decision
    eng.close()
