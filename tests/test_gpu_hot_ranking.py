"""The GPU radix sort and the reference's vertex hotness ranking.

dcg_radix_sort is checked against numpy's stable sort (keys by a bit range,
values carried: a stable sort keeps equal digits in input order).  The engine's
dcg_hot_candidates is checked against the oracle's restatement of
compute_hot_vertices_sssp + CacheCandidateCount (driver.cuh:486-579,
framework.cuh:1321-1350) on the engine's own window, unreached mask and live
degrees, then the window shift (reset_hotness) is checked word by word.
"""
import numpy as np
import pytest

import oracle
from tests.conftest import announce
from tests.helpers import rmat_graph

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 777, 4095, 4096, 4097, 100_003, 1 << 21])
@pytest.mark.parametrize("bits", [(0, 64), (0, 20), (3, 17), (60, 64), (9, 9)])
def test_radix_sort_matches_stable_sort(n, bits):
    import dycgraph_amd as d
    announce(f"radix sort n={n} bits={bits}")
    rng = np.random.default_rng(n * 7 + bits[0])
    keys = rng.integers(0, 1 << 63, n, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, n, dtype=np.uint64)
    if n > 100:
        keys[rng.integers(0, n, n // 3)] = keys[0]   # many equal keys
    vals = np.arange(n, dtype=np.uint32)
    b0, b1 = bits
    width = b1 - b0
    field = (keys >> np.uint64(b0)) & np.uint64((1 << width) - 1) if width < 64 else keys
    order = np.argsort(field, kind="stable")
    k2, v2 = d.radix_sort(keys, vals, b0, b1)
    assert np.array_equal(v2, vals[order])
    assert np.array_equal(k2, keys[order])
    k3 = d.radix_sort(keys, None, b0, b1)
    assert np.array_equal(k3, keys[order])


def _ranking_case(eng, n, cap, query, masked):
    """Compare one dcg_hot_candidates call with the oracle, then the shift."""
    win = eng.hot_window()
    rp, _ = eng.csr()
    deg = np.diff(rp).astype(np.uint32)
    value = eng.gather_value(query).view(np.uint32) if masked else None
    want_k, want_ids = oracle.hot_candidates(win, deg, cap, value)
    k, ids = eng.hot_candidates(cap, query)
    assert k == want_k, (cap, query, k, want_k)
    assert np.array_equal(ids, want_ids[:k]), (cap, query, int(np.sum(ids != want_ids[:k])))
    assert np.array_equal(eng.hot_window(), (win << np.uint32(8)).astype(np.uint32))
    return win


@pytest.mark.parametrize("placement", ["device", "host"])
def test_hot_candidates_match_oracle(placement):
    """RMAT graph, SSSP + BFS queries, 3 snapshots of 4000 deletes and inserts;
    after each snapshot one ranking at a cap inside the total, and at the end
    the cap edge cases (0, 1, total, total + 1) and a short id buffer."""
    import dycgraph_amd as d
    announce(f"hot candidates {placement}")
    n, src, dst = rmat_graph(16, 1 << 20, seed=4)
    per, snaps = 4000, 3
    adds, dels = d.gen_stream(n, src, dst, [per] * snaps, [per] * snaps, seed=9)
    rng = np.random.default_rng(5)
    sources = [int(src[i]) for i in rng.integers(0, len(src), 2)]
    place = d.PLACE_DEVICE if placement == "device" else d.PLACE_HOST
    eng = d.Engine(d.EngineConfig(app="sssp", sources=sources, query_apps=("sssp", "bfs"), weight_mode=d.W_SEEDED,
                                  weight_max=100, weight_seed=3, placement=place))
    eng.load_graph(n, src, dst)
    eng.start()
    eng.load_updates(adds, dels)
    seen = 0
    for s in range(snaps):
        b = s * per
        eng.del_edge(b, per)
        eng.add_edge(b, per)
        total = int(np.diff(eng.csr()[0]).sum())
        win = _ranking_case(eng, n, total // (3 + s), s % 2, True)
        seen += int(np.count_nonzero(win))
    assert seen > 0   # the traversals counted row expansions
    total = int(np.diff(eng.csr()[0]).sum())
    for cap in (0, 1, total, total + 1):
        _ranking_case(eng, n, cap, 0, True)
    # a short id buffer: the count is the full count, the ids its prefix
    win = eng.hot_window()
    deg = np.diff(eng.csr()[0]).astype(np.uint32)
    want_k, want_ids = oracle.hot_candidates(win, deg, total // 2, eng.gather_value(1).view(np.uint32))
    k, ids = eng.hot_candidates(total // 2, 1, limit=10)
    assert k == want_k and np.array_equal(ids, want_ids[:min(k, 10)])
    eng.close()


def test_hot_candidates_pagerank_unmasked():
    """PageRank ranks every vertex (no unreached mask)."""
    import dycgraph_amd as d
    announce("hot candidates pagerank")
    n, src, dst = rmat_graph(14, 1 << 18, seed=6)
    adds, dels = d.gen_stream(n, src, dst, [2000], [2000], seed=2)
    eng = d.Engine(d.EngineConfig(app="pr"))
    eng.load_graph(n, src, dst)
    eng.start()
    eng.load_updates(adds, dels)
    eng.update_pagerank(0, 2000, 0, 2000)
    total = int(np.diff(eng.csr()[0]).sum())
    win = _ranking_case(eng, n, total // 4, 0, False)
    assert np.count_nonzero(win) > 0
    eng.close()
