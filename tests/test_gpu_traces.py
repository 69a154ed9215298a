"""Diagnostic traces written by the engine, read back and checked.

Repair snapshot: the affected set, in-rows and states of the first delete
stage; replaying it offline (Dijkstra from the unaffected in-neighbours'
values, src/i16_repair_oracle.cpp:13-99) must reproduce the engine's repaired
distances, which must equal the oracle's from-scratch Dijkstra; every vertex
whose distance changed must be in the affected set; no state outside it may
change.  Cache-candidate trace: every record must equal the oracle's ranking
and admission over the engine's own window and degrees.
"""
import numpy as np
import pytest

import oracle
from dycgraph_amd import traces
from tests.conftest import announce
from tests.helpers import rmat_graph

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("placement", [0, 1])
def test_repair_snapshot_and_cache_trace(tmp_path, placement):
    import dycgraph_amd as d
    announce(f"traces rmat12 placement {placement}")
    n, src, dst = rmat_graph(12, 1 << 16, seed=3)
    per = [800, 600, 700]
    adds, dels = d.gen_stream(n, src, dst, per, per, seed=4)
    eng = d.Engine(d.EngineConfig(app="sssp", sources=(0,), weight_mode=d.W_REFERENCE, placement=placement,
                                  cache_bytes=2 << 18))
    eng.load_graph(n, src, dst)
    eng.start()
    eng.load_updates(adds, dels)
    store = oracle.Store(n, src, dst)
    snap, f1 = str(tmp_path / "i16.bin"), str(tmp_path / "f1.bin")
    eng.set_repair_snapshot(snap, 0)
    eng.set_cache_trace(f1)
    cap = int(np.diff(eng.csr()[0]).sum()) // 3
    want_records = []

    def rank(batch):
        win = eng.hot_window()
        deg = np.diff(eng.csr()[0]).astype(np.uint32)
        k, ids = oracle.hot_candidates(win, deg, cap, eng.gather_value(0))
        want_records.append((batch, ids[:k], deg[ids[:k]]))
        eng.hot_candidates(cap, 0, limit=0)
        eng.trace_cache_candidates(batch)

    rank(0xFFFFFFFF)
    values0 = eng.gather_value(0)
    off = 0
    for s, k in enumerate(per):
        eng.del_edge(off, k)
        store.apply(dels[off:off + k], [])
        if s == 0:
            rp, col = store.csr()
            want = oracle.sssp(n, rp, col, 0, oracle.W_REFERENCE)
            got = eng.gather_value(0)
            assert np.array_equal(got, want)
            changed = set(np.flatnonzero(got != values0).tolist())
        eng.add_edge(off, k)
        store.apply([], adds[off:off + k])
        rank(s)
        off += k
    eng.close()

    sn = traces.read_repair_snapshot(snap)
    assert sn["weight_rule"] == 1 and sn["source"] == 0 and sn["nodes"] == n and len(sn["affected"]) > 0
    aff = set(sn["affected"].tolist())
    assert changed <= aff   # every vertex the stage changed was in the affected set
    before = {int(x["vertex"]): int(x["value"]) for x in sn["before"]}
    after = {int(x["vertex"]): int(x["value"]) for x in sn["after"]}
    assert all(before[v] == traces.INF for v in aff)
    assert all(after[v] == want[v] for v in aff)
    r = traces.replay_repair_snapshot(sn)
    assert r["mismatches"] == 0 and r["external_changes"] == 0, r
    # in-rows as recorded = the store's live in-rows after the deletions
    rp, col = store.csr()   # (final store; rebuild the post-delete one for the check)
    st0 = oracle.Store(n, src, dst)
    st0.apply(dels[:per[0]], [])
    rp0, col0 = st0.csr()
    inn = [[] for _ in range(n)]
    for u in range(n):
        for v in col0[rp0[u]:rp0[u + 1]]:
            inn[int(v)].append(u)
    for i, v in enumerate(sn["affected"].tolist()):
        got = sorted(sn["incoming"][sn["offsets"][i]:sn["offsets"][i + 1]].tolist())
        assert got == sorted(inn[v]), (v, got[:8], sorted(inn[v])[:8])

    recs = traces.read_cache_trace(f1)
    assert len(recs) == len(want_records)
    for (b, c, adm), (wb, wids, wdeg) in zip(recs, want_records):
        assert b == wb and c == cap
        assert np.array_equal(adm["vertex"], wids) and np.array_equal(adm["degree"], wdeg), b
