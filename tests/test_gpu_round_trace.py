"""Closure round trace: one record per round of every persistent launch.

The records of a stage must account for exactly the rounds the stage reports
(delete: invalidation + push closure; add: push closure), every launch must
end with an END record, round numbers and wall-clock stamps must not go
backwards within a launch, and tracing must not change any value (oracle
SSSP/BFS after every stage).  A capacity smaller than the rounds keeps the
first records and reports how many were written.
"""
import numpy as np
import pytest

import oracle
from tests.conftest import announce
from tests.helpers import rmat_graph

pytestmark = pytest.mark.gpu


def _launches(recs):
    """Split records into launches (each closed by its END record)."""
    import dycgraph_amd.engine as de
    out, cur = [], []
    for r in recs:
        cur.append(r)
        if r["tag"] & de.ROUND_END:
            out.append(cur)
            cur = []
    assert not cur, "records after the last END record"
    return out


def _rounds(recs, ctl):
    import dycgraph_amd.engine as de
    flags = de.ROUND_END | de.ROUND_ADVANCE
    return sum(1 for r in recs if (r["tag"] & 0xFF) == ctl and not (r["tag"] & flags))


@pytest.mark.parametrize("graphs", [1, 0])
def test_round_trace_accounts_for_every_round(graphs):
    import dycgraph_amd as d
    import dycgraph_amd.engine as de
    announce(f"round trace rmat12 graphs {graphs}")
    n, src, dst = rmat_graph(12, 1 << 16, seed=8)
    per = [500, 700]
    adds, dels = d.gen_stream(n, src, dst, per, per, seed=9)
    eng = d.Engine(d.EngineConfig(app="sssp", sources=(0, 5), query_apps=("sssp", "bfs"),
                                  weight_mode=d.W_REFERENCE))
    eng.set_tuning(graphs=bool(graphs))
    eng.load_graph(n, src, dst)
    eng.set_round_trace(1 << 14)
    eng.start()
    recs, written = eng.round_trace()
    assert written == len(recs) > 0
    for launch in _launches(recs):
        assert all(r["tag"] & de.ROUND_STATIC for r in launch)
        body = [r for r in launch if not (r["tag"] & de.ROUND_END)]
        assert body and body[0]["round"] == 0 and body[0]["frontier"] == 2   # one source per query
        assert all(b["round"] >= a["round"] for a, b in zip(body, body[1:]))
        assert all(b["t"] >= a["t"] for a, b in zip(launch, launch[1:]))
    eng.load_updates(adds, dels)
    store = oracle.Store(n, src, dst)
    off = 0
    for k in per:
        st = eng.del_edge(off, k)
        recs, written = eng.round_trace()
        assert written == len(recs)
        modes = {(r["tag"] >> 8) & 0xF for r in recs}
        assert modes <= {0, 1, 2} and 2 in modes
        assert not any(r["tag"] & de.ROUND_STATIC for r in recs)
        assert _rounds(recs, 1) + _rounds(recs, 3) == st["rounds"]
        assert len(_launches(recs)) == 3   # invalidation, pull pass, push closure
        st = eng.add_edge(off, k)
        recs, _ = eng.round_trace()
        assert _rounds(recs, 5) == st["rounds"]
        assert {(r["tag"] >> 8) & 0xF for r in recs} == {0}
        store.apply(dels[off:off + k], adds[off:off + k])
        rp, col = store.csr()
        assert np.array_equal(eng.gather_value(0), oracle.sssp(n, rp, col, 0, oracle.W_REFERENCE))
        assert np.array_equal(eng.gather_value(1), oracle.bfs(n, rp, col, 5))
        off += k


def test_round_trace_capacity_and_off():
    import dycgraph_amd as d
    announce("round trace capacity")
    n, src, dst = rmat_graph(10, 1 << 13, seed=2)
    eng = d.Engine(d.EngineConfig(app="bfs", sources=(1,)))
    eng.load_graph(n, src, dst)
    eng.set_round_trace(3)
    eng.start()
    recs, written = eng.round_trace()
    assert len(recs) == 3 and written > 3
    assert [int(r["round"]) for r in recs] == [0, 1, 2]
    eng.set_round_trace(0)
    with pytest.raises(d.DcgError):
        eng.round_trace()
