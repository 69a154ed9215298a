"""dcg_snapshot: both stages of a snapshot with one host wait.  Values of every
query equal the oracle's after each snapshot, rows equal the oracle's multiset
model, and the stage statistics equal those of an engine that runs dcg_del_edge
then dcg_add_edge on the same stream."""
import numpy as np
import pytest

import oracle
from tests.conftest import announce
from tests.helpers import cc_parents_ok, oracle_values, rmat_graph, same_rows

pytestmark = pytest.mark.gpu

APPS = ("sssp", "bfs", "cc", "sssp", "bfs")
SOURCES = (0, 0, 0, 17, 1234)


def _engine(placement, policy, graphs=True):
    import dycgraph_amd as d
    eng = d.Engine(d.EngineConfig(app="sssp", sources=SOURCES, query_apps=APPS, weight_mode=1, weight_seed=5,
                                  weight_max=50, placement=placement, cache_bytes=1 << 18 if placement else 0,
                                  cache_policy=policy))
    eng.set_tuning(graphs=graphs)
    return eng


# (placement, policy): HBM store; host store with a 1-frame cache under the
# hotness and the learned policy
@pytest.mark.parametrize("placement,policy", [(0, 1), (1, 1), (1, 2)])
def test_snapshot_step_matches_oracle_and_stage_calls(placement, policy):
    announce(f"snapshot step: sssp x2, bfs x2, cc on rmat13, placement={placement}, policy={policy}")
    import dycgraph_amd as d
    n, src, dst = rmat_graph(13, 90000, seed=21)
    # batch sizes on both sides of the one-workgroup sort (4,096 records), and
    # empty delete / add batches
    adds_per = [600, 5000, 0, 300, 7000]
    dels_per = [500, 6000, 800, 0, 4500]
    adds, dels = d.gen_stream(n, src, dst, adds_per, dels_per, seed=13)
    step = _engine(placement, policy)
    pair = _engine(placement, policy, graphs=False)
    for e in (step, pair):
        e.load_graph(n, src, dst)
        e.start()
        e.load_updates(adds, dels)
    store = oracle.Store(n, src, dst)
    ab = db = 0
    for s, (na, nd) in enumerate(zip(adds_per, dels_per)):
        if policy == d.CACHE_LEARNED:
            step.cache_prefetch(ab, na, db, nd)
            pair.cache_prefetch(ab, na, db, nd)
        dst_s, ast_s = step.snapshot(db, nd, ab, na)
        dst_p = pair.del_edge(db, nd)
        ast_p = pair.add_edge(ab, na)
        md = store.apply(dels[db:db + nd], [])
        ma = store.apply([], adds[ab:ab + na])
        assert dst_s["missing_deletes"] == md["missing_deletes"] == dst_p["missing_deletes"], s
        assert ast_s["invalid_additions"] == ma["invalid_additions"] == ast_p["invalid_additions"], s
        for k in ("requests", "touched_sources", "changed_sources"):
            assert dst_s[k] == dst_p[k], (s, k, dst_s[k], dst_p[k])
        for k in ("requests", "touched_sources", "relocated_rows"):
            assert ast_s[k] == ast_p[k], (s, k, ast_s[k], ast_p[k])
        rp, col = store.csr()
        erp, ecol = step.csr()
        assert same_rows(erp, ecol, store), f"rows differ after snapshot {s}"
        prp, pcol = pair.csr()
        assert same_rows(prp, pcol, store), f"rows of the stage-call engine differ after snapshot {s}"
        for q, (app, so) in enumerate(zip(APPS, SOURCES)):
            want = oracle_values(app, n, rp, col, so, oracle.W_SEEDED, 5, 50)
            got = step.gather_value(q)
            assert np.array_equal(got, want), (s, q, app, int(np.sum(got != want)))
            assert np.array_equal(got, pair.gather_value(q)), (s, q)
            if app == "cc":
                assert cc_parents_ok(n, rp, col, got, step.gather_parent(q)), (s, q)
            else:
                chk = oracle.check_sssp(n, rp, col, so, got, step.gather_parent(q),
                                        oracle.W_SEEDED if app == "sssp" else oracle.W_UNIT, 5, 50)
                assert chk["invalid_parents"] == 0 and chk["relaxable_edges"] == 0, (s, q, chk)
        step.cache_maintain()
        pair.cache_maintain()
        ab += na
        db += nd
    step.close()
    pair.close()


def test_snapshot_step_argument_checks():
    announce("snapshot step: state and range errors")
    import dycgraph_amd as d
    n, src, dst = rmat_graph(10, 4000, seed=2)
    adds, dels = d.gen_stream(n, src, dst, [100], [100], seed=3)
    eng = _engine(0, 1)
    eng.load_graph(n, src, dst)
    with pytest.raises(d.DcgError):
        eng.snapshot(0, 0, 0, 0)   # before dcg_start
    eng.start()
    eng.load_updates(adds, dels)
    with pytest.raises(d.DcgError):
        eng.snapshot(0, 101, 0, 100)   # past the staged deletions
    with pytest.raises(d.DcgError):
        eng.snapshot(0, 100, 50, 100)   # past the staged additions
    dst_s, ast_s = eng.snapshot(0, 100, 0, 100)
    assert dst_s["requests"] == 100 and ast_s["requests"] == 100
    eng.close()
