"""Closures on a single workgroup: every round's new frontier and far-pile
entries exceed the workgroup's LDS output buffers (8,192 and 4,096 entries),
so the appends go through the waves' LDS staging runs, and bucket advances
move the far piles through them too.  Values of every query equal the
oracle's after the static convergence and after every stage."""
import numpy as np
import pytest

import oracle
from tests.conftest import announce
from tests.helpers import cc_parents_ok, oracle_values, rmat_graph, same_rows

pytestmark = pytest.mark.gpu

APPS = ("sssp", "bfs", "cc", "sssp")
SOURCES = (0, 3, 0, 77)
OUT_CAP = 8192   # a workgroup's LDS frontier buffer (vertex_kernels.hpp kOutCap)


@pytest.mark.parametrize("quarters", [0, 4])
def test_single_workgroup_closures_stage_overflow(quarters):
    announce(f"closures on one workgroup (staging runs), rmat16, buckets={quarters}")
    import dycgraph_amd as d
    n, src, dst = rmat_graph(16, 600000, seed=12)
    adds, dels = d.gen_stream(n, src, dst, [30000, 20000], [25000, 30000], seed=8)
    eng = d.Engine(d.EngineConfig(app="sssp", sources=SOURCES, query_apps=APPS, weight_mode=1, weight_seed=7,
                                  weight_max=100))
    eng.set_tuning(bucket_quarters=quarters)
    eng.set_closure_grid(16, 0)   # one workgroup, no solo stretches
    eng.set_round_trace(1 << 16)
    eng.load_graph(n, src, dst)
    eng.start()
    trace, _ = eng.round_trace()
    # with one workgroup, a frontier of a later round (round 0's comes from the
    # host) above the LDS buffer was appended through the staging runs
    emitted = trace["frontier"][trace["round"] > 0]
    assert emitted.size and int(emitted.max()) > OUT_CAP, int(emitted.max()) if emitted.size else 0
    store = oracle.Store(n, src, dst)

    def check(stage):
        rp, col = store.csr()
        for q, (app, s) in enumerate(zip(APPS, SOURCES)):
            want = oracle_values(app, n, rp, col, s, oracle.W_SEEDED, 7, 100)
            got = eng.gather_value(q)
            assert np.array_equal(got, want), (stage, q, app, int(np.sum(got != want)))
            if app == "cc":
                assert cc_parents_ok(n, rp, col, got, eng.gather_parent(q)), (stage, q)
            else:
                chk = oracle.check_sssp(n, rp, col, s, got, eng.gather_parent(q),
                                        oracle.W_SEEDED if app == "sssp" else oracle.W_UNIT, 7, 100)
                assert chk["invalid_parents"] == 0 and chk["relaxable_edges"] == 0, (stage, q, chk)

    check("static")
    eng.load_updates(adds, dels)
    ab = db = 0
    for s, (na, nd) in enumerate(zip([30000, 20000], [25000, 30000])):
        eng.del_edge(db, nd)
        store.apply(dels[db:db + nd], [])
        check(f"delete {s}")
        eng.add_edge(ab, na)
        store.apply([], adds[ab:ab + na])
        erp, ecol = eng.csr()
        assert same_rows(erp, ecol, store), f"rows after snapshot {s}"
        check(f"add {s}")
        ab += na
        db += nd
    eng.close()
