"""The engine on the reference's own regression scenarios, against the vectors
its Python test oracles produce (tests/golden/reference_vectors.json, written by
tests/golden/make_reference_vectors.py from tests/bfs_dynamic_smoke.py and
tests/cc_dynamic_smoke.py): values and value checksums after every delete and
add stage, with the reference's (u+v)%128+1 weights from vertex 0."""
import numpy as np
import pytest

import oracle
from tests.conftest import announce
from tests.test_oracle_golden import reference_vectors, unpack_u32

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("placement", [0, 1])
@pytest.mark.parametrize("app", ["bfs", "sssp"])
def test_bfs_sssp_kat_stages(app, placement):
    import dycgraph_amd as d
    announce(f"reference BFS/SSSP KAT vectors, app={app}, placement={placement}")
    doc = reference_vectors()["bfs_sssp"]
    n = doc["n"]
    e = np.array(doc["initial_edges"], np.uint32)
    stages = doc["stages"]
    dels = [x for st in stages for x in st["dels"]]
    adds = [x for st in stages for x in st["adds"]]
    eng = d.Engine(d.EngineConfig(app=app, sources=(0,), weight_mode=d.W_REFERENCE, placement=placement,
                                  cache_bytes=1 << 18))
    eng.load_graph(n, e[:, 0], e[:, 1])
    eng.start()
    eng.load_updates(adds, dels)
    ab = db = 0
    for st in stages:
        if st["dels"]:
            eng.del_edge(db, len(st["dels"]))
            db += len(st["dels"])
        if st["adds"]:
            eng.add_edge(ab, len(st["adds"]))
            ab += len(st["adds"])
        got = eng.gather_value()
        assert np.array_equal(got, unpack_u32(st[app]["values"])), st["stage"]
        assert oracle.checksum(got) == st[app]["checksum"], st["stage"]
    eng.close()


@pytest.mark.parametrize("placement", [0, 1])
def test_cc_kat_batches(placement):
    import dycgraph_amd as d
    announce(f"reference CC KAT vectors (20 batches), placement={placement}")
    doc = reference_vectors()["cc"]
    n = doc["n"]
    base = np.array(doc["base"], np.uint32)
    batches = doc["batches"]
    dels = [x for b in batches for x in b["dels"]]
    adds = [x for b in batches for x in b["adds"]]
    eng = d.Engine(d.EngineConfig(app="cc", placement=placement, cache_bytes=1 << 18))
    eng.load_graph(n, base[:, 0], base[:, 1])
    eng.start()
    got = eng.gather_value()
    assert np.array_equal(got, unpack_u32(doc["initial"]["labels"]))
    eng.load_updates(adds, dels)
    ab = db = 0
    for i, b in enumerate(batches):
        eng.del_edge(db, len(b["dels"]))
        db += len(b["dels"])
        got = eng.gather_value()
        assert np.array_equal(got, unpack_u32(b["after_delete"]["labels"])), (i, "delete")
        assert oracle.checksum(got) == b["after_delete"]["checksum"], (i, "delete")
        eng.add_edge(ab, len(b["adds"]))
        ab += len(b["adds"])
        got = eng.gather_value()
        assert np.array_equal(got, unpack_u32(b["after_add"]["labels"])), (i, "add")
        assert oracle.checksum(got) == b["after_add"]["checksum"], (i, "add")
    eng.close()
