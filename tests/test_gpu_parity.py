"""GPU parity: every snapshot phase bit-exact against the CPU oracle.

SSSP/BFS/CC values must equal a from-scratch Dijkstra / FIFO BFS / flood fill
after every delete stage and every add stage; the engine's rows must equal the
multiset model's rows (first-occurrence deletes, ordered appends); parents must
be tight witnesses (samples/hybrid_sssp/hybrid_sssp.cu:228-333).
"""
import numpy as np
import pytest

import oracle
from tests.conftest import announce
from tests.helpers import cc_parents_ok, oracle_values, rmat_graph, same_rows

pytestmark = pytest.mark.gpu


def _engine(**kw):
    import dycgraph_amd as d
    return d.Engine(d.EngineConfig(**kw))


def _run_stream(app, n, src, dst, adds_per, dels_per, placement, cache_bytes=0, source=0, wmode=0, seed=1,
                wmax=100, check_rows=True, stream_seed=7, stream=None, policy=1, prefetch=False):
    """Apply the stream snapshot by snapshot and compare values, rows and parents
    with the oracle after every stage.  policy 2 (learned) runs the policy's
    prefetch before each snapshot, as the benchmark does; prefetch=True copies
    the next snapshot's records ahead on the copy stream."""
    import dycgraph_amd as d
    adds, dels = stream if stream is not None else d.gen_stream(n, src, dst, adds_per, dels_per, seed=stream_seed)
    eng = _engine(app=app, sources=(source,), weight_mode=wmode, weight_seed=seed, weight_max=wmax,
                  placement=placement, cache_bytes=cache_bytes, cache_policy=policy)
    eng.load_graph(n, src, dst)
    eng.start()
    store = oracle.Store(n, src, dst)
    rp, col = store.csr()
    want = oracle_values(app, n, rp, col, source, wmode, seed, wmax)
    got = eng.gather_value()
    assert np.array_equal(got, want), f"{app} static mismatch: {np.sum(got != want)} vertices"
    eng.load_updates(adds, dels)
    ab = db = 0
    add_stats = []
    for s, (na, nd) in enumerate(zip(adds_per, dels_per)):
        if prefetch and s + 1 < len(adds_per):
            eng.prefetch_updates(ab + na, adds_per[s + 1], db + nd, dels_per[s + 1])
        if policy == d.CACHE_LEARNED:
            eng.cache_prefetch(ab, na, db, nd)
        st = eng.del_edge(db, nd)
        m = store.apply(dels[db:db + nd], [])
        assert st["missing_deletes"] == m["missing_deletes"]
        rp, col = store.csr()
        if check_rows:
            erp, ecol = eng.csr()
            assert same_rows(erp, ecol, store), f"rows differ after delete stage {s}"
        want = oracle_values(app, n, rp, col, source, wmode, seed, wmax)
        got = eng.gather_value()
        assert np.array_equal(got, want), f"{app} delete stage {s}: {np.sum(got != want)} vertices differ"
        if app != "cc":
            chk = oracle.check_sssp(n, rp, col, source, got, eng.gather_parent(),
                                    wmode if app == "sssp" else oracle.W_UNIT, seed, wmax)
            assert chk["invalid_parents"] == 0 and chk["relaxable_edges"] == 0, chk
        else:
            assert cc_parents_ok(n, rp, col, got, eng.gather_parent()), f"CC parents, delete stage {s}"
        st = eng.add_edge(ab, na)
        m = store.apply([], adds[ab:ab + na])
        assert st["invalid_additions"] == m["invalid_additions"]
        rp, col = store.csr()
        if check_rows:
            erp, ecol = eng.csr()
            assert same_rows(erp, ecol, store), f"rows differ after add stage {s}"
        want = oracle_values(app, n, rp, col, source, wmode, seed, wmax)
        got = eng.gather_value()
        assert np.array_equal(got, want), f"{app} add stage {s}: {np.sum(got != want)} vertices differ"
        if app != "cc":
            chk = oracle.check_sssp(n, rp, col, source, got, eng.gather_parent(),
                                    wmode if app == "sssp" else oracle.W_UNIT, seed, wmax)
            assert chk["invalid_parents"] == 0 and chk["relaxable_edges"] == 0, chk
        else:
            assert cc_parents_ok(n, rp, col, got, eng.gather_parent()), f"CC parents, add stage {s}"
        eng.cache_maintain()
        add_stats.append(st)
        ab += na
        db += nd
    eng.close()
    return add_stats


@pytest.mark.parametrize("app", ["sssp", "bfs", "cc"])
def test_rmat12_device(app):
    announce(f"rmat12 device {app}")
    n, src, dst = rmat_graph(12, 40000)
    _run_stream(app, n, src, dst, [300] * 4, [300] * 4, placement=0)


@pytest.mark.parametrize("policy", [1, 2])
@pytest.mark.parametrize("app", ["sssp", "bfs", "cc"])
def test_rmat12_host_cache(app, policy):
    """Host placement with a one-frame cache under the hotness policy (1) and the
    learned policy (2, with its prefetch before every snapshot)."""
    announce(f"rmat12 host-cache {app} policy={policy}")
    n, src, dst = rmat_graph(12, 40000)
    # ~40k out + 40k in slots: 1 frame of 65,536 slots leaves most chunks zero-copy
    _run_stream(app, n, src, dst, [300] * 6, [300] * 6, placement=1, cache_bytes=1 << 18, policy=policy)


@pytest.mark.parametrize("app", ["sssp", "cc"])
def test_prefetched_batches(app):
    """Each snapshot's records copied ahead on the copy stream while the previous
    snapshot runs (two device slots per direction, reused in turn); batch sizes
    change between snapshots, so slots are refilled and regrown."""
    announce(f"prefetched batches {app}")
    n, src, dst = rmat_graph(12, 40000, seed=3)
    sizes = [300, 900, 50, 2000, 300]
    _run_stream(app, n, src, dst, sizes, sizes[::-1], placement=1, cache_bytes=1 << 18, prefetch=True)


@pytest.mark.parametrize("app", ["sssp", "cc"])
def test_rmat14_large_batches(app):
    """Batches of 20,000 deletes and inserts: group heads come from the
    device-wide ordered compaction (small batches use one workgroup)."""
    announce(f"rmat14 large batches {app}")
    n, src, dst = rmat_graph(14, 200000, seed=6)
    _run_stream(app, n, src, dst, [20000] * 2, [20000] * 2, placement=1, cache_bytes=1 << 20)


@pytest.mark.parametrize("app", ["sssp", "cc"])
def test_mixed_batch_sizes_out_of_range(app):
    """A large batch (device-wide sorts, thousands of relocated rows) followed by
    small ones (one-workgroup LDS path) whose free pool of retired blocks is
    larger than one workgroup sorts, then a mid-size one; about 1% of the
    records have an id out of range (additions: source; deletions: either end)
    and count as invalid additions / missing deletes without touching a row."""
    announce(f"mixed batch sizes with out-of-range records {app}")
    import dycgraph_amd as d
    n, src, dst = rmat_graph(14, 100000, seed=9)
    sizes = [30000, 700, 5000, 900]
    adds, dels = d.gen_stream(n, src, dst, sizes, sizes, seed=13)
    rng = np.random.default_rng(21)
    ia = rng.choice(len(adds), len(adds) // 100, replace=False)
    adds[ia, 0] = n + rng.integers(0, 1000, len(ia)).astype(np.uint32)
    idl = rng.choice(len(dels), len(dels) // 100, replace=False)
    half = len(idl) // 2
    dels[idl[:half], 0] = n + rng.integers(0, 1000, half).astype(np.uint32)
    dels[idl[half:], 1] = n + rng.integers(0, 1000, len(idl) - half).astype(np.uint32)
    st = _run_stream(app, n, src, dst, sizes, sizes, placement=1, cache_bytes=1 << 20, stream=(adds, dels))
    # the retired blocks of the first batch form the second batch's free pool
    assert st[0]["relocated_rows"] > 4096, st[0]["relocated_rows"]


def test_rmat16_sssp_seeded_weights():
    announce("rmat16 sssp seeded weights, 3 snapshots x 1000/1000")
    n, src, dst = rmat_graph(16, 1 << 20)
    _run_stream("sssp", n, src, dst, [1000] * 3, [1000] * 3, placement=0, wmode=1, seed=11, wmax=100,
                check_rows=False)


@pytest.mark.parametrize("dups", [False, True])
@pytest.mark.parametrize("placement", [0, 1])
def test_large_row_deletes_exact_rows(dups, placement):
    """Rows above 1,024 entries take the workgroup delete path; with duplicate
    targets (more occurrences than requests) the exact first-occurrence path."""
    announce(f"large-row deletes dups={dups} placement={placement}")
    import dycgraph_amd as d
    n = 30000
    rng = np.random.default_rng(5)
    hub = list(range(1, 20001))
    if dups:
        hub += [5, 9, 5, 77]                     # value 5 three times, 9 and 77 twice
    rng.shuffle(hub)
    src = [0] * len(hub) + list(range(1, 1000))
    dst = hub + [int(x) for x in rng.integers(1, n, 999)]
    # the in-row of vertex 1 is large too: many sources point at 1
    src += list(range(100, 12100))
    dst += [1] * 12000
    src, dst = np.array(src, np.uint32), np.array(dst, np.uint32)
    dels = [(0, int(v)) for v in rng.choice(np.arange(10, 20000), 300, replace=False)]
    dels += [(0, 5), (0, 9), (0, 9), (0, 9), (0, 123456 % n)] + [(u, 1) for u in range(100, 400)]
    adds = [(0, 5), (3, 4)]
    eng = _engine(app="bfs", placement=placement, cache_bytes=1 << 18)
    eng.load_graph(n, src, dst)
    eng.start()
    eng.load_updates(adds, dels)
    store = oracle.Store(n, src, dst)
    st = eng.del_edge(0, len(dels))
    m = store.apply(dels, [])
    assert st["missing_deletes"] == m["missing_deletes"]
    erp, ecol = eng.csr()
    assert same_rows(erp, ecol, store)
    rp, col = store.csr()
    assert np.array_equal(eng.gather_value(), oracle.bfs(n, rp, col, 0))
    eng.add_edge(0, len(adds))
    store.apply([], adds)
    erp, ecol = eng.csr()
    assert same_rows(erp, ecol, store)
    eng.close()


@pytest.mark.parametrize("policy", [1, 2])
@pytest.mark.parametrize("graphs", [True, False])
def test_mixed_programs_one_engine(graphs, policy):
    """SSSP, BFS and CC queries (several sources) repaired together over one chunk
    store: every query bit-exact after every stage, with each stage's repair
    launches replayed from a captured graph or issued directly, under the
    hotness (1) and the learned (2) cache policy."""
    announce(f"mixed programs: sssp x2, bfs x2, cc on rmat12 host-cache, graphs={graphs}, policy={policy}")
    import dycgraph_amd as d
    n, src, dst = rmat_graph(12, 40000, seed=9)
    adds, dels = d.gen_stream(n, src, dst, [400] * 3, [400] * 3, seed=3)
    apps = ("sssp", "bfs", "cc", "sssp", "bfs")
    sources = (0, 0, 0, 17, 1234)
    eng = _engine(app="sssp", sources=sources, query_apps=apps, weight_mode=1, weight_seed=5, weight_max=50,
                  placement=1, cache_bytes=1 << 18, cache_policy=policy)
    eng.set_tuning(graphs=graphs)
    eng.load_graph(n, src, dst)
    eng.start()
    eng.load_updates(adds, dels)
    store = oracle.Store(n, src, dst)

    def check(stage):
        rp, col = store.csr()
        for q, (app, s) in enumerate(zip(apps, sources)):
            want = oracle_values(app, n, rp, col, s, oracle.W_SEEDED, 5, 50)
            got = eng.gather_value(q)
            assert np.array_equal(got, want), (stage, q, app, int(np.sum(got != want)))

    check("static")
    for s in range(3):
        if policy == d.CACHE_LEARNED:
            eng.cache_prefetch(400 * s, 400, 400 * s, 400)
        eng.del_edge(400 * s, 400)
        store.apply(dels[400 * s:400 * (s + 1)], [])
        check(f"delete {s}")
        eng.add_edge(400 * s, 400)
        store.apply([], adds[400 * s:400 * (s + 1)])
        check(f"add {s}")
        eng.cache_maintain()
    eng.close()


def test_cc_queries_at_unaligned_ranges():
    """Three CC queries between SSSP / BFS ones on 3,001 vertices: each CC
    query's slot range starts off a 1,024-slot boundary and blocks straddle two
    queries, so the parent check's block marks are shared between queries.
    Labels bit-exact and parents rooted witnesses after every stage."""
    announce("cc queries at unaligned slot ranges, n=3001")
    import dycgraph_amd as d
    _, src, dst = rmat_graph(12, 40000, seed=4)
    n = 3001
    keep = (src < n) & (dst < n)
    src, dst = src[keep], dst[keep]
    adds, dels = d.gen_stream(n, src, dst, [300] * 3, [300] * 3, seed=8)
    apps = ("cc", "sssp", "cc", "bfs", "cc")
    sources = (0, 5, 0, 7, 0)
    eng = _engine(app="sssp", sources=sources, query_apps=apps, weight_mode=1, weight_seed=2, weight_max=40)
    eng.load_graph(n, src, dst)
    eng.start()
    eng.load_updates(adds, dels)
    store = oracle.Store(n, src, dst)

    def check(stage):
        rp, col = store.csr()
        for q, (app, s) in enumerate(zip(apps, sources)):
            want = oracle_values(app, n, rp, col, s, oracle.W_SEEDED, 2, 40)
            got = eng.gather_value(q)
            assert np.array_equal(got, want), (stage, q, app, int(np.sum(got != want)))
            if app == "cc":
                assert cc_parents_ok(n, rp, col, got, eng.gather_parent(q)), (stage, q)

    check("static")
    for s in range(3):
        eng.del_edge(300 * s, 300)
        store.apply(dels[300 * s:300 * (s + 1)], [])
        check(f"delete {s}")
        eng.add_edge(300 * s, 300)
        store.apply([], adds[300 * s:300 * (s + 1)])
        check(f"add {s}")
    eng.close()


def test_autotune_keeps_values_and_picks_candidates():
    announce("autotune rmat14 sssp")
    n, src, dst = rmat_graph(14, 200000, seed=2)
    eng = _engine(app="sssp", weight_mode=1, weight_seed=3, weight_max=100)
    eng.load_graph(n, src, dst)
    eng.start()
    before = eng.gather_value()
    res = eng.autotune()
    assert res["blocks_per_cu"] in (1, 2, 3, 4) and res["big_row"] in (64, 128, 256, 512, 1024, 4096)
    assert res["piece"] in (256, 512, 1024, 2048, 4096) and res["evaluations"] >= 10
    assert np.array_equal(eng.gather_value(), before)
    rp, col = oracle.csr_from_edges(n, src, dst)
    assert np.array_equal(before, oracle.sssp(n, rp, col, 0, oracle.W_SEEDED, 3, 100))
    # an odd but legal setting still gives exact results
    import dycgraph_amd as d
    adds, dels = d.gen_stream(n, src, dst, [2000], [2000], seed=1)
    eng.set_tuning(1, 64, 256)
    eng.load_updates(adds, dels)
    eng.del_edge(0, 2000)
    eng.add_edge(0, 2000)
    store = oracle.Store(n, src, dst)
    store.apply(dels, adds)
    rp, col = store.csr()
    assert np.array_equal(eng.gather_value(), oracle.sssp(n, rp, col, 0, oracle.W_SEEDED, 3, 100))
    eng.close()


@pytest.mark.parametrize("placement", [0, 1])
def test_large_row_relocation_exact_rows(placement):
    """A row above 4,096 entries that outgrows its block moves on the workgroup path."""
    announce(f"large-row relocation placement={placement}")
    n = 20000
    hub = list(range(1, 8193))                       # exactly 8,192 entries: cap 8,192
    src = np.array([0] * len(hub) + list(range(10, 8202)), np.uint32)
    dst = np.array(hub + [5] * 8192, np.uint32)      # in-row of 5 holds 8,192 entries too
    adds = [(0, 9000), (0, 9001), (0, 1), (9, 5), (8300, 5)]
    dels = [(0, 77), (10, 5)]
    eng = _engine(app="bfs", placement=placement, cache_bytes=1 << 18)
    eng.load_graph(n, src, dst)
    eng.start()
    eng.load_updates(adds, dels)
    store = oracle.Store(n, src, dst)
    st = eng.add_edge(0, len(adds))
    store.apply([], adds)
    assert st["relocated_rows"] >= 1
    erp, ecol = eng.csr()
    assert same_rows(erp, ecol, store)
    eng.del_edge(0, len(dels))
    store.apply(dels, [])
    erp, ecol = eng.csr()
    assert same_rows(erp, ecol, store)
    rp, col = store.csr()
    assert np.array_equal(eng.gather_value(), oracle.bfs(n, rp, col, 0))
    eng.close()


@pytest.mark.parametrize("placement", [0, 1])
def test_sorted_large_rows_tombstones(placement):
    """Rows whose sorted prefix holds at least 1,024 entries delete by binary
    search and tombstones: first occurrences across the prefix and the appended
    tail, requests beyond the occurrences, tombstones carried through a
    relocation, and the compaction once tombstones pass a quarter of the row.
    Rows (tombstones skipped) and BFS levels must match the multiset model after
    every stage."""
    announce(f"sorted large rows with tombstones placement={placement}")
    import dycgraph_amd as d
    n = 60000
    hub = sorted(list(range(1, 20001)) + [5, 5, 9, 20000])    # 5 x3, 9 x2, 20000 x2, sorted
    src = [0] * len(hub) + list(range(100, 20100)) + list(range(20100, 20600))
    dst = hub + [1] * 20000 + [int(x) for x in np.random.default_rng(3).integers(2, n, 500)]
    src, dst = np.array(src, np.uint32), np.array(dst, np.uint32)
    rng = np.random.default_rng(9)
    stages = []
    # 1: prefix deletes (duplicates, over-requests, a missing value), tail appends incl. repeats
    stages.append(([(0, 5), (0, 5), (0, 9), (0, 9), (0, 9), (0, 123), (0, 50001)] +
                   [(int(u), 1) for u in rng.choice(np.arange(100, 20100), 400, replace=False)],
                   [(0, 5), (0, 123), (0, 40000), (0, 40000), (0, 9), (150, 1), (40000, 1)]))
    # 2: deletes reaching past the prefix into the tail, tail-only values
    stages.append(([(0, 5), (0, 5), (0, 40000), (0, 9), (0, 9), (0, 123), (0, 123), (40000, 1), (150, 1)],
                   [(0, 41000 + i) for i in range(300)]))
    # 3: relocation of the tombstoned row (cap 32,768 overflows)
    stages.append(([(0, int(v)) for v in rng.choice(np.arange(10, 20000), 200, replace=False)],
                   [(0, 45000 + i) for i in range(13000)]))
    # 4: enough deletes to pass a quarter tombstones -> compaction; in-row too
    big = [int(v) for v in rng.choice(np.arange(10, 20000), 9000, replace=False)]
    stages.append(([(0, v) for v in big] + [(int(u), 1) for u in range(100, 8000)],
                   [(0, 7), (0, 7), (3, 1)]))
    # 5: after the compaction: prefix (recomputed length) and tail deletes
    stages.append(([(0, 7), (0, 45001), (0, 20000), (0, 20000), (0, 20000)] +
                   [(0, v) for v in big[:50]], [(0, 2), (0, 1)]))
    adds = [e for _, a in stages for e in a]
    dels = [e for dl, _ in stages for e in dl]
    eng = _engine(app="bfs", placement=placement, cache_bytes=1 << 18)
    eng.load_graph(n, src, dst)
    eng.start()
    eng.load_updates(adds, dels)
    store = oracle.Store(n, src, dst)
    ab = db = 0
    for s, (dl, al) in enumerate(stages):
        st = eng.del_edge(db, len(dl))
        m = store.apply(dl, [])
        assert st["missing_deletes"] == m["missing_deletes"], f"stage {s}"
        erp, ecol = eng.csr()
        assert same_rows(erp, ecol, store), f"rows differ after delete stage {s}"
        rp, col = store.csr()
        assert np.array_equal(eng.gather_value(), oracle.bfs(n, rp, col, 0)), f"levels, delete stage {s}"
        st = eng.add_edge(ab, len(al))
        store.apply([], al)
        erp, ecol = eng.csr()
        assert same_rows(erp, ecol, store), f"rows differ after add stage {s}"
        rp, col = store.csr()
        assert np.array_equal(eng.gather_value(), oracle.bfs(n, rp, col, 0)), f"levels, add stage {s}"
        if s == 2:
            assert st["relocated_rows"] >= 1
        ab += len(al)
        db += len(dl)
    eng.close()


@pytest.mark.parametrize("quarters", [0, 1, 4, 64])
def test_grid_mixed_queries_bucket_widths(quarters):
    """Road-like grid (4-neighbour + diagonal shortcuts, weights 1..1000), 3 SSSP
    + 3 BFS sources + CC in one engine, lognormal batch sizes: every query
    bit-exact after every stage for every bucket width of the push closure
    (0 = plain frontier order) and CC parents rooted witnesses."""
    announce(f"grid 96x96 mixed queries, bucket quarters {quarters}")
    import dycgraph_amd as d
    from tests.helpers import cc_parents_ok
    side = 96
    src, dst = d.gen_grid(side, 0.05, seed=4)
    n = side * side
    sizes = d.gen_lognormal_sizes(4, 5.0, 1.0, 20, 600, seed=2).astype(np.int64)
    ap, dp = sizes - sizes // 2, sizes // 2
    adds, dels = d.gen_stream(n, src, dst, ap, dp, seed=8)
    apps = ("sssp", "bfs", "sssp", "bfs", "sssp", "bfs", "cc")
    sources = (0, 0, 4000, 4000, 9215, 123, 0)
    eng = _engine(app="sssp", sources=sources, query_apps=apps, weight_mode=1, weight_seed=3, weight_max=1000,
                  placement=0)
    eng.set_tuning(bucket_quarters=quarters)
    eng.load_graph(n, src, dst)
    eng.start()
    eng.load_updates(adds, dels)
    store = oracle.Store(n, src, dst)

    def check(stage):
        rp, col = store.csr()
        for q, (app, s) in enumerate(zip(apps, sources)):
            want = oracle_values(app, n, rp, col, s, oracle.W_SEEDED, 3, 1000)
            got = eng.gather_value(q)
            assert np.array_equal(got, want), (stage, q, app, int(np.sum(got != want)))
            if app == "cc":
                assert cc_parents_ok(n, rp, col, got, eng.gather_parent(q)), (stage, "cc parents")
            else:
                chk = oracle.check_sssp(n, rp, col, s, got, eng.gather_parent(q),
                                        oracle.W_SEEDED if app == "sssp" else oracle.W_UNIT, 3, 1000)
                assert chk["invalid_parents"] == 0 and chk["relaxable_edges"] == 0, (stage, q, chk)

    check("static")
    ab = db = 0
    for s in range(len(sizes)):
        eng.del_edge(db, int(dp[s]))
        store.apply(dels[db:db + dp[s]], [])
        check(f"delete {s}")
        eng.add_edge(ab, int(ap[s]))
        store.apply([], adds[ab:ab + ap[s]])
        check(f"add {s}")
        ab += int(ap[s])
        db += int(dp[s])
    eng.close()


@pytest.mark.parametrize("graph,placement", [("grid", 0), ("rmat", 1)])
def test_stage_autotune_exact(graph, placement):
    """The stage autotuner re-runs each armed stage's repair once per candidate
    from the same post-mutation state: values stay bit-exact after every stage
    (CC parents rooted witnesses), the tuner tried several candidates and kept
    one no slower than the defaults, and later stages run with its choice."""
    announce(f"stage autotune {graph} placement {placement}")
    import dycgraph_amd as d
    from tests.helpers import cc_parents_ok
    if graph == "grid":
        side = 64
        src, dst = d.gen_grid(side, 0.05, seed=5)
        n = side * side
        apps, sources, wmax = ("sssp", "bfs", "sssp", "cc"), (0, 100, 4000, 0), 1000
    else:
        n, src, dst = rmat_graph(13, 1 << 17, seed=9)
        apps, sources, wmax = ("sssp", "bfs", "cc"), (0, 1, 0), 100
    sizes = [1500, 300, 800]
    adds, dels = d.gen_stream(n, src, dst, sizes, sizes, seed=6)
    eng = _engine(app="sssp", sources=sources, query_apps=apps, weight_mode=1, weight_seed=7, weight_max=wmax,
                  placement=placement, cache_bytes=2 << 18)
    eng.load_graph(n, src, dst)
    eng.start()
    eng.load_updates(adds, dels)
    store = oracle.Store(n, src, dst)

    def check(stage):
        rp, col = store.csr()
        for q, (app, s) in enumerate(zip(apps, sources)):
            want = oracle_values(app, n, rp, col, s, oracle.W_SEEDED, 7, wmax)
            got = eng.gather_value(q)
            assert np.array_equal(got, want), (stage, q, app, int(np.sum(got != want)))
            if app == "cc":
                assert cc_parents_ok(n, rp, col, got, eng.gather_parent(q)), (stage, "cc parents")
            else:
                chk = oracle.check_sssp(n, rp, col, s, got, eng.gather_parent(q),
                                        oracle.W_SEEDED if app == "sssp" else oracle.W_UNIT, 7, wmax)
                assert chk["invalid_parents"] == 0 and chk["relaxable_edges"] == 0, (stage, q, chk)

    assert eng.last_tune()["evaluations"] == 0
    assert eng.stage_tune(0)["evaluations"] == 0 and eng.stage_tune(1)["evaluations"] == 0
    with pytest.raises(d.DcgError):
        eng.stage_tune(2)   # stages are 0 (delete) and 1 (add)
    eng.autotune_stages(3)   # both stages of snapshot 0, the delete stage of snapshot 1
    off = 0
    for s, k in enumerate(sizes):
        eng.del_edge(off, k)
        store.apply(dels[off:off + k], [])
        check(f"delete {s}")
        eng.add_edge(off, k)
        store.apply([], adds[off:off + k])
        check(f"add {s}")
        off += k
        if s == 0:
            t = eng.last_tune()
            assert t["evaluations"] >= 8, t
            assert t["best_ms"] <= t["first_ms"], t
            assert t["blocks_per_cu"] in (1, 2) and t["solo_max"] in (0, 512, 2048, 8192), t
            # each stage reports its own choice; the add stage's is the last one
            assert eng.stage_tune(1) == t, (eng.stage_tune(1), t)
            td = eng.stage_tune(0)
            assert td["evaluations"] >= 8 and td["best_ms"] <= td["first_ms"], td
            assert td["piece"] in (64, 128, 512, 2048) and td["big_row"] in (64, 128, 256, 1024, 4096), td
    # explicit closure parameters replace both stages' tuned ones
    eng.set_closure_grid(0, 2048)
    assert eng.stage_tune(0)["evaluations"] == 0 and eng.stage_tune(1)["evaluations"] == 0
    eng.close()


@pytest.mark.parametrize("placement", [0, 1])
def test_sorted_medium_rows_wave_search(placement):
    """Rows of 64..1,024 entries with sorted prefixes take the wave path's
    binary search: first occurrences across prefix and tail (duplicates,
    over-requests), tombstones, the wave compaction past a quarter of the row,
    and rows that shrink below the search length.  Rows (tombstones skipped)
    and BFS levels match the multiset model after every stage."""
    announce(f"sorted medium rows, wave search, placement={placement}")
    rng = np.random.default_rng(21)
    n = 5000
    src, dst = [], []
    for u in range(60):
        deg = int(rng.integers(70, 1000))
        row = sorted(int(x) for x in rng.integers(0, n, deg))      # sorted, with repeats
        src += [u] * len(row)
        dst += row
    src += [int(x) for x in rng.integers(60, n, 3000)]
    dst += [int(x) for x in rng.integers(0, n, 3000)]
    src, dst = np.array(src, np.uint32), np.array(dst, np.uint32)
    store = oracle.Store(n, src, dst)
    rp0, col0 = store.csr()
    stages = []
    for s in range(5):
        dl, al = [], []
        frac = 0.4 if s == 3 else 0.05
        for u in range(60):
            row = col0[rp0[u]:rp0[u + 1]]
            k = int(len(row) * frac)
            pick = [int(x) for x in rng.choice(row, k)] if k else []
            dl += [(u, v) for v in pick]                             # repeats -> over-requests
            al += [(u, int(x)) for x in rng.integers(0, n, 3)]
            al += [(u, int(row[0]))] if len(row) else []             # a prefix value again, in the tail
        dl += [(u, v) for u, v in al[:40]] if s else []              # tail hits of earlier appends
        stages.append((dl, al))
    adds = [e for _, a in stages for e in a]
    dels = [e for d_, _ in stages for e in d_]
    eng = _engine(app="bfs", placement=placement, cache_bytes=1 << 18)
    eng.load_graph(n, src, dst)
    eng.start()
    eng.load_updates(adds, dels)
    ab = db = 0
    for s, (dl, al) in enumerate(stages):
        st = eng.del_edge(db, len(dl))
        m = store.apply(dl, [])
        assert st["missing_deletes"] == m["missing_deletes"], f"stage {s}"
        erp, ecol = eng.csr()
        assert same_rows(erp, ecol, store), f"rows differ after delete stage {s}"
        rp, col = store.csr()
        assert np.array_equal(eng.gather_value(), oracle.bfs(n, rp, col, 0)), f"levels, delete stage {s}"
        eng.add_edge(ab, len(al))
        store.apply([], al)
        erp, ecol = eng.csr()
        assert same_rows(erp, ecol, store), f"rows differ after add stage {s}"
        rp, col = store.csr()
        assert np.array_equal(eng.gather_value(), oracle.bfs(n, rp, col, 0)), f"levels, add stage {s}"
        ab += len(al)
        db += len(dl)
    eng.close()


def test_relocated_blocks_are_reused():
    """A long stream of inserts and deletes: rows outgrow their blocks again and
    again.  The block a relocated row leaves joins a free pool and is reused by
    later batches (rows that need a block of that class), rows and levels stay
    exact, and the pool's accounting adds up."""
    announce("block reuse over 150 snapshots")
    import dycgraph_amd as d
    n, src, dst = rmat_graph(11, 12000, seed=4)
    per = [600] * 150
    adds, dels = d.gen_stream(n, src, dst, per, per, seed=12)
    eng = _engine(app="bfs", placement=0)
    eng.load_graph(n, src, dst)
    eng.start()
    eng.load_updates(adds, dels)
    store = oracle.Store(n, src, dst)
    relocated = 0
    ab = db = 0
    for s, na in enumerate(per):
        eng.del_edge(db, na)
        st = eng.add_edge(ab, na)
        relocated += st["relocated_rows"]
        store.apply(dels[db:db + na], adds[ab:ab + na])
        if s % 25 == 24:
            erp, ecol = eng.csr()
            assert same_rows(erp, ecol, store), f"rows differ after snapshot {s}"
        ab += na
        db += na
    info = eng.info()
    erp, ecol = eng.csr()
    assert same_rows(erp, ecol, store)
    rp, col = store.csr()
    assert np.array_equal(eng.gather_value(), oracle.bfs(n, rp, col, 0))
    print(f"relocated {relocated}, reused {info['reused_blocks']}, in pool {info['free_blocks']}", flush=True)
    assert info["reused_blocks"] > relocated // 20
    assert info["reused_blocks"] + info["free_blocks"] <= relocated
    eng.close()


@pytest.mark.parametrize("app", ["bfs", "cc"])
@pytest.mark.parametrize("placement", [0, 1])
def test_fenced_prefix_search_rows(app, placement):
    """Prefix searches through the fences: single targets (the whole wave),
    a few targets (a quarter-wave each) and many targets (a lane each); hits
    in the sorted prefix, a repeated value whose first occurrence is already a
    tombstone, a value only in the appended tail, a missing value, and small
    rows where tombstones cross the compaction point (fences refreshed by the
    compaction, then searched again).  Rows, values and the missing count
    match the multiset model after every stage."""
    announce(f"fenced prefix search, app={app}, placement={placement}")
    rng = np.random.default_rng(33)
    n = 4000
    src, dst = [], []
    for u in range(300):
        deg = int(rng.integers(4, 12)) if u < 100 else int(rng.integers(12, 400))
        row = sorted(int(x) for x in rng.integers(0, n, deg))
        row += [row[len(row) // 2]] * int(rng.integers(0, 3))          # repeats of one value
        row.sort()
        src += [u] * len(row)
        dst += row
    src += [int(x) for x in rng.integers(300, n, 4000)]
    dst += [int(x) for x in rng.integers(0, n, 4000)]
    src, dst = np.array(src, np.uint32), np.array(dst, np.uint32)
    store = oracle.Store(n, src, dst)
    stages, last = [], {}
    appended = {u: [] for u in range(300)}
    rp, col = store.csr()
    for s in range(6):
        dl, al = [], []
        for u in range(300):
            row = [int(x) for x in col[rp[u]:rp[u + 1]]]
            # one target (whole-wave search), a few (quarter-waves), many (a lane each)
            nd = 1 if u % 3 == 0 else (int(rng.integers(2, 9)) if u % 3 == 1 else int(rng.integers(17, 41)))
            for _ in range(nd):
                kind = int(rng.integers(0, 5))
                if kind == 1 and u in last:
                    v = last[u]                                       # next occurrence or missing
                elif kind == 2 and appended[u]:
                    v = appended[u][int(rng.integers(0, len(appended[u])))]
                elif kind == 3:
                    v = int(n - 1 - (u % 7))                          # usually absent
                elif row:
                    v = row[int(rng.integers(0, len(row)))]
                else:
                    continue
                dl.append((u, v))
                last[u] = v
            if rng.random() < 0.3:
                w = int(rng.integers(0, n))
                al.append((u, w))
                appended[u].append(w)
        stages.append((dl, al))
        store.apply(dl, al)
        rp, col = store.csr()
    store = oracle.Store(n, src, dst)
    adds = [e for _, a in stages for e in a]
    dels = [e for d_, _ in stages for e in d_]
    eng = _engine(app=app, placement=placement, cache_bytes=1 << 18)
    eng.load_graph(n, src, dst)
    eng.start()
    eng.load_updates(adds, dels)
    def ref(n_, rp_, col_):
        return oracle_values(app, n_, rp_, col_)
    ab = db = 0
    for s, (dl, al) in enumerate(stages):
        st = eng.del_edge(db, len(dl))
        m = store.apply(dl, [])
        assert st["missing_deletes"] == m["missing_deletes"], f"stage {s}"
        erp, ecol = eng.csr()
        assert same_rows(erp, ecol, store), f"rows differ after delete stage {s}"
        rp, col = store.csr()
        got = eng.gather_value()
        assert np.array_equal(got, ref(n, rp, col)), f"{app} values, delete stage {s}"
        if app == "cc":
            assert cc_parents_ok(n, rp, col, got, eng.gather_parent()), f"CC parents, delete stage {s}"
        eng.add_edge(ab, len(al))
        store.apply([], al)
        erp, ecol = eng.csr()
        assert same_rows(erp, ecol, store), f"rows differ after add stage {s}"
        rp, col = store.csr()
        assert np.array_equal(eng.gather_value(), ref(n, rp, col)), f"{app} values, add stage {s}"
        ab += len(al)
        db += len(dl)
    eng.close()


@pytest.mark.parametrize("placement", [0, 1])
def test_pagerank_dense_sweeps(placement):
    """PageRank batches large enough for dense sweeps (more than 1/32 of the
    vertices pending): small in-rows pulled by quarter-waves, a hub in-row of
    5,000 slots cut into three pieces, vertices without in-edges.  After the
    static solve and every batch the ranks pass the reference's checker
    (hybrid_pr_host.cu:9-56) against the multiset model's rows."""
    announce(f"pagerank dense sweeps, placement={placement}")
    import dycgraph_amd as d
    n, src, dst = rmat_graph(13, 60000, seed=8)
    rng = np.random.default_rng(3)
    hub = 7
    extra = rng.choice(np.arange(100, n), 5000, replace=False).astype(np.uint32)
    src = np.concatenate([src, extra])
    dst = np.concatenate([dst, np.full(extra.size, hub, np.uint32)])
    per = [1500, 1500, 1500]
    adds, dels = d.gen_stream(n, src, dst, per, per, seed=5)
    # the hub's mass needs eps 1e-5 to settle within the 100-round cap
    eng = _engine(app="pr", placement=placement, cache_bytes=1 << 18, pr_base=0.15, pr_epsilon=1e-5)
    eng.load_graph(n, src, dst)
    eng.start()
    store = oracle.Store(n, src, dst)
    rp, col = store.csr()
    chk = oracle.check_pagerank(n, rp, col, eng.gather_value(), eng.gather_buffer(), eps=1e-5)
    assert chk["ok"], repr(chk)
    eng.load_updates(adds, dels)
    ab = db = 0
    for s, k in enumerate(per):
        eng.update_pagerank(ab, k, db, k)
        store.apply(dels[db:db + k], [])
        store.apply([], adds[ab:ab + k])
        rp, col = store.csr()
        chk = oracle.check_pagerank(n, rp, col, eng.gather_value(), eng.gather_buffer(), eps=1e-5)
        assert chk["ok"], repr((s, chk))
        ab += k
        db += k
    eng.close()


@pytest.mark.parametrize("app", ["sssp", "cc"])
def test_reload_updates_between_snapshots(app):
    """A second dcg_load_updates after stages already ran (and their repair
    launches were captured as graphs): the captured graphs read the batch's
    device records through the stage arguments, so the replays after the reload
    see the new stream.  Values and rows match the oracle after every stage."""
    announce(f"reload updates between snapshots under captured graphs, {app}")
    import dycgraph_amd as d
    n, src, dst = rmat_graph(12, 40000, seed=11)
    eng = _engine(app=app, sources=(0,), placement=1, cache_bytes=1 << 18)
    eng.set_tuning(graphs=True)
    eng.load_graph(n, src, dst)
    eng.start()
    store = oracle.Store(n, src, dst)
    for part in range(3):
        rp, col = store.csr()
        srcs = np.repeat(np.arange(n, dtype=np.uint32), np.diff(rp).astype(np.int64))
        order = np.lexsort((col, srcs))   # the generator takes the edge set sorted by (src, dst)
        per = [250 + 100 * part] * 2
        adds, dels = d.gen_stream(n, srcs[order], col[order], per, per, seed=40 + part)
        eng.load_updates(adds, dels)   # replaces the stream: slices index the new one from 0
        for s, k in enumerate(per):
            eng.del_edge(s * k, k)
            store.apply(dels[s * k:(s + 1) * k], [])
            rp, col = store.csr()
            assert np.array_equal(eng.gather_value(), oracle_values(app, n, rp, col)), (part, s, "delete")
            eng.add_edge(s * k, k)
            store.apply([], adds[s * k:(s + 1) * k])
            rp, col = store.csr()
            assert np.array_equal(eng.gather_value(), oracle_values(app, n, rp, col)), (part, s, "add")
        erp, ecol = eng.csr()
        assert same_rows(erp, ecol, store), f"rows after part {part}"
    eng.close()


def test_barrier_timeout_ends_the_launch_and_reports():
    """A grid barrier that gives up (forced from the host with a zero patience,
    so every workgroup that waits gives up at once) ends every round of the
    launch: the stage fails with DCG_ERR_HIP and every kernel exits; a fresh
    engine on the same device is exact afterwards."""
    import dycgraph_amd as d
    announce("barrier timeout error path")
    n, src, dst = rmat_graph(13, 1 << 17, seed=4)
    adds, dels = d.gen_stream(n, src, dst, [3000], [3000], seed=2)
    eng = _engine(app="sssp", weight_mode=1, weight_seed=3, weight_max=100)
    eng.load_graph(n, src, dst)
    eng.start()
    eng.load_updates(adds, dels)
    eng.set_barrier_timeout(0)
    with pytest.raises(d.DcgError) as err:
        eng.del_edge(0, 3000)
        eng.add_edge(0, 3000)
    assert err.value.status == -2 and "grid barrier timed out" in str(err.value)
    eng.close()
    eng = _engine(app="sssp", weight_mode=1, weight_seed=3, weight_max=100)
    eng.load_graph(n, src, dst)
    eng.start()
    eng.load_updates(adds, dels)
    eng.del_edge(0, 3000)
    eng.add_edge(0, 3000)
    store = oracle.Store(n, src, dst)
    store.apply(dels, adds)
    rp, col = store.csr()
    assert np.array_equal(eng.gather_value(), oracle.sssp(n, rp, col, 0, oracle.W_SEEDED, 3, 100))
    eng.close()
