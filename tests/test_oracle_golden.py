"""Pin the CPU oracle against the reference's own known-answer tests, its fixture,
closed forms and brute force (CPU only)."""
import heapq
import itertools
import os
from collections import deque

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
INF = oracle.INF


# ----------------------------------------------------------------- store model
def test_chunk_store_kat_batch_ordering_and_duplicates():
    # tests/source_local_chunk_store_test.cpp:35-56
    src = [0, 0, 0] + [2] * 32
    dst = [1, 1, 2] + [7] * 32
    s = oracle.Store(4, src, dst)
    m = s.apply(dels=[(0, 1), (1, 99), (2, 7)], adds=[(0, 3), (1, 4), (0, 1)])
    assert m["touched_sources"] == 3 and m["changed_sources"] == 3 and m["missing_deletes"] == 1
    assert s.row(0) == [1, 2, 3, 1]
    assert s.row(1) == [4]
    assert len(s.row(2)) == 31
    assert s.edges() == 36


def test_chunk_store_kat_source_isolation():
    # tests/source_local_chunk_store_test.cpp:58-80 (row contents part)
    s = oracle.Store(3, [0, 0, 1, 1], [1, 2, 8, 9])
    m = s.apply(adds=[(0, 3), (0, 4), (0, 5)])
    assert m["changed_sources"] == 1
    assert s.row(0) == [1, 2, 3, 4, 5] and s.row(1) == [8, 9] and s.row(2) == []


def _read_fixture():
    g = np.loadtxt(os.path.join(GOLDEN, "topology_replay_graph.txt"), dtype=np.int64).reshape(-1, 2)
    sizes = np.loadtxt(os.path.join(GOLDEN, "topology_replay_batch_sizes.txt"), dtype=np.int64).reshape(-1, 2)
    lines = open(os.path.join(GOLDEN, "topology_replay_updates.txt")).read().split("\n")
    ups = [ln.split() for ln in lines if ln.strip()]
    return g, sizes, ups


def test_topology_replay_fixture_golden_line():
    g, sizes, ups = _read_fixture()
    s = oracle.Store(16, g[:, 0], g[:, 1])
    pos = 0
    out = []
    for b, (na, nd) in enumerate(sizes):
        adds, dels = [], []
        while len(adds) < na or len(dels) < nd:
            t, u, v = ups[pos][0], int(ups[pos][1]), int(ups[pos][2])
            pos += 1
            (adds if t == "a" else dels).append((u, v))
        m = s.apply(dels, adds)
        out.append(f"batch={b} additions={na} deletions={nd} touched_sources={m['touched_sources']} "
                   f"changed_sources={m['changed_sources']} missing_deletes={m['missing_deletes']} "
                   f"invalid_additions={m['invalid_additions']} graph_edges={m['graph_edges']}")
    # CMakeLists.txt:129-131
    assert ("batch=1 additions=1 deletions=1 touched_sources=2 changed_sources=1 missing_deletes=1 "
            "invalid_additions=0 graph_edges=6") in out[1]


def test_store_first_occurrence_order():
    s = oracle.Store(8, [0, 0, 0, 0], [5, 3, 5, 3])
    s.apply(dels=[(0, 5)])
    assert s.row(0) == [3, 5, 3]      # the first 5 goes, survivors keep order
    s.apply(dels=[(0, 3), (0, 3), (0, 3)])
    assert s.row(0) == [5]
    assert s.edges() == 1


def test_hash_sequence_closed_form():
    # HashSequence of the empty sequence: FNV basis ^ 0, times the prime
    assert oracle.hash_sequence([]) == (1469598103934665603 * 1099511628211) % 2**64


# -------------------------------------------------------------- BFS/SSSP KATs
def _kat_bfs_graph():
    # tests/bfs_dynamic_smoke.py:56-60
    n = 2048
    edges = {(u, (u + 1) % n) for u in range(n)}
    edges |= {(u, 1024) for u in range(1024)} | {(0, 0), (0, 2), (1, 3)}
    return n, sorted(edges)


def _csr(n, edges):
    e = np.array(edges, dtype=np.uint32).reshape(-1, 2)
    return oracle.csr_from_edges(n, e[:, 0], e[:, 1])


def test_bfs_kat_closed_form():
    n, edges = _kat_bfs_graph()
    rp, col = _csr(n, edges)
    d = oracle.bfs(n, rp, col, 0)
    assert d[0] == 0 and d[1] == 1 and d[2] == 1 and d[1024] == 1
    assert all(d[v] == v - 1 for v in range(2, 1024))
    assert all(d[v] == v - 1023 for v in range(1024, 2048))
    # delete stage of the smoke test's first batch
    dels = {(0, 1024), (20, 21), (700, 701), (1200, 1201)}
    rp, col = _csr(n, [e for e in edges if e not in dels])
    d = oracle.bfs(n, rp, col, 0)
    assert d[1024] == 2                               # 0 -> 1 -> 1024
    assert all(d[v] == INF for v in range(21, 1024))  # cycle cut at 20 -> 21
    assert all(d[v] == v - 1022 for v in range(1024, 1201))
    assert all(d[v] == INF for v in range(1201, 2048))


def _py_dijkstra(n, edges, w):
    adj = [[] for _ in range(n)]
    for u, v in edges:
        adj[u].append(v)
    dist = [INF] * n
    dist[0] = 0
    pq = [(0, 0)]
    while pq:
        du, u = heapq.heappop(pq)
        if du != dist[u]:
            continue
        for v in adj[u]:
            c = du + w(u, v)
            if c < dist[v]:
                dist[v] = c
                heapq.heappush(pq, (c, v))
    return dist


def test_sssp_kat_matches_independent_dijkstra():
    n, edges = _kat_bfs_graph()
    rp, col = _csr(n, edges)
    d = oracle.sssp(n, rp, col, 0, oracle.W_REFERENCE)
    want = _py_dijkstra(n, edges, lambda u, v: (u + v) % 128 + 1)
    assert d.tolist() == want
    # a closed form: every vertex of 2..1023 is reached along the cycle from 2
    assert d[2] == (0 + 2) % 128 + 1


def _floyd(n, edges, w):
    D = np.full((n, n), np.inf)
    np.fill_diagonal(D, 0)
    for u, v in edges:
        D[u, v] = min(D[u, v], w(u, v))
    for k in range(n):
        D = np.minimum(D, D[:, k:k + 1] + D[k:k + 1, :])
    return D


@pytest.mark.parametrize("seed", range(6))
def test_sssp_bfs_brute_force_random(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(5, 40))
    m = int(rng.integers(0, 4 * n))
    e = rng.integers(0, n, size=(m, 2))
    edges = [tuple(map(int, x)) for x in e]
    rp, col = _csr(n, edges)
    for wmode, wf in ((oracle.W_REFERENCE, lambda u, v: (u + v) % 128 + 1),
                      (oracle.W_SEEDED, lambda u, v: oracle.weight(oracle.W_SEEDED, 99, 1000, u, v))):
        D = _floyd(n, edges, wf)
        for s in (0, n - 1):
            d = oracle.sssp(n, rp, col, s, wmode, 99, 1000)
            want = [INF if np.isinf(x) else int(x) for x in D[s]]
            assert d.tolist() == want
    D = _floyd(n, edges, lambda u, v: 1)
    d = oracle.bfs(n, rp, col, 0)
    assert d.tolist() == [INF if np.isinf(x) else int(x) for x in D[0]]


def test_seeded_weight_range_and_purity():
    ws = [oracle.weight(oracle.W_SEEDED, 5, 100, u, v) for u in range(50) for v in range(50)]
    assert min(ws) >= 1 and max(ws) <= 100 and len(set(ws)) > 50
    assert oracle.weight(oracle.W_SEEDED, 5, 100, 3, 4) == oracle.weight(oracle.W_SEEDED, 5, 100, 3, 4)
    assert oracle.weight(oracle.W_REFERENCE, 0, 0, 100, 60) == (160 % 128) + 1


# ------------------------------------------------------------------- CC KATs
def _cc_base():
    # tests/cc_dynamic_smoke.py:52-66
    base = [(0, 0), (11, 11), (255, 255)]
    base += [(0, 1), (1, 2), (2, 0), (2, 3), (3, 4), (4, 5), (5, 3)]
    base += [(8, 9), (8, 9), (130, 140), (140, 150), (150, 130)]
    base += [(u, u + 1) for u in range(160, 200)]
    base += [(31, 30), (32, 30), (40, 41), (41, 42), (42, 41)]
    base += [(0, 220), (220, 210), (210, 211), (211, 210)]
    base += [(50, 51), (50, 52), (51, 53), (52, 53), (53, 54), (54, 53)]
    base += [(60, 61), (60, 62), (61, 63), (62, 63), (63, 64), (64, 63)]
    return 256, base


def test_cc_kat_hand_labels():
    n, base = _cc_base()
    e = np.array(base, dtype=np.uint32)
    s = oracle.Store(n, e[:, 0], e[:, 1])
    rp, col = s.csr()
    lab = oracle.cc(n, rp, col)
    assert lab[:6].tolist() == [0, 0, 0, 0, 0, 0]
    assert lab[9] == 8 and lab[30] == 30 and lab[41] == 40 and lab[42] == 40
    assert lab[140] == 130 and lab[199] == 160 and lab[210] == 0 and lab[53] == 50
    # batch 1 deletes (2,3),(40,41),(220,210): the 3-4-5 cycle keeps only its own seed
    s.apply(dels=[(2, 3), (40, 41), (220, 210)])
    rp, col = s.csr()
    lab = oracle.cc(n, rp, col)
    assert lab[3] == lab[4] == lab[5] == 3
    assert lab[41] == 41 and lab[42] == 41 and lab[210] == 210 and lab[211] == 210
    # duplicates: deleting one copy of (8,9) keeps 9 labelled 8; the second copy frees it
    s.apply(dels=[(8, 9)])
    rp, col = s.csr()
    assert oracle.cc(n, rp, col)[9] == 8
    s.apply(dels=[(8, 9)])
    rp, col = s.csr()
    assert oracle.cc(n, rp, col)[9] == 9


@pytest.mark.parametrize("seed", range(5))
def test_cc_brute_force_closure(seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(3, 30))
    m = int(rng.integers(0, 3 * n))
    e = rng.integers(0, n, size=(m, 2))
    rp, col = oracle.csr_from_edges(n, e[:, 0], e[:, 1])
    R = np.eye(n, dtype=bool)
    for u, v in e:
        R[u, v] = True
    for k in range(n):
        R = R | (R[:, k:k + 1] & R[k:k + 1, :])
    want = [int(np.argmax(R[:, v])) for v in range(n)]   # smallest u reaching v
    assert oracle.cc(n, rp, col).tolist() == want


# ------------------------------------------------------------------ PageRank
def test_pagerank_closed_forms():
    # 2-cycle: x = 0.15 / (1 - 0.85) = 1
    rp, col = oracle.csr_from_edges(2, [0, 1], [1, 0])
    x, r = oracle.pagerank(2, rp, col)
    assert r > 0 and np.allclose(x, 1.0, atol=1e-9)
    # star with dangling leaves: hub 0.15, leaf 0.15 + 0.85 * 0.15 / k
    k = 7
    rp, col = oracle.csr_from_edges(k + 1, [0] * k, list(range(1, k + 1)))
    x, _ = oracle.pagerank(k + 1, rp, col)
    assert abs(x[0] - 0.15) < 1e-12 and np.allclose(x[1:], 0.15 + 0.85 * 0.15 / k, atol=1e-12)


@pytest.mark.parametrize("seed", range(4))
def test_pagerank_dense_solve(seed):
    rng = np.random.default_rng(7 + seed)
    n = int(rng.integers(3, 25))
    e = rng.integers(0, n, size=(int(rng.integers(1, 4 * n)), 2))
    rp, col = oracle.csr_from_edges(n, e[:, 0], e[:, 1])
    P = np.zeros((n, n))
    for u, v in e:
        P[u, v] += 1
    deg = P.sum(1, keepdims=True)
    P = np.divide(P, deg, out=np.zeros_like(P), where=deg > 0)   # multiplicities count, dangling dropped
    want = np.linalg.solve(np.eye(n) - 0.85 * P.T, np.full(n, 0.15))
    x, r = oracle.pagerank(n, rp, col)
    assert r > 0 and np.max(np.abs(x - want)) < 1e-9


def test_pagerank_check_accepts_exact_and_rejects_perturbed():
    rng = np.random.default_rng(3)
    n = 50
    e = rng.integers(0, n, size=(200, 2))
    rp, col = oracle.csr_from_edges(n, e[:, 0], e[:, 1])
    x, _ = oracle.pagerank(n, rp, col)
    rank = x.astype(np.float32)
    res = np.zeros(n, np.float32)
    ok = oracle.check_pagerank(n, rp, col, rank, res)
    assert ok["ok"] and ok["l1"] < 1e-4
    bad = rank.copy()
    bad[5] += 0.5
    assert not oracle.check_pagerank(n, rp, col, bad, res)["ok"]


# ------------------------------------------------- the reference's own oracles
def unpack_u32(text):
    """Values of tests/golden/reference_vectors.json (base64 of zlib of uint32 LE)."""
    import base64
    import zlib
    return np.frombuffer(zlib.decompress(base64.b64decode(text)), dtype="<u4").astype(np.uint32)


def reference_vectors():
    import json
    with open(os.path.join(GOLDEN, "reference_vectors.json")) as f:
        return json.load(f)


def bfs_kat_stages(doc):
    """(stage record, current edge list) of the BFS/SSSP KAT graph after each stage."""
    b = doc["bfs_sssp"]
    edges = set(map(tuple, b["initial_edges"]))
    for st in b["stages"]:
        edges -= set(map(tuple, st["dels"]))
        edges |= set(map(tuple, st["adds"]))
        yield st, sorted(edges)


def test_reference_vectors_bfs_sssp_values_and_checksums():
    """tests/bfs_dynamic_smoke.py's oracle() (FIFO BFS / Dijkstra with (u+v)%128+1,
    then the checksum of hybrid_sssp.cu:335-347) on its KAT graph after every stage:
    the oracle's BFS, SSSP and checksum give the same values and the same 64-bit
    checksums, which exercise the <<6 and >>2 terms on non-zero distances."""
    doc = reference_vectors()
    n = doc["bfs_sssp"]["n"]
    for st, edges in bfs_kat_stages(doc):
        e = np.array(edges, np.int64)
        rp, col = oracle.csr_from_edges(n, e[:, 0], e[:, 1])
        bfs = oracle.bfs(n, rp, col, 0)
        sssp = oracle.sssp(n, rp, col, 0, oracle.W_REFERENCE)
        assert np.array_equal(bfs, unpack_u32(st["bfs"]["values"])), st["stage"]
        assert np.array_equal(sssp, unpack_u32(st["sssp"]["values"])), st["stage"]
        assert oracle.checksum(bfs) == st["bfs"]["checksum"], st["stage"]
        assert oracle.checksum(sssp) == st["sssp"]["checksum"], st["stage"]
        assert sssp.max() > 1000   # large distances: every term of the checksum matters


def test_reference_vectors_cc_labels_and_checksums():
    """tests/cc_dynamic_smoke.py's oracle() on its KAT graph and all 20 batches (the
    last six drawn with random.Random(20260923)), applied through the oracle's
    multiset store: labels and checksums after every delete and add stage."""
    doc = reference_vectors()["cc"]
    n = doc["n"]
    base = np.array(doc["base"], np.int64)
    store = oracle.Store(n, base[:, 0], base[:, 1])
    rp, col = store.csr()
    lab = oracle.cc(n, rp, col)
    assert np.array_equal(lab, unpack_u32(doc["initial"]["labels"]))
    assert oracle.checksum(lab) == doc["initial"]["checksum"]
    for i, b in enumerate(doc["batches"]):
        for phase, dels, adds in (("after_delete", b["dels"], []), ("after_add", [], b["adds"])):
            store.apply(dels=dels, adds=adds)
            rp, col = store.csr()
            lab = oracle.cc(n, rp, col)
            assert np.array_equal(lab, unpack_u32(b[phase]["labels"])), (i, phase)
            assert oracle.checksum(lab) == b[phase]["checksum"], (i, phase)


def test_mix64_is_splitmix64():
    """orc_mix64(i * 0x9e3779b97f4a7c15) is the i+1-th output of SplitMix64 seeded
    with 0 (Steele, Lea, Flood 2014); the first four published outputs."""
    g = 0x9e3779b97f4a7c15
    want = [0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4, 0x06c45d188009454f, 0xf88bb8a8724c81ec]
    got = [oracle.lib().orc_mix64((i * g) % 2**64) for i in range(4)]
    assert got == want
    # seeded weights follow: 1 + mix64(seed ^ (u << 32 | v)) % W
    assert oracle.weight(oracle.W_SEEDED, 0, 100, 0, 0) == 1 + want[0] % 100


# ------------------------------------------------------------------- checkers
def test_checksum_closed_forms():
    assert oracle.checksum(np.full(5, INF, np.uint32)) == 1469598103934665603
    one = ((1469598103934665603 ^ (0x9e3779b97f4a7c15)) * 1099511628211) % 2**64
    assert oracle.checksum(np.array([0], np.uint32)) == one
    a = np.array([0, 3, INF, 9], np.uint32)
    b = a.copy()
    b[3] = 8
    assert oracle.checksum(a) != oracle.checksum(b)


def test_sssp_checker_planted_faults():
    n, edges = _kat_bfs_graph()
    rp, col = _csr(n, edges)
    d = oracle.sssp(n, rp, col, 0)
    parent = np.full(n, INF, np.uint32)
    for u, v in edges:
        if u != v and d[u] != INF and d[u] + (u + v) % 128 + 1 == d[v] and parent[v] == INF:
            parent[v] = u
    ok = oracle.check_sssp(n, rp, col, 0, d, parent)
    assert ok["relaxable_edges"] == 0 and ok["missing_tight_witnesses"] == 0 and ok["invalid_parents"] == 0
    worse = d.copy()
    worse[500] += 1          # a relaxable edge into 500 appears, 500 loses its witness
    r = oracle.check_sssp(n, rp, col, 0, worse, parent)
    assert r["relaxable_edges"] >= 1 and r["missing_tight_witnesses"] >= 1
    bad_parent = parent.copy()
    bad_parent[700] = 5      # 5 -> 700 is not an edge
    assert oracle.check_sssp(n, rp, col, 0, d, bad_parent)["missing_parent_edges"] == 1


# ------------------------------------------------------ CPU incremental path
@pytest.mark.parametrize("app", [oracle.APP_SSSP, oracle.APP_BFS, oracle.APP_CC])
@pytest.mark.parametrize("seed", range(3))
def test_incremental_cpu_matches_recompute(app, seed):
    rng = np.random.default_rng(50 + seed)
    n = 300
    e = rng.integers(0, n, size=(1500, 2))
    rp, col = oracle.csr_from_edges(n, e[:, 0], e[:, 1])
    inc = oracle.Incr(n, rp, col, app, source=0)
    inc.converge()
    store = oracle.Store(n, e[:, 0], e[:, 1])

    def recompute():
        r, c = store.csr()
        if app == oracle.APP_SSSP:
            return oracle.sssp(n, r, c, 0)
        if app == oracle.APP_BFS:
            return oracle.bfs(n, r, c, 0)
        return oracle.cc(n, r, c)

    assert np.array_equal(inc.values()[0], recompute())
    for _ in range(6):
        r, c = store.csr()
        srcs = np.repeat(np.arange(n), np.diff(r).astype(np.int64))
        pick = rng.choice(len(c), size=min(60, len(c)), replace=False)
        dels = np.stack([srcs[pick], c[pick]], 1)
        adds = rng.integers(0, n, size=(60, 2))
        inc.delete_phase(dels)
        store.apply(dels=dels)
        assert np.array_equal(inc.values()[0], recompute())
        inc.add_phase(adds)
        store.apply(adds=adds)
        assert np.array_equal(inc.values()[0], recompute())


def test_incremental_cpu_pagerank_within_bound():
    rng = np.random.default_rng(9)
    n = 400
    e = rng.integers(0, n, size=(3000, 2))
    rp, col = oracle.csr_from_edges(n, e[:, 0], e[:, 1])
    inc = oracle.Incr(n, rp, col, oracle.APP_PR, pr_base=0.15, pr_eps=1e-6, pr_max_rounds=100)
    inc.converge()
    store = oracle.Store(n, e[:, 0], e[:, 1])
    for _ in range(3):
        r, c = store.csr()
        srcs = np.repeat(np.arange(n), np.diff(r).astype(np.int64))
        pick = rng.choice(len(c), size=40, replace=False)
        dels = np.stack([srcs[pick], c[pick]], 1)
        adds = rng.integers(0, n, size=(40, 2))
        inc.pr_batch(dels, adds)
        store.apply(dels=dels, adds=adds)
        r, c = store.csr()
        rank, res = inc.pr_values()
        chk = oracle.check_pagerank(n, r, c, rank, res, base=0.15, eps=1e-6)
        assert chk["ok"], chk


# ----------------------------------------------------------------- MLP oracle
def test_mlp_gradients_central_differences():
    rng = np.random.default_rng(0)
    shapes = [(8, 64), (64,), (64, 64), (64,), (64, 2), (2,)]
    params = [rng.normal(0, 0.3, s).astype(np.float32) for s in shapes]
    x = rng.normal(0, 1, (5, 8)).astype(np.float32)
    y = np.stack([rng.integers(0, 2, 5), rng.normal(0, 1, 5)], 1).astype(np.float32)
    _, loss, grads = oracle.mlp_forward_backward(x, y, *params)
    eps = 1e-4
    checked = 0
    for pi, (p, g) in enumerate(zip(params, grads)):
        flat = p.reshape(-1)
        for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + eps
            _, lp, _ = oracle.mlp_forward_backward(x, y, *params, grads=False)
            flat[idx] = old - eps
            _, lm, _ = oracle.mlp_forward_backward(x, y, *params, grads=False)
            flat[idx] = old
            fwd, bwd = (lp - loss) / eps, (loss - lm) / eps
            if abs(fwd - bwd) > 1e-2 * max(1.0, abs(fwd)):
                continue          # a ReLU kink lies inside the stencil
            num = (lp - lm) / (2 * eps)
            assert abs(num - g.reshape(-1)[idx]) < 1e-3 * max(1.0, abs(num)), (pi, idx, num, g.reshape(-1)[idx])
            checked += 1
    assert checked >= 24


# ------------------------------------------------- hotness ranking / admission
def _window(rng, n):
    b = rng.integers(0, 256, (n, 4)).astype(np.uint32)
    b[rng.random(n) < 0.3] = 0                     # many zero scores: ties by id
    b[rng.random(n) < 0.1] = [255, 255, 255, 255]  # the largest score, tied
    return b[:, 0] | (b[:, 1] << 8) | (b[:, 2] << 16) | (b[:, 3] << 24)


def test_hot_candidates_brute_force():
    """The ranking is the permutation sorted by (score descending, id ascending)
    with unreached vertices scored 0; the count is the largest k whose first k
    ranked degrees sum below the cap (k = n when everything fits, 0 without a
    cache), found here by trying every k."""
    rng = np.random.default_rng(12)
    for trial in range(200):
        n = int(rng.integers(1, 40))
        win = _window(rng, n)
        deg = rng.integers(0, 6, n).astype(np.uint32)
        value = np.where(rng.random(n) < 0.2, INF, rng.integers(0, 50, n)).astype(np.uint32)
        mask = None if trial % 3 == 0 else value
        score = sum(((win >> (8 * k)) & 0xFF).astype(np.int64) for k in range(4))
        if mask is not None:
            score = np.where(mask == INF, 0, score)
        order = sorted(range(n), key=lambda v: (-int(score[v]), v))
        total = int(deg.sum())
        for cap in (0, 1, int(rng.integers(1, total + 2)), total, total + 1):
            k, ids = oracle.hot_candidates(win, deg, cap, mask)
            assert ids.tolist() == order
            pref = np.concatenate([[0], np.cumsum(deg[order])])
            want = 0 if cap == 0 else max(j for j in range(n + 1) if pref[j] < cap)
            assert k == want, (trial, cap, k, want)
