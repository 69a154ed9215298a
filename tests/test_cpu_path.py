"""The parallel CPU path (oracle/cpu_path.c, the timed CPU baseline) against
the from-scratch oracle: SSSP, BFS and CC queries over one store, every delete
and add stage, with duplicate and missing deletes, out-of-range records and
rows that outgrow their blocks; rows equal the multiset model's as multisets."""
import numpy as np

import oracle


def _rmat_like(n, m, seed):
    rng = np.random.default_rng(seed)
    # skewed sources (a few hubs) and uniform targets, duplicates and self-loops kept
    src = np.minimum(rng.zipf(1.6, m) - 1, n - 1).astype(np.uint32)
    dst = rng.integers(0, n, m).astype(np.uint32)
    return src, dst


def _multiset_rows(n, rp, col):
    return [sorted(col[rp[u]:rp[u + 1]].tolist()) for u in range(n)]


def test_cpu_path_matches_from_scratch_every_stage():
    n, m = 3000, 40000
    src, dst = _rmat_like(n, m, 4)
    rng = np.random.default_rng(9)
    cpu = oracle.CpuPath(n, src, dst, threads=4)
    queries = [(oracle.APP_SSSP, 0, oracle.W_SEEDED), (oracle.APP_BFS, 17, oracle.W_REFERENCE),
               (oracle.APP_CC, 0, oracle.W_REFERENCE), (oracle.APP_SSSP, 5, oracle.W_REFERENCE)]
    for app, s, w in queries:
        q = cpu.add_query(app, s, w, seed=3, wmax=50)
        cpu.converge(q)
    store = oracle.Store(n, src, dst)

    def check(stage):
        rp, col = store.csr()
        for q, (app, s, w) in enumerate(queries):
            if app == oracle.APP_SSSP:
                want = oracle.sssp(n, rp, col, s, w, 3, 50)
            elif app == oracle.APP_BFS:
                want = oracle.bfs(n, rp, col, s)
            else:
                want = oracle.cc(n, rp, col)
            got, par = cpu.values(q)
            assert np.array_equal(got, want), (stage, q, int(np.sum(got != want)))
            if app != oracle.APP_CC:
                chk = oracle.check_sssp(n, rp, col, s, got, par, w if app == oracle.APP_SSSP else oracle.W_UNIT,
                                        3, 50)
                assert chk["invalid_parents"] == 0 and chk["relaxable_edges"] == 0, (stage, q, chk)
        crp, ccol = cpu.csr()
        assert _multiset_rows(n, crp, ccol) == _multiset_rows(n, rp, col), stage

    check("static")
    for s in range(5):
        rpc, colc = store.csr()
        srcs = np.repeat(np.arange(n), np.diff(rpc).astype(np.int64))
        pick = rng.choice(len(colc), 800, replace=True)          # repeats: duplicate deletes
        dels = np.stack([srcs[pick], colc[pick]], 1).astype(np.int64)
        dels = np.concatenate([dels, [[n + 5, 1], [2, n + 9], [7, 7]]])   # out of range / maybe missing
        adds = np.stack([rng.integers(0, n, 900), rng.integers(0, n, 900)], 1)
        adds[::50, 0] = 1                                        # one row grows past its block
        adds = np.concatenate([adds, [[n + 1, 3]]])
        miss = cpu.delete_phase(dels)
        mm = store.apply(dels=dels, adds=[])
        assert miss == mm["missing_deletes"], (miss, mm)
        check(f"delete {s}")
        cpu.add_phase(adds)
        store.apply(dels=[], adds=adds)
        check(f"add {s}")
    cpu.close()


def test_cpu_path_mutate_only_then_installed_state():
    """A store brought forward by mutations alone, with the oracle's values of
    that point installed, repairs the next snapshots exactly (how the timed CPU
    baseline starts at the first timed snapshot)."""
    n, m = 2000, 25000
    src, dst = _rmat_like(n, m, 6)
    rng = np.random.default_rng(2)
    cpu = oracle.CpuPath(n, src, dst, threads=4)
    q = cpu.add_query(oracle.APP_SSSP, 0, oracle.W_SEEDED, seed=3, wmax=50)
    store = oracle.Store(n, src, dst)

    def batch():
        rpc, colc = store.csr()
        srcs = np.repeat(np.arange(n), np.diff(rpc).astype(np.int64))
        pick = rng.choice(len(colc), 500, replace=False)
        dels = np.stack([srcs[pick], colc[pick]], 1).astype(np.int64)
        adds = np.stack([rng.integers(0, n, 600), rng.integers(0, n, 600)], 1)
        return dels, adds

    for _ in range(3):   # untimed part of the stream: mutations only
        dels, adds = batch()
        cpu.mutate_only(dels, adds)
        store.apply(dels=dels, adds=adds)
    rp, col = store.csr()
    crp, ccol = cpu.csr()
    assert _multiset_rows(n, crp, ccol) == _multiset_rows(n, rp, col)
    want = oracle.sssp(n, rp, col, 0, oracle.W_SEEDED, 3, 50)
    ref = oracle.Incr(n, rp, col, oracle.APP_SSSP, 0, oracle.W_SEEDED, 3, 50)
    ref.converge()
    _, par = ref.values()
    cpu.set_state(q, want, par)
    for s in range(3):
        dels, adds = batch()
        cpu.delete_phase(dels)
        cpu.add_phase(adds)
        store.apply(dels=dels, adds=adds)
        rp, col = store.csr()
        got, _ = cpu.values(q)
        assert np.array_equal(got, oracle.sssp(n, rp, col, 0, oracle.W_SEEDED, 3, 50)), s
    cpu.close()


def test_cpu_path_pagerank_within_tolerance():
    """The all-core PageRank path (amends around the mutation, dense rounds):
    the static solve and three batches each within the FP64 solve's tolerance
    at the probability scale (max-abs 1e-6 and the checker's L1 bound)."""
    n, m = 3000, 30000
    src, dst = _rmat_like(n, m, 11)
    base, eps = 0.15 / n, 1e-7 / n
    cpu = oracle.CpuPath(n, src, dst, threads=4)
    cpu.pr_init(base, eps, 400)
    store = oracle.Store(n, src, dst)
    rng = np.random.default_rng(5)

    def check(stage):
        rp, col = store.csr()
        rank, res = cpu.pr_values()
        chk = oracle.check_pagerank(n, rp, col, rank, res, base=base, eps=eps, max_abs=1e-6)
        assert chk["ok"] and chk["max_error"] <= 1e-6 and chk["l1"] <= chk["bound"], (stage, chk)
        assert float(np.max(np.abs(res))) <= eps, stage

    rounds = cpu.pr_batch([], [])
    assert 0 < rounds < 400
    check("static")
    for s in range(3):
        rpc, colc = store.csr()
        srcs = np.repeat(np.arange(n), np.diff(rpc).astype(np.int64))
        pick = rng.choice(len(colc), 300, replace=False)
        dels = np.stack([srcs[pick], colc[pick]], 1).astype(np.int64)
        adds = np.stack([rng.integers(0, n, 300), rng.integers(0, n, 300)], 1)
        assert cpu.pr_batch(dels, adds) < 400
        store.apply(dels=dels, adds=adds)
        check(f"batch {s}")
    cpu.close()
