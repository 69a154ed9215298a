"""PageRank at the probability scale against an FP64 from-scratch power iteration.

Ranks are FP32 on the device; the acceptance is the reference's checker
(samples/hybrid_pr/hybrid_pr_host.cu:9-56: converged FP64 solve, invariant
||(base + 0.85 P^T x) - x - r||_1 <= 64 FLT_EPSILON max(1, mass), every
|residual| <= eps, L1 <= (n eps + roundoff) / 0.15) plus max |rank - FP64| <= 1e-6.
Scale: base 0.15/n and eps 1e-7/n (ranks sum to 1), after the static solve and
after every batch, on both solve paths: batches that leave few vertices above eps
run the residual push (k_persistent<3>), batches that leave more than 1/32 of
them run dense sweeps (k_prd_*).
"""
import numpy as np
import pytest

import oracle
from tests.conftest import announce
from tests.helpers import rmat_graph

pytestmark = pytest.mark.gpu

MAX_ABS = 1e-6


def _accept(chk, converged, what):
    """North-star bounds at every stage (max-abs, the checker's L1 bound, the
    invariant); the residual threshold once the solve finished below the round
    cap (the reference caps every solve at pr_max_rounds too, framework.cuh:2681-2736,
    and carries the unfinished frontier into the next batch)."""
    assert chk["finite"] and chk["converged"], (what, chk)
    assert chk["max_error"] <= MAX_ABS, (what, chk)
    assert chk["l1"] <= chk["bound"], (what, chk)
    assert chk["defect"] <= chk["roundoff"], (what, chk)
    if converged:
        assert chk["ok"], (what, chk)


def _low_degree_stream(n, src, dst, per, snaps, seed):
    """Batches whose records all leave low-out-degree sources (at most 4 out-edges),
    so the amend passes leave only a few dozen vertices above eps and the batch
    runs the push solve: deletions of existing edges, insertions of absent pairs."""
    rng = np.random.default_rng(seed)
    deg = np.bincount(src, minlength=n)
    low = np.flatnonzero((deg >= 1) & (deg <= 4))
    present = set(zip(src.tolist(), dst.tolist()))
    by_src = {}
    for u, v in zip(src.tolist(), dst.tolist()):
        if deg[u] <= 4:
            by_src.setdefault(u, []).append(v)
    adds, dels = [], []
    for s in range(snaps):
        while len(dels) < (s + 1) * per:
            u = int(rng.choice(low))
            row = by_src[u]
            if row:
                v = row.pop(int(rng.integers(0, len(row))))
                dels.append((u, v))
                present.discard((u, v))
        while len(adds) < (s + 1) * per:
            u, v = int(rng.choice(low)), int(rng.integers(0, n))
            if u != v and (u, v) not in present:
                adds.append((u, v))
                present.add((u, v))
                by_src[u].append(v)
    return np.array(adds, np.uint32), np.array(dels, np.uint32)


def _solve_path(before, after):
    """'push' if the batch ran the persistent push solve (it expands rows and
    counts them), 'dense' if only dense sweeps ran."""
    return "push" if after["expansions"] > before["expansions"] else "dense"


@pytest.mark.parametrize("placement", [0, 1])
@pytest.mark.parametrize("path,per", [("push", 6), ("dense", 1500)])
def test_pagerank_probability_scale(path, per, placement):
    import dycgraph_amd as d
    announce(f"pagerank probability scale, {path} path, {per}+{per} x 3, placement={placement}")
    n, src, dst = rmat_graph(13, 80000, seed=4)
    base, eps = 0.15 / n, 1e-7 / n
    if path == "push":
        adds, dels = _low_degree_stream(n, src, dst, per, 3, seed=6)
    else:
        adds, dels = d.gen_stream(n, src, dst, [per] * 3, [per] * 3, seed=6)
    eng = d.Engine(d.EngineConfig(app="pr", placement=placement, cache_bytes=1 << 18, pr_base=base,
                                  pr_epsilon=eps, pr_max_rounds=100))
    eng.load_graph(n, src, dst)
    eng.start()
    store = oracle.Store(n, src, dst)
    rp, col = store.csr()
    chk = oracle.check_pagerank(n, rp, col, eng.gather_value(), eng.gather_buffer(), base=base, eps=eps,
                                max_abs=MAX_ABS)
    _accept(chk, False, "static")
    eng.load_updates(adds, dels)
    paths = []
    for s in range(3):
        k0 = eng.kernel_stats()["pr_solve"]
        st = eng.update_pagerank(s * per, per, s * per, per)
        paths.append(_solve_path(k0, eng.kernel_stats()["pr_solve"]))
        store.apply(dels[s * per:(s + 1) * per], [])
        store.apply([], adds[s * per:(s + 1) * per])
        rp, col = store.csr()
        rank = eng.gather_value()
        chk = oracle.check_pagerank(n, rp, col, rank, eng.gather_buffer(), base=base, eps=eps, max_abs=MAX_ABS)
        print(f"[pr] batch {s} path={paths[-1]} rounds={st['rounds']} max_abs={chk['max_error']:.3e} "
              f"l1={chk['l1']:.3e} bound={chk['bound']:.3e} max_res={chk['max_residual']:.3e}", flush=True)
        converged = st["rounds"] < 100
        _accept(chk, converged, f"batch {s}")
        if s == 2:
            assert converged, "the last batch's solve reached the round cap"
        # the FP32 ranks are close to FP64 relative to the rank scale itself
        assert chk["max_error"] <= 1e-3 * float(rank.max()), chk
    assert paths.count(path) >= 2, paths
    eng.close()
