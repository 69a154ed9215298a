"""Incremental streaming-graph benchmark (one GPU).

A step is one snapshot of the update stream: ingest of its records + delete stage +
add stage + cache maintenance (the reference's P0-TIMER window,
samples/hybrid_sssp/hybrid_sssp.cu:444-548).  The stream sits in pinned host memory;
each stage copies its records to the device inside the stage, or waits for the copy
that the previous snapshot started on the copy stream (double buffering).  `value`
is updates/s over the timed snapshots; p50/p90 are per-snapshot repair latencies.

Default: c4, the configuration the headline metric (learned vs LRU cache) is
quoted on.  Workloads (seeded, generated in-process):
  c1  RMAT-16 (65,536 v, 1,048,576 e), SSSP, weights 1+mix64 % 100, source 0,
      1,000 deletes + 1,000 inserts per snapshot, topology in HBM.
  c2  RMAT-22 (4,194,304 v, 67,108,864 e), PageRank (probability scale), 50k+50k per snapshot.
  c3  RMAT-26, SSSP + BFS + CC batched in one engine (sources drawn from the edge list with a
      fixed seed), 4e9-byte device chunk cache over pinned host memory, hotness/LRU policy,
      500k+500k per snapshot.
  c4  c3 with the learned cache policy.  Measured beside it on the same graph and stream:
      the batched engine with hotness/LRU, and the plain form (one single-query engine per
      program, hotness/LRU, run one after the other; its snapshot time is the sum).
  c5  4096x4096 grid, 32 SSSP + 32 BFS + CC in one engine, lognormal batch sizes.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "incremental updates/s + p50 repair latency/batch; learned vs LRU cache, 1 GPU"
HBM_PEAK_GBS = 8000.0
# timing category -> device kernel (rocprofv3 names)
# the closures are instantiated per program set; the built-in set is SSSP, BFS, CC
_PS = "dcg::ProgramSet<dcg::apps::SSSP, dcg::apps::BFS, dcg::apps::CC>"
KERNEL_SYMBOL = {"push": f"k_persistent<{_PS}, 0u, 0u>", "pull": f"k_persistent<{_PS}, 1u, 0u>",
                 "invalidate": f"k_persistent<{_PS}, 2u, 0u>",
                 "pr_solve": f"k_persistent<{_PS}, 3u, 0u> (push) / k_prd_consume + k_prd_pull_rows + "
                             "k_prd_pull_big + k_prd_rmax (dense sweeps)",
                 "pr_amend": f"k_persistent<{_PS}, 4u, 0u>", "delete_rows": "k_del_rows", "delete_big": "k_del_big",
                 "append_rows": "k_add_rows + k_add_big",
                 "static": f"k_persistent<{_PS}, 0u, 1u> / k_persistent<{_PS}, 3u, 1u>"}
# records that are not one kernel's launches: several kernels together, or copies
NOT_A_KERNEL = ("delete_pass", "repair_delete", "repair_add", "ingest", "static")
# HBM traffic per launch measured with rocprofv3 PMC counters (FETCH_SIZE and
# WRITE_SIZE runs, scaled by the counters of a copy kernel of known size)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic.json")


def host_link(k, host_lps, delete_pass=None):
    """The dominant kernel against the host link's random-read ceiling: 64-byte lines
    it read from pinned host memory, their rate, and the measured ceiling.  The delete
    mutation's out-row and in-row passes run concurrently, so for them the rate is
    taken over the wall time of both passes together (one record per snapshot)."""
    src, scope = k, "per launch"
    if delete_pass and delete_pass.get("launches") and delete_pass.get("host_lines"):
        src, scope = delete_pass, "both delete passes of a snapshot, run concurrently (wall time)"
    if not src.get("host_lines") or not src["ms"]:
        return None
    rate = src["host_lines"] / (src["ms"] / 1e3)
    return {"bound": "random 64-byte reads of pinned host memory", "scope": scope,
            "lines_per_unit": round(src["host_lines"] / max(src["launches"], 1)),
            "ms_per_unit": round(src["ms"] / max(src["launches"], 1), 4), "achieved_lines_per_s": round(rate),
            "peak_lines_per_s": round(host_lps), "frac": round(rate / host_lps, 4) if host_lps else None}


def hbm_lines(k, hbm_lps):
    """A repair closure against HBM's random-line ceiling: each expanded row
    reads its descriptor and its adjacency (at least two 64-byte lines at random
    places) and each adjacency entry the destination's state slot (one line at a
    random place), so (2 x expansions + edges) are the random lines the launches
    touch when no line is shared (neighbouring slots that share a line make the
    true count smaller); their rate against the ceiling the same process
    measured."""
    if not k.get("edges") or not k["ms"] or not hbm_lps:
        return None
    lines = 2 * k["expansions"] + k["edges"]
    rate = lines / (k["ms"] / 1e3)
    return {"bound": "random 64-byte accesses to HBM", "lines_per_launch": round(lines / max(k["launches"], 1)),
            "achieved_lines_per_s": round(rate), "peak_lines_per_s": round(hbm_lps), "frac": round(rate / hbm_lps, 4)}


def measured_traffic(config, kernel, lines_per_launch=None):
    """HBM bytes per launch from the committed counter runs: the measured bytes
    per launch, or, for a closure whose counter runs covered other launches than
    this run's, their bytes per random line times this run's lines per launch."""
    try:
        with open(TRAFFIC_FILE) as f:
            entries = json.load(f)
    except (OSError, ValueError):
        return None
    for e in entries:
        if e.get("config") == config and e.get("kernel") == kernel:
            if e.get("bytes_per_line") and lines_per_launch:
                return round(e["bytes_per_line"] * lines_per_launch)
            return e.get("bytes_per_launch")
    return None

CONFIGS = {
    "c1": dict(workload="RMAT-16 incremental SSSP (device-resident topology)", scale=16, edges=1 << 20,
               app="sssp", adds=1000, dels=1000, placement=0, policy=1, cache_bytes=0, snapshots=10),
    "c2": dict(workload="RMAT-22 incremental PageRank (device-resident topology)", scale=22, edges=1 << 26,
               app="pr", adds=50000, dels=50000, placement=0, policy=1, cache_bytes=0, snapshots=20),
    "c3": dict(workload="RMAT-26 incremental SSSP+BFS+CC batched, 4e9-byte chunk cache, hotness/LRU", scale=26,
               edges=1 << 30, app="sssp", programs=("sssp", "bfs", "cc"), adds=500000, dels=500000, placement=1,
               policy=1, cache_bytes=4_000_000_000, snapshots=25, warmup=5),
    "c4": dict(workload="RMAT-26 incremental SSSP+BFS+CC batched, 4e9-byte chunk cache, learned policy",
               scale=26, edges=1 << 30, app="sssp", programs=("sssp", "bfs", "cc"), adds=500000, dels=500000,
               placement=1, policy=2, cache_bytes=4_000_000_000, snapshots=25, warmup=5),
    "c5": dict(workload="4096x4096 grid + 5% diagonals, 32 SSSP + 32 BFS sources + CC in one engine, "
                        "lognormal(8,1.5) batches", grid=4096, app="sssp", placement=0, policy=1, cache_bytes=0,
               # the first snapshot after the static convergence re-settles most
               # distances (its first shortcut edges); the stage autotuner runs on
               # the second, the last warm-up snapshot
               snapshots=30, queries=64, warmup=2),
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def batch_sizes(cfg, snapshots, seed=1):
    """(adds per snapshot, deletes per snapshot)."""
    import dycgraph_amd as d
    if "grid" in cfg:
        tot = d.gen_lognormal_sizes(snapshots, 8.0, 1.5, 100, 100000, seed=seed + 3).astype(np.int64)
        return tot - tot // 2, tot // 2
    return np.full(snapshots, cfg["adds"], np.int64), np.full(snapshots, cfg["dels"], np.int64)


def make_inputs(cfg, snapshots, seed=1):
    import dycgraph_amd as d
    t0 = time.time()
    if "grid" in cfg:
        src, dst = d.gen_grid(cfg["grid"], 0.05, seed=seed)
        n = cfg["grid"] * cfg["grid"]
    else:
        src, dst = d.gen_rmat(cfg["scale"], cfg["edges"], seed=seed)
        n = 1 << cfg["scale"]
    ap, dp = batch_sizes(cfg, snapshots, seed)
    adds, dels = d.gen_stream(n, src, dst, ap, dp, seed=seed + 6)
    log(f"inputs: n={n} m={len(src)} snapshots={snapshots} updates={int(ap.sum() + dp.sum())} "
        f"({time.time() - t0:.1f}s)")
    return n, src, dst, adds, dels, ap, dp


def queries_of(cfg, n, src=None, seed=1):
    """(sources, programs): one SSSP query from vertex 0; for `programs` one query per
    program whose source is the tail of a seeded random edge (so it has out-edges; CC
    ignores its source); for the grid 32 SSSP + 32 BFS seeded sources plus one CC query."""
    if "programs" in cfg:
        rng = np.random.default_rng(seed + 23)
        picks = [int(src[i]) for i in rng.integers(0, len(src), len(cfg["programs"]))]
        return tuple(0 if p == "cc" else v for p, v in zip(cfg["programs"], picks)), tuple(cfg["programs"])
    if "queries" not in cfg:
        return (0,), ()
    rng = np.random.default_rng(seed + 17)
    src = [int(x) for x in rng.integers(0, n, cfg["queries"])]
    half = cfg["queries"] // 2
    return tuple(src + [0]), tuple(["sssp"] * half + ["bfs"] * (cfg["queries"] - half) + ["cc"])


def engine_config(cfg, n, policy, sources, programs, placement=None):
    import dycgraph_amd as d
    wmax = 1000 if "grid" in cfg else 100
    app = cfg["app"] if not programs or len(programs) > 1 else programs[0]
    return d.EngineConfig(app=app, sources=sources, query_apps=programs if len(programs) > 1 else (),
                          weight_mode=d.W_SEEDED, weight_max=wmax, weight_seed=11,
                          placement=cfg["placement"] if placement is None else placement, cache_policy=policy,
                          cache_bytes=cfg["cache_bytes"], pr_base=0.15 / n, pr_epsilon=1e-7 / n, pr_max_rounds=100)


def run_gpu(cfg, n, src, dst, adds, dels, ap, dp, warmup, steps, policy, queries=None, keep_initial=None):
    """keep_initial: query indices whose (values, parents) are gathered after the
    warm-up snapshots, where the timed snapshots (and the CPU baseline) start."""
    import dycgraph_amd as d
    pr = cfg["app"] == "pr"
    sources, programs = queries if queries is not None else queries_of(cfg, n, src)
    ec = engine_config(cfg, n, policy, sources, programs)
    eng = d.Engine(ec)
    if cfg.get("buckets") is not None:
        eng.set_tuning(bucket_quarters=cfg["buckets"])
    t0 = time.time()
    eng.load_graph(n, src, dst)
    eng.start()
    eng.load_updates(adds, dels)
    eng.sync()
    log(f"load+start {time.time() - t0:.1f}s")
    aoff = np.concatenate([[0], np.cumsum(ap)]).astype(np.int64)
    doff = np.concatenate([[0], np.cumsum(dp)]).astype(np.int64)
    lat, hits, host = [], 0, 0
    closure = {"affected": 0, "rounds": 0, "edges_scanned": 0, "bucket_advances": 0, "far_entries": 0}

    last = len(ap) - 1

    def snapshot(s):
        ab, db, na, nd = int(aoff[s]), int(doff[s]), int(ap[s]), int(dp[s])
        if s < last:   # the next snapshot's records go to the device beside this one
            eng.prefetch_updates(int(aoff[s + 1]), int(ap[s + 1]), int(doff[s + 1]), int(dp[s + 1]))
        if pr:
            st = eng.update_pagerank(ab, na, db, nd)
            phases = [st]
        else:
            if policy == d.CACHE_LEARNED:
                eng.cache_prefetch(ab, na, db, nd)
            phases = list(eng.snapshot(db, nd, ab, na))
        if cfg["placement"] == 1:   # a device-resident store has no chunk cache to maintain
            eng.cache_maintain()
        return phases

    for s in range(warmup):
        if s == warmup - 1 and cfg.get("tune") and not pr:
            # the stage autotuner: the last warm-up snapshot's delete and add
            # stages each run their repair once per candidate, from the same state
            eng.autotune_stages(2)
        snapshot(s)
    eng.sync()
    # the state at the first timed snapshot: the CPU baseline starts from it
    # (PageRank: ranks and residuals)
    if keep_initial and pr:
        initial = {0: (eng.gather_value(), eng.gather_buffer())}
    elif keep_initial:
        initial = {q: (eng.gather_value(q), eng.gather_parent(q)) for q in keep_initial}
    else:
        initial = None
    tune = ({"delete_stage": eng.stage_tune(0), "add_stage": eng.stage_tune(1)}
            if cfg.get("tune") and not pr and warmup else None)
    eng.reset_kernel_stats()
    t_all = time.perf_counter()
    for s in range(warmup, warmup + steps):
        t = time.perf_counter()
        phases = snapshot(s)
        eng.sync()
        lat.append((time.perf_counter() - t) * 1e3)
        for p in phases:
            hits += p["cached_edge_reads"]
            host += p["host_edge_reads"]
            for k in closure:
                closure[k] += p.get(k, 0)
    total = time.perf_counter() - t_all
    ks = eng.kernel_stats()
    info = eng.info()
    values = [eng.gather_value(q) for q in range(eng.queries)]
    return dict(total_s=total, lat=lat, ks=ks, info=info, hits=hits, host=host, eng=eng, values=values,
                queries=(sources, programs), closure={k: round(v / steps, 1) for k, v in closure.items()},
                initial=initial, tune=tune)


def run_plain(cfg, n, src, dst, adds, dels, ap, dp, warmup, steps, queries):
    """The plain GPU form: one single-query engine per program, hotness/LRU cache, each
    applying the whole stream; engines run one after the other and a snapshot of the
    plain form costs the sum of their snapshot times."""
    import dycgraph_amd as d
    total, lat, hits, host, values = 0.0, np.zeros(steps), 0, 0, []
    for s_q, p_q in zip(*queries):
        r = run_gpu(cfg, n, src, dst, adds, dels, ap, dp, warmup, steps, d.CACHE_HOTNESS, ((s_q,), (p_q,)))
        r["eng"].close()
        total += r["total_s"]
        lat += np.array(r["lat"])
        hits += r["hits"]
        host += r["host"]
        values += r["values"]
        log(f"plain {p_q} from {s_q}: {r['total_s'] * 1e3:.1f} ms")
    return dict(total_s=total, lat=list(lat), hits=hits, host=host, values=values)


def verify_from_scratch(cfg, n, eng, queries, values):
    """Exactness at full size: rebuild an engine from the incremental engine's final
    topology (device placement), converge it from scratch and compare every query's
    values bit for bit."""
    import dycgraph_amd as d
    rp, col = eng.csr()
    srcs = np.repeat(np.arange(n, dtype=np.uint32), np.diff(rp).astype(np.int64))
    ec = engine_config(cfg, n, d.CACHE_HOTNESS, queries[0], queries[1], placement=d.PLACE_DEVICE)
    with d.Engine(ec) as fresh:
        fresh.load_graph(n, srcs, col)
        fresh.start()
        fresh.sync()
        return all(np.array_equal(fresh.gather_value(q), values[q]) for q in range(fresh.queries))


def pagerank_accuracy(n, src, dst, adds, dels, ap, dp, eng):
    """The last snapshot's FP32 ranks against the oracle's FP64 from-scratch power
    iteration on the oracle's final multiset: max-abs and L1 error, the largest
    residual, and the reference checker's L1 bound (hybrid_pr_host.cu:25-56) at the
    probability scale (base 0.15/n, eps 1e-7/n)."""
    import oracle
    t = time.time()
    store = oracle.Store(n, src, dst)
    aoff = np.concatenate([[0], np.cumsum(ap)]).astype(np.int64)
    doff = np.concatenate([[0], np.cumsum(dp)]).astype(np.int64)
    for s in range(len(ap)):
        store.apply(dels[doff[s]:doff[s + 1]], [])
        store.apply([], adds[aoff[s]:aoff[s + 1]])
    rp, col = store.csr()
    rank, res = eng.gather_value(), eng.gather_buffer()
    chk = oracle.check_pagerank(n, rp, col, rank, res, base=0.15 / n, eps=1e-7 / n, max_abs=1e-6)
    return {"max_abs": chk["max_error"], "l1": chk["l1"], "l1_bound": chk["bound"],
            "max_residual": chk["max_residual"], "eps": 1e-7 / n, "invariant_l1": chk["defect"],
            "invariant_bound": chk["roundoff"], "max_abs_bound": 1e-6, "within_bounds": bool(chk["ok"]),
            "mean_rank": float(np.mean(rank)), "check_s": round(time.time() - t, 1),
            "reference": "FP64 Jacobi power iteration to sum|dx| <= 1e-12 (probability scale)"}


def host_threads():
    """Host cores this process may use: the CPU affinity set, capped by
    OMP_NUM_THREADS when that is set."""
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        cores = min(cores, int(env))
    return max(1, cores)


def cpu_baseline(cfg, n, src, dst, adds, dels, ap, dp, queries, initial, start, budget_s=20.0):
    """The reference's CPU path on all host cores (oracle/cpu_path.c: grouped row
    mutation by a worker pool, parallel invalidation, pull repair and push
    closures; the queries over one store) on the same graph and stream, timed on
    the same snapshots as the GPU: the store is brought to the first timed
    snapshot by the warm-up snapshots' mutations alone (untimed), the GPU's state
    at that point is installed, and snapshots from there are timed one by one
    until ~budget_s of CPU work.  `initial` maps the query indices that run to
    their state; with fewer queries than the GPU ran (the grid), the rate is
    scaled by the query count (the repair dominates and is per query)."""
    import oracle
    codes = {"sssp": oracle.APP_SSSP, "bfs": oracle.APP_BFS, "cc": oracle.APP_CC}
    sources, programs = queries
    programs = programs or (cfg["app"],) * len(sources)
    threads = host_threads()
    wmax = 1000 if "grid" in cfg else 100
    t0 = time.time()
    cpu = oracle.CpuPath(n, src, dst, threads=threads)
    aoff = np.concatenate([[0], np.cumsum(ap)]).astype(np.int64)
    doff = np.concatenate([[0], np.cumsum(dp)]).astype(np.int64)
    for s in range(start):
        cpu.mutate_only(dels[doff[s]:doff[s + 1]], adds[aoff[s]:aoff[s + 1]])
    sample = sorted(initial)
    for q in sample:
        k = cpu.add_query(codes[programs[q]], sources[q], oracle.W_SEEDED, 11, wmax)
        cpu.set_state(k, initial[q][0], initial[q][1])
    build_s = time.time() - t0
    done, spent, upd, times = start, 0.0, 0, []
    while done < len(ap) and (done == start or spent < budget_s):
        t = time.perf_counter()
        cpu.delete_phase(dels[doff[done]:doff[done + 1]])
        cpu.add_phase(adds[aoff[done]:aoff[done + 1]])
        dt = time.perf_counter() - t
        times.append(round(dt * 1e3, 1))
        spent += dt
        upd += int(ap[done] + dp[done])
        done += 1
        log(f"cpu snapshot {done - 1}: {dt * 1e3:.1f} ms ({threads} threads, {len(sample)} queries)")
    cpu.close()
    rate = upd / spent if spent else 0.0
    scale = len(sample) / len(sources)
    names = "+".join(programs[q] for q in sample)
    text = (f"the same graph and stream: snapshots {start}..{done - 1} (the GPU's first timed snapshots; "
            f"{upd} updates; per-snapshot ms {times}), from the GPU's state at snapshot {start} over a store "
            f"brought there by the earlier snapshots' mutations (untimed; store built in {build_s:.1f}s); "
            f"oracle/cpu_path.c on {threads} threads, queries {names} over one store")
    if scale < 1:
        text += (f"; {len(sample)} of the {len(sources)} queries ran: value = their rate x {len(sample)}/"
                 f"{len(sources)} (measured {rate:.1f} updates/s for the sample)")
    return dict(value=rate * scale, unit="updates/s", cores=threads, kind="port", nproc=os.cpu_count(),
                sample=text)


def cpu_baseline_pr(cfg, n, src, dst, adds, dels, ap, dp, initial, start, budget_s=20.0):
    """PageRank on all host cores (oracle/cpu_path.c orc_cpu_pr_*: the amends
    around the batch's mutation, then dense Jacobi rounds, the engine's
    large-batch path) on the same graph and snapshots as the GPU, from the GPU's
    ranks and residuals at the first timed snapshot over a store brought there
    by the earlier snapshots' mutations (untimed)."""
    import oracle
    threads = host_threads()
    t0 = time.time()
    cpu = oracle.CpuPath(n, src, dst, threads=threads)
    aoff = np.concatenate([[0], np.cumsum(ap)]).astype(np.int64)
    doff = np.concatenate([[0], np.cumsum(dp)]).astype(np.int64)
    for s in range(start):
        cpu.mutate_only(dels[doff[s]:doff[s + 1]], adds[aoff[s]:aoff[s + 1]])
    cpu.pr_init(0.15 / n, 1e-7 / n, 100)
    cpu.pr_set_state(initial[0][0], initial[0][1])
    build_s = time.time() - t0
    done, spent, upd, times, rounds = start, 0.0, 0, [], []
    while done < len(ap) and (done == start or spent < budget_s):
        t = time.perf_counter()
        rounds.append(cpu.pr_batch(dels[doff[done]:doff[done + 1]], adds[aoff[done]:aoff[done + 1]]))
        dt = time.perf_counter() - t
        times.append(round(dt * 1e3, 1))
        spent += dt
        upd += int(ap[done] + dp[done])
        done += 1
        log(f"cpu snapshot {done - 1}: {dt * 1e3:.1f} ms ({threads} threads, {rounds[-1]} rounds)")
    cpu.close()
    return dict(value=upd / spent if spent else 0.0, unit="updates/s", cores=threads, kind="port",
                nproc=os.cpu_count(),
                sample=f"the same graph and stream: snapshots {start}..{done - 1} (the GPU's first timed snapshots; "
                       f"{upd} updates; per-snapshot ms {times}; dense rounds {rounds}), from the GPU's ranks and "
                       f"residuals at snapshot {start} over a store brought there by the earlier snapshots' "
                       f"mutations (untimed; store built in {build_s:.1f}s); oracle/cpu_path.c PageRank on "
                       f"{threads} threads (amends, mutation, dense Jacobi rounds to eps)")


def cpu_single_thread(cfg, n, src, dst, adds, dels, ap, dp, queries, initial, start, budget_s=5.0):
    """The single-threaded incremental port (oracle.c orc_incr_*) of the one SSSP
    query on the same snapshots as cpu_baseline."""
    import oracle
    sources, _ = queries
    store = oracle.Store(n, src, dst)
    aoff = np.concatenate([[0], np.cumsum(ap)]).astype(np.int64)
    doff = np.concatenate([[0], np.cumsum(dp)]).astype(np.int64)
    for s in range(start):
        store.apply(dels[doff[s]:doff[s + 1]], adds[aoff[s]:aoff[s + 1]])
    rp, col = store.csr()
    inc = oracle.Incr(n, rp, col, oracle.APP_SSSP, sources[0], oracle.W_SEEDED, 11, 100)
    inc.set_state(initial[0][0], initial[0][1])
    done, spent, upd = start, 0.0, 0
    while done < len(ap) and (done == start or spent < budget_s):
        t = time.perf_counter()
        inc.delete_phase(dels[doff[done]:doff[done + 1]])
        inc.add_phase(adds[aoff[done]:aoff[done + 1]])
        spent += time.perf_counter() - t
        upd += int(ap[done] + dp[done])
        done += 1
    return {"value": round(upd / spent, 1) if spent else 0.0, "unit": "updates/s", "cores": 1, "kind": "port",
            "sample": f"oracle.c's single-threaded incremental port, snapshots {start}..{done - 1} from the same "
                      f"state as cpu_baseline"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed snapshots first (default: 5 for the RMAT-26 streams, whose cache policies "
                         "settle over the first snapshots; 1 otherwise)")
    ap.add_argument("--config", default=os.environ.get("DCG_BENCH_CONFIG", "c4"), choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true",
                    help="skip the exactness check of the final state (from-scratch convergence or the oracle)")
    ap.add_argument("--no-compare", action="store_true",
                    help="with the learned policy, skip the hotness/LRU run on the same stream")
    ap.add_argument("--buckets", type=int, default=None,
                    help="push-closure bucket width in quarters (0 off; default: the engine's automatic choice)")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the stage autotuner on the last warm-up snapshot")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    cfg = dict(CONFIGS[args.config], buckets=args.buckets, tune=not args.no_tune and args.buckets is None)
    warmup = args.warmup if args.warmup is not None else cfg.get("warmup", 1)
    steps = args.steps if args.steps is not None else cfg["snapshots"] - warmup
    snapshots = warmup + steps
    n, src, dst, adds, dels, ap, dp = make_inputs(cfg, snapshots)
    import dycgraph_amd as d
    hbm_gbs, host_gbs = d.calibrate()
    host_lps, hbm_lps = d.calibrate_lines()
    log(f"calibration: HBM copy {hbm_gbs:.0f} GB/s, pinned-host zero-copy read {host_gbs:.1f} GB/s, "
        f"random 64-byte lines: host {host_lps / 1e6:.0f} M/s, HBM {hbm_lps / 1e6:.0f} M/s")
    cpu_wanted = not args.no_cpu
    nq = len(queries_of(cfg, n, src)[0])
    # the grid's 65 queries would keep the CPU path busy for minutes per
    # snapshot: it runs 2 SSSP + 2 BFS queries and the CC query
    cpu_sample = [0, 1, (nq - 1) // 2, (nq - 1) // 2 + 1, nq - 1] if "grid" in cfg else list(range(nq))
    res = run_gpu(cfg, n, src, dst, adds, dels, ap, dp, warmup, steps, cfg["policy"],
                  keep_initial=cpu_sample if cpu_wanted else None)
    upd = int(ap[warmup:].sum() + dp[warmup:].sum())
    value = upd / res["total_s"]
    lat = np.array(res["lat"])
    ks = res["ks"]
    # records spanning several kernels (or copies) are not candidates
    dom = max(((n, v) for n, v in ks.items() if n not in NOT_A_KERNEL), key=lambda kv: kv[1]["ms"])
    kname, k = dom
    achieved = (k["algo_bytes"] / (k["ms"] / 1e3) / 1e9) if k["ms"] > 0 else 0.0
    reads = res["hits"] + res["host"]
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "updates/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": round(res["total_s"] * 1e3 / steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32" if cfg["app"] == "pr" else "u32",
        "data": "synthetic (seeded RMAT graph and update stream generated in-process)",
        "config": {"workload": cfg["workload"], "vertices": n, "edges": len(src), "app": cfg["app"],
                   "updates_per_snapshot": round(upd / steps, 1), "placement":
                   "device" if cfg["placement"] == 0 else "pinned-host+device-cache",
                   "cache_bytes": cfg["cache_bytes"], "policy": ["none", "hotness", "learned"][cfg["policy"]],
                   "parallelism": "single GPU"},
        "p50_ms": round(float(np.percentile(lat, 50)), 3), "p90_ms": round(float(np.percentile(lat, 90)), 3),
        "cache_hit_rate": round(res["hits"] / reads, 4) if reads else None,
        # device memory held by the engine at the end of the run (bytes): the
        # chunk cache's frames (the capped part), the rest of the store in HBM
        # (row descriptors, sorted-prefix lengths, fences, read counters), the
        # per-query state, and everything together with batch scratch
        "device_footprint": {"frames": res["info"]["frame_bytes"],
                             "store": res["info"]["store_device_bytes"],
                             "state": res["info"]["state_device_bytes"],
                             "total": res["info"]["device_bytes"],
                             "host_store": res["info"]["host_bytes"]},
        "host_link_edge_bytes": int(res["host"] * 4),
        "calibration": {"hbm_copy_GBps": round(hbm_gbs, 1), "host_zero_copy_read_GBps": round(host_gbs, 2),
                        "host_random_lines_per_s": round(host_lps), "hbm_random_lines_per_s": round(hbm_lps)},
        "roofline": {"kernel": kname, "hip_kernel": KERNEL_SYMBOL.get(kname, kname), "bound": "hbm",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(args.config, kname, (hbm_lines(k, hbm_lps) or {}).get("lines_per_launch")),
                     "launches": k["launches"], "avg_launch_ms": round(k["ms"] / max(k["launches"], 1), 4),
                     "algo_bytes_per_launch": round(k["algo_bytes"] / max(k["launches"], 1), 1),
                     "host_link": host_link(k, host_lps, ks.get("delete_pass") if kname == "delete_rows" else None),
                     "hbm_lines": hbm_lines(k, hbm_lps)},
        "per_snapshot": res["closure"],
        # parameters the stage autotuner chose on the last warm-up snapshot, per
        # stage (best_ms / first_ms: the tuned and the starting parameters'
        # repair time of that snapshot's delete / add stage)
        "autotune": res.get("tune"),
        # per timed snapshot: each record's time, launches and algorithmic rate
        # (stage records span all of a stage's repair launches; ingest is the
        # records' host-to-device copies)
        "kernels": {name: {"ms_per_snapshot": round(v["ms"] / steps, 4), "launches": v["launches"],
                           "avg_launch_ms": round(v["ms"] / v["launches"], 4),
                           "GBps": round(v["algo_bytes"] / (v["ms"] / 1e3) / 1e9, 1) if v["ms"] else 0.0,
                           "symbol": KERNEL_SYMBOL.get(name, "several kernels" if name in NOT_A_KERNEL[:3]
                                                       else "hipMemcpyAsync")}
                    for name, v in ks.items() if v["launches"]},
    }
    queries = res["queries"]
    if "programs" in cfg or "queries" in cfg:
        out["config"]["queries"] = [f"{p}@{s}" for s, p in zip(*queries)]
        out["config"]["app"] = "+".join(queries[1])
    vals_main = res["values"]

    def summary(r, label):
        reads = r["hits"] + r["host"]
        v = upd / r["total_s"]
        return {"form": label, "value": round(v, 1),
                "p50_ms": round(float(np.percentile(r["lat"], 50)), 3),
                "p90_ms": round(float(np.percentile(r["lat"], 90)), 3),
                "cache_hit_rate": round(r["hits"] / reads, 4) if reads else None,
                "host_link_edge_bytes": int(r["host"] * 4),
                "headline_over_this": round(value / v, 4) if v else None,
                "identical_values": all(np.array_equal(a, b) for a, b in zip(vals_main, r["values"]))}

    if cfg["policy"] == 2 and not args.no_compare:
        # the hotness/LRU policy on the identical graph, stream and batched queries
        lru = run_gpu(cfg, n, src, dst, adds, dels, ap, dp, warmup, steps, 1, queries)
        lru["eng"].close()
        out["compare_lru"] = summary(lru, "batched engine, hotness/LRU cache")
    if "programs" in cfg and not args.no_compare:
        plain = run_plain(cfg, n, src, dst, adds, dels, ap, dp, warmup, steps, queries)
        out["compare_plain"] = summary(plain, "plain: one single-query engine per program, hotness/LRU cache")
    # the from-scratch check runs after the comparison runs (a large device-placement
    # engine before them slowed the next engine by ~3%); the measured engine stays open
    if cfg["app"] == "pr":
        out["pagerank_accuracy"] = pagerank_accuracy(n, src, dst, adds, dels, ap, dp, res["eng"])
        log(f"pagerank accuracy: {out['pagerank_accuracy']}")
    if cfg["app"] != "pr" and not args.no_check:
        if "programs" in cfg or "queries" in cfg:
            t = time.time()
            out["exact"] = verify_from_scratch(cfg, n, res["eng"], queries, res["values"])
            out["exact_check"] = "final values of every query == from-scratch GPU convergence on the final topology"
            log(f"from-scratch check: {out['exact']} ({time.time() - t:.1f}s)")
        else:
            import oracle
            store = oracle.Store(n, src, dst)
            aoff = np.concatenate([[0], np.cumsum(ap)]).astype(np.int64)
            doff = np.concatenate([[0], np.cumsum(dp)]).astype(np.int64)
            for s in range(snapshots):
                store.apply(dels[doff[s]:doff[s + 1]], adds[aoff[s]:aoff[s + 1]])
            rp, col = store.csr()
            want = oracle.sssp(n, rp, col, 0, oracle.W_SEEDED, 11, 100)
            out["exact"] = bool(np.array_equal(res["values"][0], want))
            out["exact_check"] = "final SSSP distances == oracle Dijkstra on the oracle's final multiset"
    res["eng"].close()
    cpu = None
    if cpu_wanted:
        if cfg["app"] == "pr":
            cpu = cpu_baseline_pr(cfg, n, src, dst, adds, dels, ap, dp, res["initial"], warmup)
        else:
            cpu = cpu_baseline(cfg, n, src, dst, adds, dels, ap, dp, res["queries"], res["initial"], warmup)
        out["cpu_baseline"] = {k: (round(v, 1) if isinstance(v, float) else v) for k, v in cpu.items()}
        out["gpu_over_cpu"] = round(value / cpu["value"], 1) if cpu["value"] else None
        if n <= 1 << 20 and "grid" not in cfg and cfg["app"] != "pr":
            # small graphs: the single-threaded incremental port beside the
            # all-core path (its per-level thread overhead dominates tiny batches)
            st1 = cpu_single_thread(cfg, n, src, dst, adds, dels, ap, dp, res["queries"], res["initial"], warmup)
            out["cpu_single_thread"] = st1
            out["gpu_over_cpu_single_thread"] = round(value / st1["value"], 2) if st1["value"] else None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
