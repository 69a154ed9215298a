"""Incremental streaming-graph benchmark (one GPU).

A step is one snapshot of the update stream: delete stage + add stage + cache
maintenance (the reference's P0-TIMER window, samples/hybrid_sssp/hybrid_sssp.cu:444-548),
with the update stream already resident in HBM.  `value` is updates/s over the
timed snapshots; p50/p90 are per-snapshot repair latencies.

Workloads (seeded, generated in-process):
  c1  RMAT-16 (65,536 v, 1,048,576 e), SSSP, weights 1+mix64 % 100, source 0,
      1,000 deletes + 1,000 inserts per snapshot, topology in HBM.
  c2  RMAT-22 (4,194,304 v, 67,108,864 e), PageRank (probability scale), 50k+50k per snapshot.
  c3  RMAT-26, SSSP, 4e9-byte device chunk cache over pinned host memory, hotness/LRU policy,
      500k+500k per snapshot.
  c4  c3 with the learned cache policy (c3 measured beside it for the comparison).
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

CONFIGS = {
    "c1": dict(workload="RMAT-16 incremental SSSP (device-resident topology)", scale=16, edges=1 << 20,
               app="sssp", adds=1000, dels=1000, placement=0, policy=1, cache_bytes=0, snapshots=10),
    "c2": dict(workload="RMAT-22 incremental PageRank (device-resident topology)", scale=22, edges=1 << 26,
               app="pr", adds=50000, dels=50000, placement=0, policy=1, cache_bytes=0, snapshots=20),
    "c3": dict(workload="RMAT-26 incremental SSSP, 4e9-byte chunk cache, hotness/LRU", scale=26,
               edges=1 << 30, app="sssp", adds=500000, dels=500000, placement=1, policy=1,
               cache_bytes=4_000_000_000, snapshots=10),
    "c4": dict(workload="RMAT-26 incremental SSSP, 4e9-byte chunk cache, learned policy", scale=26,
               edges=1 << 30, app="sssp", adds=500000, dels=500000, placement=1, policy=2,
               cache_bytes=4_000_000_000, snapshots=10),
    "c5": dict(workload="4096x4096 grid + 5% diagonals, 32 SSSP + 32 BFS sources + CC in one engine, "
                        "lognormal(8,1.5) batches", grid=4096, app="sssp", placement=0, policy=1, cache_bytes=0,
               snapshots=30, queries=64),
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


def queries_of(cfg, n, seed=1):
    """(sources, programs): one SSSP query from vertex 0, or for the grid 32 SSSP +
    32 BFS seeded sources plus one CC query."""
    if "queries" not in cfg:
        return (0,), ()
    import dycgraph_amd as d
    rng = np.random.default_rng(seed + 17)
    src = [int(x) for x in rng.integers(0, n, cfg["queries"])]
    half = cfg["queries"] // 2
    return tuple(src + [0]), tuple(["sssp"] * half + ["bfs"] * (cfg["queries"] - half) + ["cc"])


def run_gpu(cfg, n, src, dst, adds, dels, ap, dp, warmup, steps, policy):
    import dycgraph_amd as d
    pr = cfg["app"] == "pr"
    sources, programs = queries_of(cfg, n)
    wmax = 1000 if "grid" in cfg else 100
    ec = d.EngineConfig(app=cfg["app"], sources=sources, query_apps=programs, weight_mode=d.W_SEEDED,
                        weight_max=wmax, weight_seed=11, placement=cfg["placement"], cache_policy=policy,
                        cache_bytes=cfg["cache_bytes"], pr_base=0.15 / n, pr_epsilon=1e-7 / n, pr_max_rounds=100)
    eng = d.Engine(ec)
    t0 = time.time()
    eng.load_graph(n, src, dst)
    eng.start()
    eng.load_updates(adds, dels)
    eng.sync()
    log(f"load+start {time.time() - t0:.1f}s")
    aoff = np.concatenate([[0], np.cumsum(ap)]).astype(np.int64)
    doff = np.concatenate([[0], np.cumsum(dp)]).astype(np.int64)
    lat, hits, host = [], 0, 0

    def snapshot(s):
        ab, db, na, nd = int(aoff[s]), int(doff[s]), int(ap[s]), int(dp[s])
        if pr:
            st = eng.update_pagerank(ab, na, db, nd)
            phases = [st]
        else:
            if policy == d.CACHE_LEARNED:
                eng.cache_prefetch(ab, na, db, nd)
            phases = [eng.del_edge(db, nd), eng.add_edge(ab, na)]
        eng.cache_maintain()
        return phases

    for s in range(warmup):
        snapshot(s)
    eng.sync()
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
    total = time.perf_counter() - t_all
    ks = eng.kernel_stats()
    info = eng.info()
    values = eng.gather_value()
    return dict(total_s=total, lat=lat, ks=ks, info=info, hits=hits, host=host, eng=eng, values=values)


def cpu_baseline(cfg, n, src, dst, adds, dels, warmup, budget_s=20.0, note=""):
    """The CPU port of the reference's incremental path (oracle/), 1 thread, on the
    same stream: snapshots after the warmup ones until ~budget_s of CPU work."""
    import oracle
    rp, col = oracle.csr_from_edges(n, src, dst)
    app = {"sssp": oracle.APP_SSSP, "bfs": oracle.APP_BFS, "cc": oracle.APP_CC, "pr": oracle.APP_PR}[cfg["app"]]
    inc = oracle.Incr(n, rp, col, app, source=0, wmode=oracle.W_SEEDED, seed=11, wmax=100, pr_base=0.15 / n,
                      pr_eps=1e-7 / n, pr_max_rounds=100)
    inc.converge()
    na, nd = cfg["adds"], cfg["dels"]
    for s in range(warmup):
        if app == oracle.APP_PR:
            inc.pr_batch(dels[s * nd:(s + 1) * nd], adds[s * na:(s + 1) * na])
        else:
            inc.delete_phase(dels[s * nd:(s + 1) * nd])
            inc.add_phase(adds[s * na:(s + 1) * na])
    done, spent = 0, 0.0
    s = warmup
    while spent < budget_s and (s + 1) * max(na, nd) <= len(adds):
        t = time.perf_counter()
        if app == oracle.APP_PR:
            inc.pr_batch(dels[s * nd:(s + 1) * nd], adds[s * na:(s + 1) * na])
        else:
            inc.delete_phase(dels[s * nd:(s + 1) * nd])
            inc.add_phase(adds[s * na:(s + 1) * na])
        spent += time.perf_counter() - t
        done += 1
        s += 1
    return dict(value=done * (na + nd) / spent if spent else 0.0, unit="updates/s", cores=1, kind="port",
                sample=f"{note}{done} snapshots after {warmup} warm-up snapshots, single-threaded oracle port "
                       f"(delete stage + add stage), {spent:.2f}s", inc=inc)


# The CPU port needs minutes just to build its rows and converge on 2^30 edges,
# so the RMAT-26 workloads are sampled on RMAT-22 with the same update/edge ratio.
CPU_SAMPLE = {"c3": dict(scale=22, edges=1 << 26, app="sssp", adds=31250, dels=31250),
              "c4": dict(scale=22, edges=1 << 26, app="sssp", adds=31250, dels=31250)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default=os.environ.get("DCG_BENCH_CONFIG", "c1"), choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--check", action="store_true", help="verify the final state against the oracle")
    ap.add_argument("--no-compare", action="store_true",
                    help="with the learned policy, skip the hotness/LRU run on the same stream")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    cfg = CONFIGS[args.config]
    warmup = args.warmup
    steps = args.steps if args.steps is not None else cfg["snapshots"] - warmup
    snapshots = warmup + steps
    n, src, dst, adds, dels, ap, dp = make_inputs(cfg, snapshots)
    import dycgraph_amd as d
    hbm_gbs, host_gbs = d.calibrate()
    log(f"calibration: HBM copy {hbm_gbs:.0f} GB/s, pinned-host zero-copy read {host_gbs:.1f} GB/s")
    res = run_gpu(cfg, n, src, dst, adds, dels, ap, dp, warmup, steps, cfg["policy"])
    upd = int(ap[warmup:].sum() + dp[warmup:].sum())
    value = upd / res["total_s"]
    lat = np.array(res["lat"])
    ks = res["ks"]
    dom = max(ks.items(), key=lambda kv: kv[1]["ms"])
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
        "host_link_edge_bytes": int(res["host"] * 4),
        "calibration": {"hbm_copy_GBps": round(hbm_gbs, 1), "host_zero_copy_read_GBps": round(host_gbs, 2)},
        "roofline": {"kernel": kname, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "launches": k["launches"], "avg_launch_ms": round(k["ms"] / max(k["launches"], 1), 4),
                     "algo_bytes_per_launch": round(k["algo_bytes"] / max(k["launches"], 1), 1)},
        "kernels": {name: {"ms": round(v["ms"], 3), "launches": v["launches"],
                           "GBps": round(v["algo_bytes"] / (v["ms"] / 1e3) / 1e9, 1) if v["ms"] else 0.0}
                    for name, v in ks.items() if v["launches"]},
    }
    if cfg["policy"] == 2 and not args.no_compare:
        # the plain policy on the identical graph and stream (the judge of the learned one)
        vals_learned = res["values"]
        res["eng"].close()
        lru = run_gpu(cfg, n, src, dst, adds, dels, ap, dp, warmup, steps, 1)
        lreads = lru["hits"] + lru["host"]
        lval = upd / lru["total_s"]
        out["compare_lru"] = {
            "policy": "hotness/LRU", "value": round(lval, 1),
            "p50_ms": round(float(np.percentile(lru["lat"], 50)), 3),
            "p90_ms": round(float(np.percentile(lru["lat"], 90)), 3),
            "cache_hit_rate": round(lru["hits"] / lreads, 4) if lreads else None,
            "host_link_edge_bytes": int(lru["host"] * 4),
            "learned_over_lru": round(value / lval, 4) if lval else None,
            "identical_values": bool(np.array_equal(vals_learned, lru["values"])),
        }
        res = lru
    cpu = None
    if not args.no_cpu and cfg["app"] != "pr" and "grid" not in cfg:
        if args.config in CPU_SAMPLE:
            sc = CPU_SAMPLE[args.config]
            sn, ss, sd, sa, sdl, _, _ = make_inputs(sc, warmup + 4)
            cpu = cpu_baseline(sc, sn, ss, sd, sa, sdl, warmup,
                               note=f"RMAT-{sc['scale']} sample ({sc['edges']} edges, {sc['adds'] + sc['dels']} "
                                    f"updates/snapshot, same generator and update/edge ratio): ")
        else:
            cpu = cpu_baseline(cfg, n, src, dst, adds, dels, warmup)
        out["cpu_baseline"] = {k: (round(v, 1) if isinstance(v, float) else v) for k, v in cpu.items() if k != "inc"}
    if args.check and cfg["app"] != "pr" and "queries" not in cfg:
        import oracle
        store = oracle.Store(n, src, dst)
        aoff = np.concatenate([[0], np.cumsum(ap)]).astype(np.int64)
        doff = np.concatenate([[0], np.cumsum(dp)]).astype(np.int64)
        for s in range(snapshots):
            store.apply(dels[doff[s]:doff[s + 1]], adds[aoff[s]:aoff[s + 1]])
        rp, col = store.csr()
        want = oracle.sssp(n, rp, col, 0, oracle.W_SEEDED, 11, 100)
        out["exact"] = bool(np.array_equal(res["values"], want))
    res["eng"].close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
