"""The drop-in command lines (hybrid_sssp/bfs/cc/pr) on the reference's own smoke
scenarios (tests/bfs_dynamic_smoke.py, cc_dynamic_smoke.py, pr_dynamic_smoke.py):
every phase checksum the binary logs must equal the oracle's, as those tests require."""
import os
import re
import subprocess
from collections import Counter

import numpy as np
import pytest

import oracle
from dycgraph_amd import traces
from tests.conftest import announce

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "dycgraph_amd", "bin")


def _run(app, tmp, extra, env=None):
    cmd = [os.path.join(BIN, f"hybrid_{app}"), f"--graphfile={tmp}/graph", f"--updatefile={tmp}/updates",
           f"--update_size={tmp}/sizes", "--format=market_big", "--source_node=0", "--SEGMENT=32", "--n_stream=3",
           "--hybrid=0", "--check=true", "--verbose=false"] + extra
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120,
                       env=None if env is None else dict(os.environ, **env))
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return r.stdout


def _csr_of(n, counter):
    src, dst = [], []
    for (u, v), c in sorted(counter.items()):
        src += [u] * c
        dst += [v] * c
    return oracle.csr_from_edges(n, np.array(src, np.uint32), np.array(dst, np.uint32))


HINTS = {"default": None, "ordered-large": {"CG_ORDERED_REPAIR": "1", "CG_BATCH_MAINTENANCE": "large"},
         "thread-auto": {"CG_INSERTION_SCHEDULE": "thread", "CG_BATCH_MAINTENANCE": "auto"}}


@pytest.mark.parametrize("app", ["bfs", "sssp"])
@pytest.mark.parametrize("placement", ["device", "host"])
@pytest.mark.parametrize("hints", sorted(HINTS))
def test_bfs_sssp_smoke_scenario(tmp_path, app, placement, hints):
    """Also under the reference's algorithm hints (CG_ORDERED_REPAIR,
    CG_INSERTION_SCHEDULE, CG_BATCH_MAINTENANCE): same checksums."""
    announce(f"cli {app} smoke scenario placement={placement} hints={hints}")
    n = 2048
    edges = {(u, (u + 1) % n) for u in range(n)}
    edges |= {(u, 1024) for u in range(1024)} | {(0, 0), (0, 2), (1, 3)}
    (tmp_path / "graph").write_text("".join(f"{u} {v}\n" for u, v in sorted(edges)))
    deleted = [(0, 1024), (20, 21), (700, 701), (1200, 1201)]
    added = [(1, 1900), (500, 1700), (900, 2000), (3, 1500)]
    cur = Counter(edges)
    expected, updates = [], []
    tag = app.upper()
    for dels, adds in [(deleted, added), (added, deleted), (deleted, added)]:
        for u, v in dels:
            cur[(u, v)] -= 1
            updates.append(f"d {u} {v} 77\n")
        rp, col = _csr_of(n, +cur)
        d = oracle.bfs(n, rp, col, 0) if app == "bfs" else oracle.sssp(n, rp, col, 0)
        expected.append((f"{tag}-DELETE-STAGE-CHECK", oracle.checksum(d)))
        for u, v in adds:
            cur[(u, v)] += 1
            updates.append(f"a {u} {v} 91\n")
        rp, col = _csr_of(n, +cur)
        d = oracle.bfs(n, rp, col, 0) if app == "bfs" else oracle.sssp(n, rp, col, 0)
        expected.append((f"{tag}-BATCH-CHECK", oracle.checksum(d)))
    (tmp_path / "updates").write_text("".join(updates))
    (tmp_path / "sizes").write_text("4 4\n" * 3)
    out = _run(app, tmp_path, [f"--{app}_max_batches=3", f"--{app}_print_checksum=true", f"--placement={placement}",
                               "--cache=0.0003", f"--output={tmp_path}/out"], env=HINTS[hints])
    got = [(t, int(v)) for t, v in re.findall(rf"\[({tag}-DELETE-STAGE-CHECK|{tag}-BATCH-CHECK)\][^\n]*"
                                              rf"distance_checksum=(\d+)", out)]
    assert got == expected
    assert len(re.findall(rf"\[{tag}-(?:DELETE-STAGE|BATCH)-CHECK\][^\n]* passed ", out)) == 6
    rows = [list(map(int, ln.split())) for ln in (tmp_path / "out").read_text().splitlines()]
    rp, col = _csr_of(n, +cur)
    want = oracle.bfs(n, rp, col, 0) if app == "bfs" else oracle.sssp(n, rp, col, 0)
    assert [r[1] for r in rows] == want.tolist()
    for node, dist, parent, _ in rows:
        if node and dist != 2 ** 32 - 1:
            w = 1 if app == "bfs" else (parent + node) % 128 + 1
            assert cur[(parent, node)] > 0 and want[parent] + w == dist


def test_cc_smoke_scenario(tmp_path):
    announce("cli cc smoke scenario")
    nodes = 256
    base = [(0, 0), (11, 11), (255, 255)]
    base += [(0, 1), (1, 2), (2, 0), (2, 3), (3, 4), (4, 5), (5, 3)]
    base += [(8, 9), (8, 9), (130, 140), (140, 150), (150, 130)]
    base += [(u, u + 1) for u in range(160, 200)]
    base += [(31, 30), (32, 30), (40, 41), (41, 42), (42, 41)]
    base += [(0, 220), (220, 210), (210, 211), (211, 210)]
    base += [(50, 51), (50, 52), (51, 53), (52, 53), (53, 54), (54, 53)]
    base += [(60, 61), (60, 62), (61, 63), (62, 63), (63, 64), (64, 63)]
    original = Counter(base)
    (tmp_path / "graph").write_text("".join(f"{u} {v} 777\n" * c for (u, v), c in sorted(original.items())))
    batches = [([], []), ([(2, 3), (40, 41), (220, 210)], []), ([(0, 1)], []),
               ([(8, 9)], []), ([(8, 9)], [(5, 8)]),
               ([(3, 4), (3, 5)], [(2, 4), (9, 130)]),
               ([(180, 181)], [(150, 190)]),
               ([(210, 211), (11, 11)], [(10, 11)]),
               ([], [(5, 8)]), ([(5, 8)], [(2, 3)])]
    batches += [([(51, 53)], []), ([(50, 52), (61, 63), (62, 63)], []), ([], [(49, 53)]), ([(49, 53)], [])]
    rng = np.random.default_rng(20260923)   # literal-free seeded batches (the reference used Python's random)
    for _ in range(6):
        batches.append(([tuple(map(int, rng.integers(0, nodes, 2))) for _ in range(8)],
                        [tuple(map(int, rng.integers(0, nodes, 2))) for _ in range(8)]))
    cur = original.copy()
    expected, updates, sizes = [], [], []
    for dels, adds in batches:
        for u, v in dels:
            cur[(u, v)] = max(0, cur[(u, v)] - 1)
            updates.append(f"d {u} {v} 999\n")
        rp, col = _csr_of(nodes, +cur)
        expected.append(("CC-DELETE-STAGE-CHECK", oracle.checksum(oracle.cc(nodes, rp, col))))
        for u, v in adds:
            cur[(u, v)] += 1
            updates.append(f"a {u} {v} 123\n")
        rp, col = _csr_of(nodes, +cur)
        expected.append(("CC-BATCH-CHECK", oracle.checksum(oracle.cc(nodes, rp, col))))
        sizes.append(f"{len(adds)} {len(dels)}\n")
    (tmp_path / "updates").write_text("".join(updates))
    (tmp_path / "sizes").write_text("".join(sizes))
    for placement in ("device", "host"):
        out = _run("cc", tmp_path, [f"--cc_max_batches={len(batches)}", "--cc_print_checksum=true",
                                    f"--placement={placement}", "--cache=0.0003"])
        got = [(t, int(v)) for t, v in re.findall(r"\[(CC-DELETE-STAGE-CHECK|CC-BATCH-CHECK)\][^\n]*"
                                                  r"label_checksum=(\d+)", out)]
        assert got == expected, placement
        assert "[CC-FINAL-CHECK] passed errors=0" in out


def test_pr_smoke_scenario(tmp_path):
    announce("cli pr smoke scenario")
    n = 4096
    base = [(0, 1), (1, 2), (2, 0), (3, 4), (4, 3), (8, 9), (8, 9), (10, 10), (n - 1, n - 1)]
    base += [(20, v) for v in range(21, 1800)]
    base += [(v, 20) for v in range(21, 1800)]
    base += [(v, v + 1) for v in range(2000, 2020)]
    (tmp_path / "graph").write_text("".join(f"{u} {v} 777\n" for u, v in sorted(base)))
    batches = [([], []), ([(2, 0)], []), ([], [(2, 3)]), ([(3, 4)], [(3, 4)]), ([(8, 9)], []),
               ([(8, 9), (10, 10)], [(9, 8)]),
               ([(20, v) for v in range(21, 1800, 3)], [(20, 2000), (2020, 20)]),
               ([(2000, 2001), (4000, 4001)], [(0, 8), (3, 10)]),
               ([], [(10, 10), (10, 3), (8, 9), (8, 9)]),
               ([(20, 2000), (2020, 20)], [(2000, 2001)])]
    updates, sizes = [], []
    for dels, adds in batches:
        updates += [f"d {u} {v} 999\n" for u, v in dels] + [f"a {u} {v} 123\n" for u, v in adds]
        sizes.append(f"{len(adds)} {len(dels)}\n")
    (tmp_path / "updates").write_text("".join(updates))
    (tmp_path / "sizes").write_text("".join(sizes))
    out = _run("pr", tmp_path, ["--error=1e-6", f"--pr_max_batches={len(batches)}", "--placement=device"])
    checks = re.findall(r"\[PR-CHECK\] stage=(\w+) batch=(-?\d+) (passed|failed)", out)
    assert len(checks) == len(batches) + 2 and all(c[2] == "passed" for c in checks), out[-2000:]


def test_cli_trace_flags(tmp_path):
    """hybrid_sssp with --i16_repair_snapshot and --f1_cache_trace_file writes
    both files and logs their lines; the snapshot replays without mismatches."""
    announce("cli trace flags")
    n = 1024
    rng = np.random.default_rng(5)
    edges = sorted({(int(u), int(v)) for u, v in rng.integers(0, n, (6000, 2)) if u != v})
    (tmp_path / "graph").write_text("".join(f"{u} {v}\n" for u, v in edges))
    dl = [edges[i] for i in rng.choice(len(edges), 200, replace=False)]
    ad = [(int(u), int(v)) for u, v in rng.integers(0, n, (200, 2))]
    (tmp_path / "updates").write_text("".join(f"d {u} {v} 1\n" for u, v in dl) +
                                      "".join(f"a {u} {v} 1\n" for u, v in ad))
    (tmp_path / "sizes").write_text("200 200\n")
    snap, f1 = tmp_path / "i16.bin", tmp_path / "f1.bin"
    cmd = [os.path.join(BIN, "hybrid_sssp"), f"--graphfile={tmp_path}/graph",
           f"--updatefile={tmp_path}/updates", f"--update_size={tmp_path}/sizes", "--format=market_big",
           "--source_node=0", "--check=true", f"--i16_repair_snapshot={snap}", f"--f1_cache_trace_file={f1}",
           "--cache=1"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "[I16-SNAPSHOT][batch 0] states=" in r.stdout and r.stdout.count("[F1-CACHE-TRACE]") == 2
    sn = traces.read_repair_snapshot(snap)
    res = traces.replay_repair_snapshot(sn)
    assert res["mismatches"] == 0 and res["external_changes"] == 0
    recs = traces.read_cache_trace(f1)
    assert [b for b, _, _ in recs] == [0xFFFFFFFF, 0] and all(c == 268435455 for _, c, _ in recs)
    assert len(recs[0][2]) == n   # everything fits a 1-unit cache on this graph


@pytest.mark.parametrize("mode", ["capacity", "map"])
def test_cli_cpu_gpu_coexecution(tmp_path, mode):
    """hybrid_sssp with --sssp_cpu_partition_capacity (the CPU takes the
    partitions with the most insertion seeds of each batch) or
    --sssp_cpu_domain_map (u16 per vertex, 0 = CPU): checks pass, the
    distances equal the oracle's and the host did work."""
    announce(f"cli co-execution {mode}")
    n = 2048
    rng = np.random.default_rng(9)
    edges = sorted({(int(u), int(v)) for u, v in rng.integers(0, n, (12000, 2)) if u != v})
    (tmp_path / "graph").write_text("".join(f"{u} {v}\n" for u, v in edges))
    dl = [edges[i] for i in rng.choice(len(edges), 300, replace=False)]
    ad = [(int(u), int(v)) for u, v in rng.integers(0, n, (300, 2)) if u != v][:300]
    (tmp_path / "updates").write_text("".join(f"d {u} {v} 1\n" for u, v in dl) +
                                      "".join(f"a {u} {v} 1\n" for u, v in ad))
    (tmp_path / "sizes").write_text(f"{len(ad)} {len(dl)}\n")
    if mode == "map":
        dom = np.ones(n, np.uint16)
        dom[n // 4: n // 2] = 0
        dom.tofile(tmp_path / "domain")
        extra = [f"--sssp_cpu_domain_map={tmp_path}/domain"]
    else:
        extra = ["--sssp_cpu_partition_capacity=3", "--SEGMENT=8"]
    out = _run("sssp", tmp_path, ["--sssp_max_batches=1", f"--output={tmp_path}/out"] + extra)
    assert "[E2-COEXEC][batch 0]" in out
    relax = int(re.search(r"cpu_relaxations=(\d+)", out).group(1))
    assert relax > 0
    cur = Counter(edges)
    for u, v in dl:
        cur[(u, v)] -= 1
    for u, v in ad:
        cur[(u, v)] += 1
    rp, col = _csr_of(n, +cur)
    want = oracle.sssp(n, rp, col, 0)
    rows = [list(map(int, ln.split())) for ln in (tmp_path / "out").read_text().splitlines()]
    assert [r[1] for r in rows] == want.tolist()

