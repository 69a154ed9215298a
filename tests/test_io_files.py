"""Graph and update files (no device): the threaded market_big reader, the
binary edge list, the update/size files and the dcg_gen tool.

The expected contents come from numpy's own text parsing of the same files
and from the in-memory generators with the same seeds, so a reader that drops,
reorders or misparses a line fails."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GEN = os.path.join(ROOT, "dycgraph_amd", "bin", "dcg_gen")


def test_market_big_text_and_binary_round_trip(tmp_path):
    import dycgraph_amd as d
    rng = np.random.default_rng(3)
    m = 300_000   # > 1 MiB of text: the reader splits the file over threads
    src = rng.integers(0, 5_000_000, m).astype(np.uint32)
    dst = rng.integers(0, 5_000_000, m).astype(np.uint32)
    txt, binf = tmp_path / "g.txt", tmp_path / "g.bin"
    n = int(max(src.max(), dst.max())) + 1
    d.write_graph(txt, n, src, dst)
    d.write_graph(binf, n, src, dst, binary=True)
    ref = np.loadtxt(txt, dtype=np.uint64).astype(np.uint32)
    assert np.array_equal(ref[:, 0], src) and np.array_equal(ref[:, 1], dst)
    for threads in (1, 3, 8):
        n2, s2, d2 = d.read_graph(txt, threads)
        assert n2 == n and np.array_equal(s2, src) and np.array_equal(d2, dst)
    n3, s3, d3 = d.read_graph(binf, 4)
    assert n3 == n and np.array_equal(s3, src) and np.array_equal(d3, dst)
    assert os.path.getsize(binf) == 24 + 8 * m


def test_market_big_line_forms(tmp_path):
    """Weights in a third column, CRLF, tabs, blank and malformed lines, no final
    newline: n = largest id + 1, edges in file order, lines without two
    integers skipped."""
    import dycgraph_amd as d
    p = tmp_path / "g.txt"
    p.write_bytes(b"0 1 7\r\n\n3\t2\n# header\n  5 4 9 9\n1 1\n9 0")
    n, s, t = d.read_graph(p)
    assert n == 10 and s.tolist() == [0, 3, 5, 1, 9] and t.tolist() == [1, 2, 4, 1, 0]
    p.write_bytes(b"")
    assert d.read_graph(p)[0] == 0
    with pytest.raises(d.DcgError):
        d.read_graph(tmp_path / "missing.txt")
    p.write_bytes(b"DCGEDGE1" + np.array([1, 3], "<u4").tobytes() + np.array([2], "<u8").tobytes() +
                  np.array([0, 1, 2, 5], "<u4").tobytes())   # id 5 >= num_nodes 3
    with pytest.raises(d.DcgError):
        d.read_graph(p)


def test_update_files_round_trip(tmp_path):
    import dycgraph_amd as d
    rng = np.random.default_rng(4)
    sizes = np.array([[3, 2], [0, 4], [5, 0]], np.uint64)
    adds = rng.integers(0, 1000, (8, 2)).astype(np.uint32)
    dels = rng.integers(0, 1000, (6, 2)).astype(np.uint32)
    sp, up = tmp_path / "sizes", tmp_path / "updates"
    d.write_updates(sp, up, sizes, adds, dels)
    lines = up.read_text().splitlines()
    assert lines[0].startswith("d ") and len(lines) == 14
    u, v, w = map(int, lines[0].split()[1:])
    assert (u, v) == tuple(dels[0]) and w == (u + v) % 128 + 1   # the reference's weight rule
    sz, a, dl = d.read_updates(sp, up, threads=2)
    assert np.array_equal(sz, sizes) and np.array_equal(a, adds) and np.array_equal(dl, dels)
    # Loader.h: additions and deletions each in file order, whatever the interleaving
    up.write_text(" a 1 2 5\nd 3 4 1\nx 9 9 9\na 5 6\n\nd 7 8 2\n")
    sp.write_text("2 2\n")
    sz, a, dl = d.read_updates(sp, up)
    assert sz.tolist() == [[2, 2]] and a.tolist() == [[1, 2], [5, 6]] and dl.tolist() == [[3, 4], [7, 8]]


@pytest.mark.parametrize("binary", [False, True])
def test_generator_tool_writes_the_in_memory_inputs(tmp_path, binary):
    import dycgraph_amd as d
    g, up, sp = tmp_path / "graph", tmp_path / "updates", tmp_path / "sizes"
    cmd = [GEN, "--graph=rmat", "--scale=10", "--edges=8000", "--seed=5", f"--out={g}", f"--updates={up}",
           f"--sizes={sp}", "--snapshots=3", "--adds=100", "--dels=60", "--stream_seed=9"]
    r = subprocess.run(cmd + (["--binary"] if binary else []), capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    n, src, dst = d.read_graph(g)
    s0, d0 = d.gen_rmat(10, 8000, seed=5)
    assert np.array_equal(src, s0) and np.array_equal(dst, d0) and n <= 1 << 10
    adds, dels = d.gen_stream(1 << 10, s0, d0, [100] * 3, [60] * 3, seed=9)
    sz, a, dl = d.read_updates(sp, up)
    assert sz.tolist() == [[100, 60]] * 3
    assert np.array_equal(a, np.asarray(adds).reshape(-1, 2)) and np.array_equal(dl, np.asarray(dels).reshape(-1, 2))
    # grid graphs and lognormal batch sizes
    r = subprocess.run([GEN, "--graph=grid", "--side=32", f"--out={g}", f"--updates={up}", f"--sizes={sp}",
                        "--snapshots=4", "--lognormal=5,1,10,300"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    n, src, dst = d.read_graph(g)
    gs, gd = d.gen_grid(32, 0.05, seed=1)
    assert np.array_equal(src, gs) and np.array_equal(dst, gd)
    tot = d.gen_lognormal_sizes(4, 5.0, 1.0, 10, 300, seed=4).astype(np.int64)
    sz, _, _ = d.read_updates(sp, up)
    assert sz.tolist() == [[t - t // 2, t // 2] for t in tot.tolist()]
