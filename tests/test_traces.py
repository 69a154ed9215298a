"""The trace readers and the offline repair replay, on the CPU.

The cache-candidate layout is pinned by the reference's own test
(tests/cache_patch_trace_test.cpp:8-39: two records, the second "unchanged");
the repair-snapshot replay by a hand-built snapshot whose repaired distances
follow by hand, and by planted damage that the readers must reject.
"""
import struct

import numpy as np
import pytest

from dycgraph_amd import traces


def test_cache_trace_reference_vector(tmp_path):
    p = tmp_path / "f1.bin"
    with open(p, "wb") as f:
        f.write(struct.pack("<Q", traces.CACHE_TRACE_MAGIC))
        f.write(struct.pack("<IQQ", 7, 100, 2) + struct.pack("<IIII", 3, 40, 9, 50))
        f.write(struct.pack("<IQQ", 8, 100, 0))
    recs = traces.read_cache_trace(p)
    assert [(b, c, len(a)) for b, c, a in recs] == [(7, 100, 2), (8, 100, 2)]
    assert recs[0][2]["vertex"].tolist() == [3, 9] and recs[0][2]["degree"].tolist() == [40, 50]
    assert recs[1][2]["vertex"].tolist() == [3, 9]   # count 0 repeats the previous list
    with open(p, "r+b") as f:
        f.write(struct.pack("<Q", 1))
    with pytest.raises(ValueError):
        traces.read_cache_trace(p)


def _snapshot(path, affected, rows, before, after, source=0, nodes=8, rule=1):
    offs = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype("<u8")
    inc = np.array([x for r in rows for x in r], "<u4")

    def vec(a, dt):
        a = np.asarray(a, dt)
        return struct.pack("<Q", len(a)) + a.tobytes()

    st = lambda d: np.array([(v, x, x, p) for v, (x, p) in sorted(d.items())], traces.STATE)
    with open(path, "wb") as f:
        f.write(struct.pack("<QIII", traces.REPAIR_MAGIC, 1, 0xFFFFFFFF, rule))
        f.write(struct.pack("<III", 0, source, nodes))
        f.write(vec(affected, "<u4") + vec(offs, "<u8") + vec(inc, "<u4"))
        f.write(vec(st(before), traces.STATE) + vec(st(after), traces.STATE))


def test_repair_replay_hand_example(tmp_path):
    """Vertices 3, 4, 5 lost their paths: 3 <- {1, 4}, 4 <- {3}, 5 <- {4, 2};
    with unit weights the outside offers are 1 + value: 3 gets 6 from 1 (value 5),
    4 = 3 + 1, 5 = min(4 + 1, 2's value 9 + 1) = 8."""
    inf = traces.INF
    before = {0: (0, inf), 1: (5, 0), 2: (9, 0), 3: (inf, inf), 4: (inf, inf), 5: (inf, inf)}
    after = dict(before)
    after.update({3: (6, 1), 4: (7, 3), 5: (8, 4)})
    p = tmp_path / "i16.bin"
    _snapshot(p, [3, 4, 5], [[1, 4], [3], [4, 2]], before, after)
    s = traces.read_repair_snapshot(p)
    assert s["affected"].tolist() == [3, 4, 5] and s["offsets"].tolist() == [0, 2, 3, 5]
    r = traces.replay_repair_snapshot(s, weight=lambda u, v: 1)
    assert r["distance"].tolist() == [6, 7, 8] and r["mismatches"] == 0 and r["external_changes"] == 0
    # a wrong repaired value and a changed outside vertex are both reported
    after[5] = (9, 2)
    after[1] = (4, 0)
    _snapshot(p, [3, 4, 5], [[1, 4], [3], [4, 2]], before, after)
    r = traces.replay_repair_snapshot(traces.read_repair_snapshot(p), weight=lambda u, v: 1)
    assert r["mismatches"] == 1 and r["external_changes"] == 1
    # the reference's weight rule: (u + v) % 128 + 1
    assert traces.reference_weight(100, 30) == 3


def test_repair_snapshot_rejects_damage(tmp_path):
    p = tmp_path / "i16.bin"
    _snapshot(p, [1], [[0]], {0: (0, 0), 1: (traces.INF, traces.INF)}, {0: (0, 0), 1: (1, 0)})
    data = p.read_bytes()
    p.write_bytes(data + b"\0")
    with pytest.raises(ValueError):
        traces.read_repair_snapshot(p)
    p.write_bytes(data[:-3])
    with pytest.raises(ValueError):
        traces.read_repair_snapshot(p)
    p.write_bytes(data[:12] + struct.pack("<I", 7) + data[16:])   # foreign weight rule
    with pytest.raises(ValueError):
        traces.read_repair_snapshot(p)
