"""Readers and the offline replay for the engine's diagnostic traces.

Both files use the reference's layouts, so traces written by either side can be
read here:

* repair snapshot: include/framework/i16_repair_snapshot.h:9-52 (written by
  ``Engine.set_repair_snapshot``); ``replay_repair_snapshot`` recomputes the
  affected vertices' repaired distances from the snapshot alone, as the
  reference's offline tool does (src/i16_repair_oracle.cpp:13-99);
* cache-candidate trace: include/framework/cache_patch_trace.h:15-60 (written
  by ``Engine.set_cache_trace`` + ``Engine.trace_cache_candidates``).
"""
from __future__ import annotations

import heapq

import numpy as np

REPAIR_MAGIC = 0x4931365354415445
CACHE_TRACE_MAGIC = 0x4631434143484531
STATE = np.dtype([("vertex", "<u4"), ("value", "<u4"), ("buffer", "<u4"), ("parent", "<u4")])
INF = 0xFFFFFFFF


class _Reader:
    def __init__(self, data: bytes):
        self.b, self.p = data, 0

    def scalar(self, dtype):
        dt = np.dtype(dtype)
        if self.p + dt.itemsize > len(self.b):
            raise ValueError("truncated trace")
        v = np.frombuffer(self.b, dt, 1, self.p)[0]
        self.p += dt.itemsize
        return int(v)

    def vector(self, dtype):
        count = self.scalar("<u8")
        dt = np.dtype(dtype)
        if count > (len(self.b) - self.p) // dt.itemsize:
            raise ValueError("invalid vector count in trace")
        v = np.frombuffer(self.b, dt, count, self.p).copy()
        self.p += count * dt.itemsize
        return v


def read_repair_snapshot(path):
    """The snapshot as a dict; raises ValueError on a foreign or damaged file
    (the checks of i16_repair_snapshot.h:40-52; weight rule 2 is this engine's
    seeded weights, which the reference's loader does not accept)."""
    with open(path, "rb") as f:
        r = _Reader(f.read())
    magic, version, infinity, rule = r.scalar("<u8"), r.scalar("<u4"), r.scalar("<u4"), r.scalar("<u4")
    if magic != REPAIR_MAGIC or version != 1 or infinity != INF or rule not in (1, 2):
        raise ValueError("unsupported repair snapshot contract")
    s = {"weight_rule": rule, "batch": r.scalar("<u4"), "source": r.scalar("<u4"), "nodes": r.scalar("<u4")}
    s["affected"] = r.vector("<u4")
    s["offsets"] = r.vector("<u8")
    s["incoming"] = r.vector("<u4")
    s["before"] = r.vector(STATE)
    s["after"] = r.vector(STATE)
    if r.p != len(r.b):
        raise ValueError("trailing bytes in repair snapshot")
    return s


def reference_weight(u, v):
    """Weight rule 1: (u + v) % 128 + 1."""
    return (int(u) + int(v)) % 128 + 1


def replay_repair_snapshot(s, weight=reference_weight):
    """Recompute the affected vertices' distances from the snapshot: every
    affected vertex starts at the best offer from its unaffected in-neighbours
    (their values in `before`), then Dijkstra runs over the edges among the
    affected vertices.  Returns a dict with the distances (in the order of
    s["affected"]), the mismatches against s["after"] and the count of states
    outside the affected set whose value changed."""
    aff = s["affected"]
    n = len(aff)
    off, inc = s["offsets"], s["incoming"]
    if len(off) != n + 1 or (n and (off[0] != 0 or off[-1] != len(inc))) or np.any(np.diff(off.astype(np.int64)) < 0):
        raise ValueError("invalid in-row offsets")
    local = {int(v): i for i, v in enumerate(aff)}
    before = {int(x["vertex"]): int(x["value"]) for x in s["before"]}
    after = {int(x["vertex"]): int(x["value"]) for x in s["after"]}
    dist = [before.get(int(v), INF) for v in aff]
    out = [[] for _ in range(n)]   # internal edges src -> dst, by local index
    for i in range(n):
        dst = int(aff[i])
        for e in range(int(off[i]), int(off[i + 1])):
            src = int(inc[e])
            j = local.get(src)
            if j is not None:
                out[j].append(i)
            elif before.get(src, INF) != INF:
                dist[i] = min(dist[i], before[src] + weight(src, dst))
    heap = [(d, i) for i, d in enumerate(dist) if d != INF]
    heapq.heapify(heap)
    while heap:
        d, u = heapq.heappop(heap)
        if d != dist[u]:
            continue
        for v in out[u]:
            c = d + weight(aff[u], aff[v])
            if c < dist[v]:
                dist[v] = c
                heapq.heappush(heap, (c, v))
    dist = [d if d < INF else INF for d in dist]
    mismatches = sum(1 for i, v in enumerate(aff) if after.get(int(v)) != dist[i])
    outside = sum(1 for v, x in before.items() if v not in local and after.get(v) != x)
    return {"distance": np.array(dist, np.uint64), "mismatches": mismatches, "external_changes": outside}


def read_cache_trace(path):
    """[(batch, capacity_edges, admitted (vertex, degree) array)]; a record of
    count 0 repeats the previous list (cache_patch_trace.h:33-58)."""
    with open(path, "rb") as f:
        r = _Reader(f.read())
    if r.scalar("<u8") != CACHE_TRACE_MAGIC:
        raise ValueError("invalid cache candidate trace magic")
    rec = np.dtype([("vertex", "<u4"), ("degree", "<u4")])
    out, prev = [], np.zeros(0, rec)
    while r.p < len(r.b):
        batch, cap, count = r.scalar("<u4"), r.scalar("<u8"), r.scalar("<u8")
        if count > (len(r.b) - r.p) // rec.itemsize:
            raise ValueError("truncated cache candidate trace records")
        adm = np.frombuffer(r.b, rec, count, r.p).copy()
        r.p += count * rec.itemsize
        if count == 0 and out:
            adm = prev
        prev = adm
        out.append((batch, cap, adm))
    return out
