"""Write golden vectors from the reference's own Python test oracles.

Run in a container that holds the reference checkout (default /root/reference):

    python tests/golden/make_reference_vectors.py [REFERENCE_ROOT]

It imports `oracle()` from the reference's tests/bfs_dynamic_smoke.py (FIFO BFS or
Dijkstra with (u+v)%128+1 weights from vertex 0, then the value checksum) and
tests/cc_dynamic_smoke.py (increasing-seed directed flood fill, then the checksum),
evaluates them on the graphs those tests build (their KAT graph after every delete
and add stage, and the seeded random CC batches drawn with random.Random(20260923)),
and writes the inputs and outputs to reference_vectors.json next to this script.
Only the vectors are kept; the reference's code is not copied and is not needed
to run the tests that read them.
"""
import base64
import importlib.util
import json
import os
import random
import struct
import sys
import zlib
from collections import Counter

HERE = os.path.dirname(os.path.abspath(__file__))


def load(ref, name):
    path = os.path.join(ref, "tests", name + ".py")
    spec = importlib.util.spec_from_file_location("ref_" + name, path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def pack(values):
    """uint32 little-endian, zlib, base64 (tests/test_oracle_golden.py unpacks it)."""
    raw = struct.pack("<%dI" % len(values), *values)
    return base64.b64encode(zlib.compress(raw, 9)).decode()


def bfs_cases(mod):
    """The BFS/SSSP KAT graph (tests/bfs_dynamic_smoke.py:61-75): n = 2048, a
    directed cycle, every u < 1024 into 1024, a self-loop and short witnesses;
    three snapshots of four deletions and four additions."""
    n = 2048
    edges = {(u, (u + 1) % n) for u in range(n)}
    edges |= {(u, 1024) for u in range(1024)} | {(0, 0), (0, 2), (1, 3)}
    deleted = [(0, 1024), (20, 21), (700, 701), (1200, 1201)]
    added = [(1, 1900), (500, 1700), (900, 2000), (3, 1500)]
    out = {"graph": "bfs_kat", "n": n, "initial_edges": sorted(edges), "stages": []}

    def snap(stage, dels, adds):
        rec = {"stage": stage, "dels": dels, "adds": adds}
        for weighted in (False, True):
            dist, chk = mod.oracle(sorted(edges), n, weighted)
            rec["sssp" if weighted else "bfs"] = {"values": pack(dist), "checksum": chk}
        out["stages"].append(rec)

    snap("initial", [], [])
    for s, (dels, adds) in enumerate([(deleted, added), (added, deleted), (deleted, added)]):
        for e in dels:
            edges.remove(e)
        snap(f"delete{s}", dels, [])
        for e in adds:
            edges.add(e)
        snap(f"add{s}", [], adds)
    return out


def cc_cases(mod):
    """The CC KAT graph and batches (tests/cc_dynamic_smoke.py:52-77), including the
    six seeded random batches of random.Random(20260923), applied as multiset
    deletions (one occurrence, none if absent) and additions."""
    nodes = 256
    base = [(0, 0), (11, 11), (255, 255)]
    base += [(0, 1), (1, 2), (2, 0), (2, 3), (3, 4), (4, 5), (5, 3)]
    base += [(8, 9), (8, 9), (130, 140), (140, 150), (150, 130)]
    base += [(u, u + 1) for u in range(160, 200)]
    base += [(31, 30), (32, 30), (40, 41), (41, 42), (42, 41)]
    base += [(0, 220), (220, 210), (210, 211), (211, 210)]
    base += [(50, 51), (50, 52), (51, 53), (52, 53), (53, 54), (54, 53)]
    base += [(60, 61), (60, 62), (61, 63), (62, 63), (63, 64), (64, 63)]
    batches = [([], []), ([(2, 3), (40, 41), (220, 210)], []), ([(0, 1)], []),
               ([(8, 9)], []), ([(8, 9)], [(5, 8)]),
               ([(3, 4), (3, 5)], [(2, 4), (9, 130)]),
               ([(180, 181)], [(150, 190)]),
               ([(210, 211), (11, 11)], [(10, 11)]),
               ([], [(5, 8)]), ([(5, 8)], [(2, 3)])]
    batches += [([(51, 53)], []), ([(50, 52), (61, 63), (62, 63)], []),
                ([], [(49, 53)]), ([(49, 53)], [])]
    rng = random.Random(20260923)
    for _ in range(6):
        batches.append(([(rng.randrange(nodes), rng.randrange(nodes)) for _ in range(8)],
                        [(rng.randrange(nodes), rng.randrange(nodes)) for _ in range(8)]))
    edges = Counter(base)
    labels, chk = mod.oracle(edges, nodes)
    out = {"graph": "cc_kat", "n": nodes, "base": base, "batches": [],
           "initial": {"labels": pack(labels), "checksum": chk}}
    for dels, adds in batches:
        for u, v in dels:
            edges[u, v] = max(0, edges[u, v] - 1)
        ld, cd = mod.oracle(edges, nodes)
        for u, v in adds:
            edges[u, v] += 1
        la, ca = mod.oracle(edges, nodes)
        out["batches"].append({"dels": dels, "adds": adds, "after_delete": {"labels": pack(ld), "checksum": cd},
                               "after_add": {"labels": pack(la), "checksum": ca}})
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    bfs = load(ref, "bfs_dynamic_smoke")
    cc = load(ref, "cc_dynamic_smoke")
    doc = {"source": "reference tests/bfs_dynamic_smoke.py and tests/cc_dynamic_smoke.py oracle()",
           "bfs_sssp": bfs_cases(bfs), "cc": cc_cases(cc)}
    path = os.path.join(HERE, "reference_vectors.json")
    with open(path, "w") as f:
        json.dump(doc, f, separators=(",", ":"))
    print(f"wrote {path}: {len(doc['bfs_sssp']['stages'])} BFS/SSSP stages, {len(doc['cc']['batches'])} CC batches")


if __name__ == "__main__":
    main()
