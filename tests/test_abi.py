"""The C ABI library loads and exports every symbol include/*.h declares; host
logic that needs no GPU (config defaults, generators) behaves as specified."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            text = open(os.path.join(inc, f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(dcg_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    import dycgraph_amd as d
    lib = d.lib()
    declared = _declared()
    assert len(declared) >= 25
    out = subprocess.run(["nm", "-D", "--defined-only", d.LIB_PATH], capture_output=True, text=True, check=True)
    exported = {ln.split()[-1] for ln in out.stdout.splitlines() if ln.strip()}
    missing = sorted(declared - exported)
    assert not missing, f"declared but not exported: {missing}"
    for name in declared:
        assert hasattr(lib, name)


def test_kernel_ids_match_the_header():
    """The Python mirror's kernel-id table (kernel_stats keys) lists exactly the
    header's dcg_kernel_id values below DCG_K_NUM."""
    from dycgraph_amd.engine import KERNELS
    text = open(os.path.join(ROOT, "include", "dcg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    body = re.search(r"typedef enum \{(.*?)\} dcg_kernel_id;", text, flags=re.S).group(1)
    ids = {name: int(v) for name, v in re.findall(r"DCG_K_([A-Z_]+)\s*=\s*(\d+)", body)}
    num = ids.pop("NUM")
    assert sorted(ids.values()) == list(range(num))
    assert {k.upper(): v for k, v in KERNELS.items()} == ids


def test_config_defaults_and_status_strings():
    import dycgraph_amd as d
    from dycgraph_amd.engine import Config
    c = Config()
    assert d.lib().dcg_config_init(C.byref(c)) == 0
    assert c.pr_max_rounds == 100 and abs(c.pr_base - 0.15) < 1e-12 and c.cache_policy == d.CACHE_HOTNESS
    assert d.lib().dcg_status_string(-4) == b"capacity exhausted"
    assert d.lib().dcg_config_init(None) == -1


def test_create_rejects_bad_pagerank_config_without_gpu():
    # samples/hybrid_pr/hybrid_pr.cu:133-137: error > 0, 1 <= pr_max_rounds <= 100
    import dycgraph_amd as d
    from dycgraph_amd.engine import Config
    c = Config()
    d.lib().dcg_config_init(C.byref(c))
    c.app = d.APP_PR
    c.pr_max_rounds = 101
    h = C.c_void_p()
    assert d.lib().dcg_create(C.byref(c), C.byref(h)) == -1
    c.pr_max_rounds = 10
    c.pr_epsilon = 0.0
    assert d.lib().dcg_create(C.byref(c), C.byref(h)) == -1


def test_rmat_generator_properties():
    import dycgraph_amd as d
    src, dst = d.gen_rmat(12, 30000, seed=5)
    keys = (src.astype(np.uint64) << 32) | dst
    assert len(np.unique(keys)) == 30000                  # distinct
    assert np.all(src != dst)                             # no self-loops
    assert np.all(np.diff(keys.astype(np.float64)) > 0)   # sorted by (src, dst)
    assert src.max() < 4096 and dst.max() < 4096
    s2, d2 = d.gen_rmat(12, 30000, seed=5)
    assert np.array_equal(src, s2) and np.array_equal(dst, d2)
    # a=0.57 skew: vertex 0 has the largest out-degree
    deg = np.bincount(src, minlength=4096)
    assert deg.argmax() == 0 and deg[0] > 20 * deg.mean()
    # prefix property: the first k distinct draws are a subset of the first k+m
    s3, d3 = d.gen_rmat(12, 20000, seed=5)
    small = set(((s3.astype(np.uint64) << 32) | d3).tolist())
    assert small <= set(keys.tolist())


def test_stream_generator_properties():
    import dycgraph_amd as d
    import oracle
    n = 1 << 10
    src, dst = d.gen_rmat(10, 6000, seed=2)
    adds, dels = d.gen_stream(n, src, dst, [300] * 4, [300] * 4, seed=9)
    present = set(zip(src.tolist(), dst.tolist()))
    for s in range(4):
        dl = [tuple(x) for x in dels[300 * s:300 * (s + 1)].tolist()]
        ad = [tuple(x) for x in adds[300 * s:300 * (s + 1)].tolist()]
        assert len(set(dl)) == 300 and all(e in present for e in dl)     # existing, without replacement
        assert len(set(ad)) == 300 and all(u != v and (u, v) not in present for u, v in ad)
        present -= set(dl)
        present |= set(ad)
    # the multiset model sees no missing deletes on this stream
    store = oracle.Store(n, src, dst)
    for s in range(4):
        m = store.apply(dels[300 * s:300 * (s + 1)], adds[300 * s:300 * (s + 1)])
        assert m["missing_deletes"] == 0 and m["invalid_additions"] == 0
    assert store.edges() == 6000
