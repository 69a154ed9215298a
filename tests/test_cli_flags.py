"""The hybrid_* front ends parse the reference's flags and reject bad ones
before touching the GPU (CPU only)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "dycgraph_amd", "bin")


def _run(app, *args):
    return subprocess.run([os.path.join(BIN, f"hybrid_{app}"), *args], capture_output=True, text=True, timeout=30)


@pytest.mark.parametrize("app", ["sssp", "bfs", "cc", "pr"])
def test_unknown_flag_is_rejected(app):
    r = _run(app, "--graphfile=/nonexistent", "--no_such_flag=1")
    assert r.returncode == 1 and "unknown command line flag 'no_such_flag'" in r.stderr


def test_pr_flag_contract():
    # samples/hybrid_pr/hybrid_pr.cu:133-141
    r = _run("pr", "--graphfile=/x", "--pr_max_rounds=101")
    assert r.returncode == 1 and "1 <= pr_max_rounds <= 100" in r.stderr
    r = _run("pr", "--graphfile=/x", "--error=0")
    assert r.returncode == 1
    r = _run("pr", "--graphfile=/x", "--updatefile=/u")
    assert r.returncode == 1 and "both updatefile and update_size" in r.stderr


def test_reference_command_line_parses(tmp_path):
    # README "Execute Algorithms" form: single-dash flags with '=' and bare booleans
    g = tmp_path / "g"
    g.write_text("")
    r = _run("sssp", f"-graphfile={g}", "-format", "market_big", "-hybrid", "0", "-SEGMENT", "32", "-weight_num", "1",
             "-weight", "-cache", "4", "-source_node", "0", "--check=true", "--sssp_max_batches=10", "-nosparse")
    # the graph is empty, so the run stops at the graph read, after the flags parsed
    assert r.returncode == 1 and "cannot read graph" in r.stderr


@pytest.mark.parametrize("var,value,msg", [
    ("CG_INSERTION_SCHEDULE", "warp", "CG_INSERTION_SCHEDULE must be block, thread or ordered"),
    ("CG_BATCH_MAINTENANCE", "huge", "CG_BATCH_MAINTENANCE must be regular, large or auto"),
])
def test_algorithm_hints_are_validated(var, value, msg):
    # framework.cuh:3775-3783, effective_update_batch.h:40-45: a bad value is an error
    env = dict(os.environ, **{var: value})
    r = subprocess.run([os.path.join(BIN, "hybrid_sssp"), "--graphfile=/nonexistent"], capture_output=True,
                       text=True, timeout=30, env=env)
    assert r.returncode == 1 and msg in r.stderr
    # CG_ORDERED_REPAIR=1 forces the ordered schedule, so a stale schedule value is not read
    if var == "CG_INSERTION_SCHEDULE":
        env["CG_ORDERED_REPAIR"] = "1"
        r = subprocess.run([os.path.join(BIN, "hybrid_sssp"), "--graphfile=/nonexistent"], capture_output=True,
                           text=True, timeout=30, env=env)
        assert r.returncode == 1 and "cannot read graph" in r.stderr


def test_coexecution_flag_contract(tmp_path):
    # framework.cuh:925-936: a map and a capacity together, a capacity above SEGMENT
    r = _run("sssp", "--graphfile=/x", "--sssp_cpu_domain_map=/m", "--sssp_cpu_partition_capacity=2")
    assert r.returncode == 1 and "domain_map_conflicts_with_capacity" in r.stdout + r.stderr
    r = _run("sssp", "--graphfile=/x", "--sssp_cpu_partition_capacity=9", "--SEGMENT=8")
    assert r.returncode == 1 and "invalid_cpu_capacity" in r.stdout + r.stderr
    r = _run("pr", "--graphfile=/x", "--sssp_cpu_partition_capacity=1", "--SEGMENT=8")
    assert r.returncode == 1
    # a map of the wrong size, or with an owner other than 0/1, stops before the device is used
    g = tmp_path / "g"
    g.write_text("0 1\n1 2\n")
    m = tmp_path / "m"
    m.write_bytes(b"\0\0\1\0")
    r = _run("sssp", f"--graphfile={g}", f"--sssp_cpu_domain_map={m}")
    assert r.returncode == 1 and "invalid_map" in r.stdout + r.stderr and "expected_bytes=6" in r.stdout + r.stderr
    m.write_bytes(b"\0\0\1\0\2\0")
    r = _run("sssp", f"--graphfile={g}", f"--sssp_cpu_domain_map={m}")
    assert r.returncode == 1 and "invalid_owner vertex=2 owner=2" in r.stdout + r.stderr
