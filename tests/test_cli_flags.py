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
