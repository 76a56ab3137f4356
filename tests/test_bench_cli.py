"""CPU checks of bench.py's contract and host logic (no GPU needed): more
than one GPU is refused with a message, stored counter traffic is used only
for the plan shape it was measured on, and the selection label names how a
plan was chosen."""
import importlib.util
import json
import os
import subprocess
import sys
from types import SimpleNamespace

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_stored_traffic_matches_only_its_plan_shape():
    b = _bench()
    with open(os.path.join(ROOT, "profiles", "traffic_stencil.json")) as f:
        tj = json.load(f)
    info = SimpleNamespace(variant=tj["variant"], ipt=tj["ipt"], grid=tj["grid"], lanes=tj.get("lanes", 64))
    assert b.stored_traffic("stencil", info) == tj["hbm_bytes_per_launch"]
    other = SimpleNamespace(variant="S", ipt=3, grid=-1, lanes=64)
    assert b.stored_traffic("stencil", other) is None
    assert b.stored_traffic("no_such_workload", info) is None


def test_selection_labels():
    b = _bench()
    g = SimpleNamespace(variant="G", ipt=8, grid=-1, lanes=64)
    assert b.selection_label(None, None, None, g, 1) == "forced"
    assert b.selection_label(None, None, "rule", g, 1) == "rule (pattern rule)"
    assert b.selection_label(None, None, "autotune", g, 1) == "autotune"
    assert b.selection_label(None, None, "learned", g, 8) == "learned"
