"""Host logic of bench.py that needs no GPU: the one-GPU guard and the per-step HBM traffic
lookup from the committed counter summary (profiles/traffic.json)."""
import argparse
import importlib.util
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_host", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_step_traffic_lookup():
    b = _bench()
    tab = json.load(open(os.path.join(ROOT, "profiles", "traffic.json")))
    for math, visc in (("exact", "sutherland"), ("exact", "constant"), ("fp32", "sutherland")):
        t = b.step_traffic(None, argparse.Namespace(math=math, viscosity=visc), 256)
        per = tab["per_step"][f"{math}_256_{visc}"]
        assert t["bytes_per_step"] == per
        assert t["bytes_per_cell_step"] == round(per / 256 ** 3, 1)
        assert "committed lookup" in t["source"]
        # a step moves more than the ideal fused 320 B per cell-step and less than the sum of
        # every launch of the counter runs
        assert 320 * 256 ** 3 < per < sum(tab["per_launch"][f"{math}_256_{visc}"].values()) * 3
    # no entry: no traffic figure rather than a wrong one
    t = b.step_traffic(None, argparse.Namespace(math="fast", viscosity="sutherland"), 256)
    assert t["bytes_per_step"] is None
    t = b.step_traffic(None, argparse.Namespace(math="exact", viscosity="sutherland"), 128)
    assert t["bytes_per_step"] is None
