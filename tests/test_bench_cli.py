"""bench.py's host-side contract: more than one GPU is refused with a
message, and the workload line is read from the case file."""
import importlib.util
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_module", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


def test_workload_line_follows_the_case_file():
    b = _bench()
    d = b.describe(os.path.join(ROOT, "cases", "shockvortex256"), "shockvortex256", 16581375)
    assert d == ("shockvortex256: shock-vortex interaction, 16581375 interior fluid nodes, WENO5-LLF-RK3 split, "
                 "inviscid, plain FP64 path")
    assert "viscous Taylor-Green" in b.describe(os.path.join(ROOT, "cases", "tgv128"), "tgv128", 1)
    assert ", viscous," in b.describe(os.path.join(ROOT, "cases", "tgv128"), "tgv128", 1)
    assert "Roe-LLF" in b.describe(os.path.join(ROOT, "cases", "v_roe"), "v_roe", 1)
    assert "-SW-" in b.describe(os.path.join(ROOT, "cases", "v_sw"), "v_sw", 1)
