"""bench.py parses and keeps its contract on the CPU: every configuration
table loads, and --gpus above 1 exits with code 2 and a message."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_configurations_load():
    sys.path.insert(0, ROOT)
    import bench
    assert "sw025" in bench.CONFIGS and "ffsl05" in bench.ADV_CONFIGS
    assert "ens64" in bench.ENS_CONFIGS and "adj1" in bench.ADJ_CONFIGS
    members = bench.ENS_CONFIGS["ens64"]["members"]
    assert sum(m[3] for m in members) == 64


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
