"""bench.py's command line and workload table, without a GPU: every --workload name resolves
to a function, and --gpus > 1 exits with a message before touching the device."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_workload_table_resolves():
    sys.path.insert(0, ROOT)
    import bench
    table = bench.workloads()
    assert "learned_cg" in table and all(callable(f) for f in table.values())


def test_multi_gpu_request_is_refused():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=300)
    assert out.returncode == 2
    assert "one GPU" in json.loads(out.stdout.strip().splitlines()[-1])["error"]
