"""CPU checks of bench.py's command-line contract (no GPU needed).

* `--gpus N` with N > 1 exits with a message and a non-zero status before
  touching a device: the build measures one GPU.
* The workloads bench.py offers build their problems on the host.
"""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_multi_gpu_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr


def test_workload_table():
    sys.path.insert(0, ROOT)
    import bench
    import csipm.workloads as wl
    assert set(bench.WORKLOADS) >= {"1d1024", "2d255", "3d255", "uneven"}
    probs = bench.make_problems(wl, "2d255", 2)
    assert len(probs) == 2 and probs[0].dims == (255, 255)
