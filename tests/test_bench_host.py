"""bench.py's host-side parts, on the CPU: the CPU baseline's timing lines
come from a child process (its OpenMP threads bound) and leave the calling
process's CPU affinity as it was; more than one GPU is refused with a
message."""
import json
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def test_cpu_rate_lines_in_a_bound_child():
    oracle = pytest.importorskip("oracle")
    if not hasattr(oracle, "time_advance_f32"):
        pytest.skip("the oracle library is not built")
    import bench
    before = os.sched_getaffinity(0)
    lines, bind = bench.cpu_rate_lines_bound("elastic_sphere", 5, 1, c1_steps=4)
    assert os.sched_getaffinity(0) == before
    assert bind == (f"{os.environ.get('OMP_PROC_BIND', 'close')}, places "
                    f"{os.environ.get('OMP_PLACES', 'cores')}")
    assert [(l["workload"], l["threads"], l["substeps"], l["repeats"]) for l in lines] == [
        ("elastic_sphere", 1, 1, 5), ("elastic_sphere", 1, 1, 3), ("elastic_sphere", 1, 4, 5),
        ("elastic_sphere", 1, 4, 5)]
    for l in lines:
        assert l["particles"] == 10000 and l["value"] > 0 and l["median_s"] > 0


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU only" in r.stderr and r.stdout == ""
