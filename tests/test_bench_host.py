"""CPU tests of the measurement contract: bench.py runs one GPU only and says so for more, the CPU
baseline uses the CPUs this job may run on, and the entry module exposes build() and smoke()."""

import os
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=300, cwd=ROOT)
    assert r.returncode == 2
    assert "single GPU" in r.stderr and not r.stdout.strip()


def test_cpu_baseline_threads_follow_the_job_share(monkeypatch):
    sys.path.insert(0, str(ROOT))
    import bench

    usable = len(os.sched_getaffinity(0))
    monkeypatch.setenv("OMP_NUM_THREADS", "1")
    assert bench.host_threads() == 1
    monkeypatch.setenv("OMP_NUM_THREADS", str(usable + 100))
    assert bench.host_threads() == usable
    monkeypatch.delenv("OMP_NUM_THREADS")
    assert bench.host_threads() == usable
    info = bench.host_info()
    assert info["usable_cpus"] == usable and info["os_cpu_count"] == os.cpu_count()


def test_entry_module_exposes_build_and_smoke():
    sys.path.insert(0, str(ROOT))
    import __graft_entry__ as entry

    assert callable(entry.build) and callable(entry.smoke)
