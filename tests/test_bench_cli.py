"""bench.py's contract on the host side: it measures one MI355X and says so
when asked for more, before touching any device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one MI355X" in r.stderr
    assert r.stdout == ""


def test_traffic_figure_comes_from_the_committed_counter_summary():
    sys.path.insert(0, ROOT)
    import bench
    t = bench.measured_traffic()
    assert t is not None and 1e8 < t < 1e11
