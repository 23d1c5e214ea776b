"""bench.py's contract on the CPU: it refuses more than one GPU with a message and a
nonzero status (before touching a device), and the committed counter summary carries
the per-launch HBM traffic lines it reports as roofline.traffic."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and "--gpus 1" in r.stderr
    assert r.stdout == ""


def test_traffic_lines_in_the_committed_profile():
    import bench

    for kernel in ("splat", "combine", "blur", "meanfield"):
        traffic, src = bench.measured_traffic(kernel)
        assert src == os.path.join("profiles", "meanfield_1024x768x21.txt")
        assert traffic is not None and traffic > 0, kernel
    # the splat moves ~141 MB of algorithmic bytes per launch at 1024x768x21; the
    # measured HBM traffic is of that order (not a stray unit)
    assert 1.0e8 < bench.measured_traffic("splat")[0] < 4.0e8
    assert bench.measured_traffic("no-such-kernel") == (None, None)
