"""bench.py's command-line contract, on the CPU: more than one GPU is refused with a
message before anything touches a device, and the per-size leg lists the grid sizes
of the workload's configurations."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_config_runs_cover_the_workload_sizes():
    import bench
    sizes = [n for _, n, _, _, _ in bench.CONFIG_RUNS]
    assert sizes == [32, 128, 512, 1024]
    # the 32^3 configuration is the fixed-dt Beltrami case; the large grids start from
    # the random-phase field
    assert bench.CONFIG_RUNS[0][2:4] == ("beltrami", 0.01)
    assert all(ic == "random" for _, n, ic, _, _ in bench.CONFIG_RUNS if n >= 512)
