"""bench.py's contract on the CPU: the one-GPU guard and the committed
traffic summary it reads for the roofline (per launch shape)."""
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_traffic_per_launch_shape_from_the_committed_profile():
    # FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (KiB), as the profile's table lists them
    assert bench.profiled_traffic("k_xyz_sweep", "sc_heis_afm_32", 16) == round((48316.7 * 2 + 37190.7) * 1024)
    assert bench.profiled_traffic("k_xyz_sweep", "sc_heis_afm_32", 4) == round((70677.2 * 2 + 38271.8) * 1024)
    assert bench.profiled_traffic("k_xyz_sweep", "sc_heis_afm_32", 3) is None
    assert bench.profiled_traffic("k_ising_sweep", "sc_ising_16", 1) is None


def test_more_than_one_gpu_exits_with_a_message():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode != 0
    assert "one GPU" in p.stderr
