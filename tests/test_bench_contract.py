"""bench.py's contract on the CPU: the one-GPU guard and the committed
traffic summary it reads for the roofline (per launch shape)."""
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_traffic_per_launch_shape_from_the_committed_profile():
    # FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (KiB), as the profile's table lists them
    key = "traffic_bytes_per_launch[spw={},xcd={}]".format
    assert bench.profiled_traffic("k_xyz_sweep", "sc_heis_afm_32", key(4, 2)) == round((43371.7 * 2 + 38427.8) * 1024)
    assert bench.profiled_traffic("k_xyz_sweep", "sc_heis_afm_32", key(4, 1)) is None   # another shape
    assert bench.profiled_traffic("k_xyz_sweep", "sc_heis_afm_32", key(3, 2)) is None
    assert bench.profiled_traffic("k_xyz_sweep", "pyro_pt_16", key(2, 0)) == round((151782.8 * 2 + 12534.6) * 1024)
    # the resident Ising kernel: per sweep (a launch runs the 10 sweeps between samples)
    assert bench.profiled_traffic("k_ising_resident", "sc_ising_16", "traffic_bytes_per_sweep") == \
        int((3657.0 * 2 + 6150.7) * 1024 / 10)
    assert bench.profiled_traffic("k_ising_sweep", "sc_ising_16", key(1, 2)) is None


def test_more_than_one_gpu_exits_with_a_message():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode != 0
    assert "one GPU" in p.stderr
