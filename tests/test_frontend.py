"""Input front-end extensions: gpu.ether and batch manifests.

gpu.ether holds the engine settings the reference has no tags for; in.ether
stays a valid reference input, and the reference's rule that an unknown key
stops the run (src/read_input.f90:144-146) holds in both files.
"""
import os
import subprocess
import sys

import pytest

from ether_amd import workloads
from ether_amd.inputs import (EtherInputError, GpuTags, parse_batch_manifest, parse_gpu_ether,
                              read_directory)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_gpu_tags_and_defaults():
    g = parse_gpu_ether("# engine settings\nladder = learned 2000 3\nautotune = 5 ! timed sweeps\n"
                        "launch = 8, 0\nlanes = 32\nGraphs = .F.\ncompat = 3\nchunk = 250\n")
    assert g == GpuTags(ladder="learned", train_mcs=2000, windows=3, autotune=5, spw=8, xcd=0,
                        lanes=32, graphs=False, compat=3, chunk=250)
    assert parse_gpu_ether("") == GpuTags()
    assert parse_gpu_ether("ladder = linear").ladder == "linear"


@pytest.mark.parametrize("text", ["pt = .T. 11", "ladder = geometric", "launch = 3 1", "lanes = 65",
                                  "compat = 16", "chunk = -1", "autotune", "ladder = learned 10 0"])
def test_gpu_tags_errors(text):
    with pytest.raises(EtherInputError):
        parse_gpu_ether(text)


def test_directory_reads_gpu_file_and_in_ether_stays_strict(tmp_path):
    workloads.write_dir("sc_heis_afm_32", str(tmp_path), sc="4 4 4")
    assert read_directory(str(tmp_path)).gpu == GpuTags()
    (tmp_path / "gpu.ether").write_text("launch = 2 1\ngraphs = .F.\n")
    g = read_directory(str(tmp_path)).gpu
    assert (g.spw, g.xcd, g.graphs) == (2, 1, False)
    # an engine tag inside in.ether is an unknown key, as the reference would say
    with open(tmp_path / "in.ether", "a") as f:
        f.write("autotune = 3\n")
    with pytest.raises(EtherInputError, match="unknown"):
        read_directory(str(tmp_path))


def test_batch_manifest(tmp_path):
    dirs = parse_batch_manifest("runs/a\n# comment\n  runs/b   ! second\n/abs/c\n", str(tmp_path))
    assert dirs == [str(tmp_path / "runs" / "a"), str(tmp_path / "runs" / "b"), "/abs/c"]
    with pytest.raises(EtherInputError):
        parse_batch_manifest("a\na\n", str(tmp_path))
    with pytest.raises(EtherInputError):
        parse_batch_manifest("# nothing\n", str(tmp_path))


def test_cli_terminates_on_bad_input(tmp_path):
    workloads.write_dir("sc_ising_16", str(tmp_path), sc="4 4 4")
    (tmp_path / "gpu.ether").write_text("warp = 32\n")
    r = subprocess.run([sys.executable, "-m", "ether_amd", str(tmp_path)], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 1
    assert "ETHER (GPU) stopped" in r.stdout and "gpu.ether" in r.stdout
