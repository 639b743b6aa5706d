"""Command-line driver: python -m ether_amd [run_dir] | --batch manifest.

Runs an ETHER input directory (in.ether, structure.vasp, j_exchange, and the
optional single_ion_anisotropy, staggered and gpu.ether) on the GPU engine and
writes the reference's result files into it.  With --batch, every run
directory listed in the manifest (batch.ether) is run, the runs with equal
schedules together in one batch.  A failure prints a remark and exits with
status 1, like the reference's terminate (src/terminate.f90:20-35).
"""
from __future__ import annotations

import argparse
import os
import sys

from . import driver
from .engine import EtherError
from .inputs import BATCH_FILE, EtherInputError, parse_batch_manifest, read_directory


def _summary(path: str, res) -> str:
    rate = res.updates / res.wall_s if res.wall_s > 0 else float("nan")
    return (f" ==> {path}: {len(res.temps)} T, ladder {res.ladder_source}, "
            f"{res.wall_s:.3f} s, {rate:.2e} spin updates/s")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m ether_amd", description=__doc__.split("\n")[0])
    ap.add_argument("run_dir", nargs="?", default=".")
    ap.add_argument("--batch", metavar="MANIFEST", help=f"a {BATCH_FILE} listing run directories")
    a = ap.parse_args(argv)
    try:
        if a.batch:
            with open(a.batch) as f:
                dirs = parse_batch_manifest(f.read(), os.path.dirname(os.path.abspath(a.batch)))
            models = [read_directory(d) for d in dirs]
            for d, res in zip(dirs, driver.run_batch(models)):
                res.write(d)
                print(_summary(d, res), flush=True)
        else:
            res = driver.run(read_directory(a.run_dir))
            res.write(a.run_dir)
            print(_summary(a.run_dir, res), flush=True)
    except (EtherInputError, EtherError, OSError) as e:
        print(f"\n ==> ETHER (GPU) stopped: {e}\n", flush=True)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
