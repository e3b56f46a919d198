"""Builds the package's native libraries in-tree (see vqm_amd/Makefile).

    python -m vqm_amd.build          # libvqm_hip.so (gfx950) and libvqm_synth.so
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs: int = 8, force: bool = False) -> None:
    if force:
        subprocess.run(["make", "-C", HERE, "clean"], check=True, stdout=subprocess.DEVNULL)
    subprocess.run(["make", "-C", HERE, f"-j{jobs}", "all"], check=True)


if __name__ == "__main__":
    build(force="--force" in sys.argv)
