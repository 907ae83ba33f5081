"""Build of libcup.so: hipcc for gfx950 through the Makefile in csrc/ (in-tree, so the
built library travels with the source tree)."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")


def build(jobs: int = 8, force: bool = False) -> str:
    """Compile every HIP/C++ source of the library; returns the path of libcup.so."""
    cmd = ["make", "-C", CSRC, f"-j{jobs}"]
    if force:
        subprocess.run(["make", "-C", CSRC, "clean"], check=True, stdout=subprocess.DEVNULL)
    subprocess.run(cmd, check=True)
    return os.path.join(HERE, "libcup.so")
