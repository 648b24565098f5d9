"""Build libgtps.so in-tree with hipcc for gfx950 (csrc/Makefile)."""
from __future__ import annotations

import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(_HERE, "csrc")


def build(jobs: int = 8, force: bool = True) -> str:
    """Compile every HIP source of the library; returns the .so path."""
    args = ["make", "-C", CSRC, f"-j{jobs}"]
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True)
    return os.path.join(_HERE, "libgtps.so")


if __name__ == "__main__":
    print(build())
