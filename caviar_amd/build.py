"""Build the in-tree HIP library (gfx950) and the C++ host layer on top of it:
``python -m caviar_amd.build``."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs: int = 8) -> str:
    for sub in ("csrc", "host"):
        subprocess.run(["make", "-s", "-C", os.path.join(HERE, sub), f"-j{jobs}"], check=True)
    return os.path.join(HERE, "libcaviar_amd.so")


if __name__ == "__main__":
    print(build(int(sys.argv[1]) if len(sys.argv) > 1 else 8))
