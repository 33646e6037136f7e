"""Build the HIP library ``liblbm_hip.so`` in-tree (gfx950)."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "liblbm_hip.so")


def build(jobs: int = 4, force: bool = True) -> str:
    """Compile every HIP translation unit for gfx950 and link the library."""
    if force:
        for name in os.listdir(CSRC):
            if name.endswith(".o"):
                os.remove(os.path.join(CSRC, name))
    subprocess.run(["make", "-s", "-j", str(jobs), "-C", CSRC], check=True)
    return LIB


if __name__ == "__main__":
    print(build())
