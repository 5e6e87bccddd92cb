"""In-tree build of the gfx950 engine (hipcc via dtcf_amd/csrc/Makefile)."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")


def build(jobs: int = 8, force: bool = True) -> None:
    """Compile every HIP source for gfx950 into dtcf_amd/libdtcf.so and the
    C++ drop-in library dtcf_amd/libdtcf_dropin.so."""
    if force:
        subprocess.run(["make", "-C", CSRC, "clean"], check=True, stdout=subprocess.DEVNULL)
    subprocess.run(["make", "-C", CSRC, f"-j{jobs}", "all"], check=True)


if __name__ == "__main__":
    build()
