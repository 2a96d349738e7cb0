"""Build recipe for libtensile_hip.so (hipcc, gfx950), driven by the Makefile
beside this file."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs: int = 8, force: bool = False) -> str:
    if force:
        subprocess.run(["make", "-s", "-C", HERE, "clean"], check=True)
    subprocess.run(["make", "-s", "-C", HERE, f"-j{jobs}"], check=True)
    return os.path.join(HERE, "libtensile_hip.so")


if __name__ == "__main__":
    print(build())
