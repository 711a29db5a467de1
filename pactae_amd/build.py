"""Compile the HIP library for gfx950 (hipcc cross-compiles; no GPU needed)."""
from __future__ import annotations

import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(_HERE, "csrc")


def build(jobs: int = 8, force: bool = True) -> str:
    """Build pactae_amd/_lib/libsmp_hip.so; force=True recompiles every source."""
    cmd = ["make", "-C", CSRC, f"-j{jobs}"]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, check=True)
    return os.path.join(_HERE, "_lib", "libsmp_hip.so")


if __name__ == "__main__":
    print(build())
