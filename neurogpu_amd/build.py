"""Build the native parts in-tree: libneurogpu (hipcc, gfx950) and the CPU
oracle used by the tests (gcc)."""
from __future__ import annotations

import os
import subprocess
import sys

_HERE = os.path.dirname(os.path.abspath(__file__))
_ROOT = os.path.dirname(_HERE)


def build(jobs: int = 8, force: bool = False, oracle: bool = True) -> str:
    """Compile every HIP source for gfx950 into neurogpu_amd/_build/libneurogpu.so
    (always re-runs make; ``force`` rebuilds from scratch)."""
    csrc = os.path.join(_HERE, "csrc")
    args = ["make", f"-j{min(jobs, 16)}"] + (["-B"] if force else [])
    subprocess.run(args, cwd=csrc, check=True)
    if oracle:
        subprocess.run(["make", "-s"] + (["-B"] if force else []), cwd=os.path.join(_ROOT, "oracle"), check=True)
    return os.path.join(_HERE, "_build", "libneurogpu.so")


if __name__ == "__main__":
    print(build(force="--force" in sys.argv))
