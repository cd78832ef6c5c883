"""Build recipe for the oracle's C restatement (test infrastructure only).

    python oracle/build_oracle.py      ->  oracle/_build/libceep_oracle.so

gcc -O2 -fopenmp -ffp-contract=off: no fused multiply-add and no reassociation,
so the FP64 results are bitwise those of the NumPy oracle.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_build", "libceep_oracle.so")


def build(verbose: bool = False) -> str:
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    cmd = ["gcc", "-O2", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
           "-std=c11", os.path.join(HERE, "fdtd_ref.c"), "-o", OUT + ".tmp"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
