"""Build recipe of the oracle's C restatement (test infrastructure only).

Compiles oracle/yee_ref.c into oracle/liboref.so with gcc: -O2, no multiply-add
contraction and no fast-math, so each operation rounds once as in NumPy; OpenMP
splits independent cells only. Run ``python -m oracle.build``; __graft_entry__.build()
calls it.
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "yee_ref.c")
LIB = os.path.join(HERE, "liboref.so")


def build():
    tmp = LIB + ".tmp"
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-Wall",
           SRC, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
