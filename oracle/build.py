"""Build recipe for the CPU oracle (liboracle.so, plain C, gcc).

The reference itself cannot be built here: it needs nvcc (CUDA 5.5, sm_20)
and its util/ submodule, whose sources are absent, so there is no oracle/_ref.
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build():
    tmp = LIB + ".tmp"
    subprocess.run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-Wall", "-Wextra",
                    os.path.join(HERE, "oracle.c"), "-o", tmp], check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
