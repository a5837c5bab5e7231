"""Build recipe for the CPU oracle: liboracle.so (the literal restatement,
plain C) and libbitsliced.so (the bit-sliced OpenMP grader), both gcc.

The reference itself cannot be built here: it needs nvcc (CUDA 5.5, sm_20)
and its util/ submodule, whose sources are absent, so there is no oracle/_ref.
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
BSL_LIB = os.path.join(HERE, "libbitsliced.so")


def _cc(src: str, out: str, extra: list[str]) -> None:
    tmp = out + ".tmp"
    subprocess.run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-Wall", "-Wextra", *extra,
                    os.path.join(HERE, src), "-o", tmp], check=True)
    os.replace(tmp, out)


def build():
    _cc("oracle.c", LIB, [])
    _cc("bitsliced.c", BSL_LIB, ["-fopenmp"])
    return LIB


if __name__ == "__main__":
    print(build())
