"""Builds the oracle's C restatement into oracle/_build/liboracle.so (gcc).

Test infrastructure only.  -ffp-contract=off and plain SSE2 arithmetic give one
IEEE rounding per operation, as the restatement is written.

The reference's own sampler (nect/sampling/ct_sampling.cu) is not built: it
needs nvcc, CUDA ATen and curand, none of which exist here, so there is no
oracle/_ref build.
"""
from __future__ import annotations

import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "_build"
LIB = OUT / "liboracle.so"


def build() -> Path:
    OUT.mkdir(exist_ok=True)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
           str(HERE / "sampler.c"), "-o", str(tmp), "-lm"]
    subprocess.run(cmd, check=True)
    tmp.replace(LIB)
    return LIB


if __name__ == "__main__":
    print(build())
