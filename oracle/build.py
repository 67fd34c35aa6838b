"""Compile the oracle's C restatement (emc_cpu.c) into oracle/lib/libemc_cpu.so.

Test infrastructure: the library is the CPU baseline of the benchmark and a
checker in the tests; the product never loads it. gcc with OpenMP, x86-64-v3
(AVX2 + FMA) so the same build runs on the GPU hosts.
"""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB = HERE / "lib" / "libemc_cpu.so"


def build_cpu() -> Path:
    LIB.parent.mkdir(parents=True, exist_ok=True)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-fopenmp", "-std=gnu11", "-Wall", "-shared", "-fPIC",
           str(HERE / "emc_cpu.c"), "-o", str(tmp), "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"gcc failed for emc_cpu.c:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_cpu())
