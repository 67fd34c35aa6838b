"""Build libemc_hip.so (HIP for gfx950 + rocFFT) in-tree.

Usage: python -m emclarity_amd.csrc.build [--jobs N]

Compiles every ``*.hip`` source of this directory with hipcc for gfx950 into
``emclarity_amd/lib/libemc_hip.so``. The object files go to ``_build/`` at
the repository root. Always recompiles (the build check must compile).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
PKG = HERE.parent
ROOT = PKG.parent
OBJ = ROOT / "_build" / "emc_hip"
LIB = PKG / "lib" / "libemc_hip.so"

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

CFLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # packed FP32 (v_pk_*_f32) forms are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    f"-I{ROOT / 'include'}",
]
LDFLAGS = ["-shared", f"-L{ROCM}/lib", "-lrocfft", f"-Wl,-rpath,{ROCM}/lib"]


def sources():
    return sorted(HERE.glob("*.hip"))


def compile_one(src: Path) -> tuple[Path, str]:
    obj = OBJ / (src.stem + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stdout}\n{r.stderr}")
    warn = "\n".join(l for l in (r.stderr or "").splitlines()
                     if "packed-fp32-ops" not in l and l.strip())
    return obj, warn


def build(jobs: int | None = None, verbose: bool = False) -> Path:
    OBJ.mkdir(parents=True, exist_ok=True)
    LIB.parent.mkdir(parents=True, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, len(srcs))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(compile_one, srcs))
    for obj, warn in results:
        if verbose and warn:
            print(f"[{obj.name}] {warn}", file=sys.stderr)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, *[str(o) for o, _ in results], "-o", str(tmp), *LDFLAGS]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_harness()
    return LIB


HARNESS_SRC = PKG / "harness" / "emc_harness.c"
HARNESS_BIN = PKG / "lib" / "emc_harness"


def build_harness() -> Path:
    """The plain C harness, linked against the library (gcc, no HIP headers)."""
    cmd = ["gcc", "-O2", "-Wall", f"-I{ROOT / 'include'}", str(HARNESS_SRC), f"-L{LIB.parent}",
           "-lemc_hip", f"-Wl,-rpath,{LIB.parent}", "-lm", "-o", str(HARNESS_BIN)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"harness build failed:\n{r.stdout}\n{r.stderr}")
    return HARNESS_BIN


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))
