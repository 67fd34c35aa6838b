"""ctypes wrapper of oracle/lib/libemc_cpu.so (emc_cpu.c): the FP32 OpenMP
restatement of the plain two-stage alignment body. Test infrastructure only:
the benchmark's CPU baseline and a checker in the tests."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from .emc_oracle import define_matrix

LIB = Path(__file__).resolve().parent / "lib" / "libemc_cpu.so"
_L = None


def load():
    global _L
    if _L is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} is missing: build it with `python -m oracle.build`")
        _L = C.CDLL(str(LIB))
        f = C.POINTER(C.c_float)
        d = C.POINTER(C.c_double)
        _L.emc_cpu_align_grid.restype = C.c_int
        _L.emc_cpu_align_grid.argtypes = [C.c_int, f, f, f, f, f, f, d, C.c_int, d, C.c_int, d, d]
        _L.emc_cpu_max_threads.restype = C.c_int
    return _L


def max_threads() -> int:
    return int(load().emc_cpu_max_threads())


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def align_grid(particle, ref, vol_mask, peak_mask, bandpass, wedge, A, grid_angles, est0=(0, 0, 0), threads=0):
    """CCC and shifts of every orientation A @ defineMatrix([phi, theta,
    psi - phi], 'Bah', 'inv') of grid_angles (the coarse loop of
    BH_alignRaw3d_v2.m:1119-1419, window = calc size, no reference weight,
    C1). threads 0 = the OpenMP default."""
    lib = load()
    n = int(np.asarray(particle).shape[0])
    A = np.asarray(A, dtype=np.float64).reshape(3, 3)
    pts = np.asarray(grid_angles, dtype=np.float64).reshape(-1, 3)
    mats = np.ascontiguousarray(np.stack([A @ define_matrix([p, t, s - p], "Bah", "inv") for p, t, s in pts]))
    vols = [_f(v) for v in (particle, ref, vol_mask, peak_mask, bandpass, wedge)]
    ccc = np.zeros(len(pts))
    est = np.zeros((len(pts), 3))
    e0 = np.ascontiguousarray(np.asarray(est0, dtype=np.float64).reshape(3))
    pf = C.POINTER(C.c_float)
    pd = C.POINTER(C.c_double)
    st = lib.emc_cpu_align_grid(n, *[v.ctypes.data_as(pf) for v in vols], mats.ctypes.data_as(pd), len(pts),
                                e0.ctypes.data_as(pd), int(threads), ccc.ctypes.data_as(pd), est.ctypes.data_as(pd))
    if st != 0:
        raise IndexError("failed to box out COM calc")
    return ccc, est
