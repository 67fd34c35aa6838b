"""Euler matrices and angular grids (host, FP64, through the C ABI).

BH_defineMatrix -> emc_define_matrix (coordinates/BH_defineMatrix.m:51-126)
eulerSearch     -> emc_euler_search  (coordinates/eulerSearch.m:38-177)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

_CONV = {"Bah": 0, "TILT": 1, "SPIDER": 2, "Helical": 3, "IMOD": 4}
_DIR = {"forward": 0, "inv": 1, "forwardvector": 2, "invvector": 3}


def BH_defineMatrix(angles, convention="Bah", direction="forward") -> np.ndarray:
    """3x3 rotation matrix (numpy, FP64) for Euler angles in degrees."""
    if convention not in _CONV:
        raise ValueError(f"Convention must be Bah,SPI,Helical, not {convention}")
    d = direction.lower()
    if d not in _DIR:
        raise ValueError(f"Direction must be forward or inv, not {direction}")
    a = np.zeros(3)
    v = np.atleast_1d(np.asarray(angles, dtype=np.float64))
    a[: v.size] = v
    out = (C.c_double * 9)()
    _lib.call("emc_define_matrix", _lib.d_arr(a), _CONV[convention], _DIR[d], out)
    return np.array(out[:], dtype=np.float64).reshape(3, 3, order="F")


def eulerSearch(symmetry, theta_max, theta_step, psi_max, psi_step) -> np.ndarray:
    """(N, 3) array of (phi, theta, psi) in the reference loop order
    theta -> phi -> psi (random start off, as parity requires)."""
    n = C.c_int64()
    sym = symmetry.encode()
    _lib.call("emc_euler_search", sym, float(theta_max), float(theta_step), float(psi_max),
              float(psi_step), None, C.c_int64(0), C.byref(n))
    out = np.zeros((max(n.value, 1), 3))
    _lib.call("emc_euler_search", sym, float(theta_max), float(theta_step), float(psi_max),
              float(psi_step), out.ctypes.data_as(C.POINTER(C.c_double)), C.c_int64(n.value),
              C.byref(n))
    return out[: n.value]


def symmetry_matrices(symmetry, convention="Bah"):
    """C_n operators of interpolator.m:188-198 (S_i = BH_defineMatrix([0,0,i*360/n]))."""
    if symmetry[0] != "C":
        raise NotImplementedError(f"symmetry {symmetry}: only C_n is implemented")
    n = int(symmetry[1:])
    return [BH_defineMatrix([0.0, 0.0, i * 360.0 / n], convention, "forward") for i in range(n)]
