"""MRC files: the @MRCImage surface the alignment and template-search path
uses (header, (sub)volume read, SAVE_IMG), over the C ABI.

Volumes come back as numpy (nz, ny, nx) float32 arrays: the same linear
order as MATLAB's (nx, ny, nz) column-major array (x fastest).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from .device import DeviceArray


class MrcHeader(C.Structure):
    """emc_mrc_header of include/emc_hip.h."""
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("mode", C.c_int32),
        ("nxstart", C.c_int32), ("nystart", C.c_int32), ("nzstart", C.c_int32),
        ("mx", C.c_int32), ("my", C.c_int32), ("mz", C.c_int32),
        ("cell", C.c_float * 3), ("cell_angle", C.c_float * 3),
        ("mapc", C.c_int32), ("mapr", C.c_int32), ("maps", C.c_int32),
        ("dmin", C.c_float), ("dmax", C.c_float), ("dmean", C.c_float),
        ("space_group", C.c_int16), ("n_symmetry_bytes", C.c_int16),
        ("n_bytes_extended", C.c_int32),
        ("creator_id", C.c_int16), ("n_bytes_per_section", C.c_int16), ("serialem_type", C.c_int16),
        ("imod_stamp", C.c_int32), ("imod_flags", C.c_int32),
        ("idtype", C.c_int16), ("lens", C.c_int16), ("nd1", C.c_int16), ("nd2", C.c_int16),
        ("vd1", C.c_int16), ("vd2", C.c_int16),
        ("tilt_angles", C.c_float * 6), ("origin", C.c_float * 3),
        ("map", C.c_ubyte * 4), ("machine_stamp", C.c_ubyte * 4),
        ("rms", C.c_float), ("n_labels", C.c_int32), ("labels", (C.c_ubyte * 80) * 10),
        ("swapped", C.c_int32), ("data_offset", C.c_int64),
    ]


MODE_DTYPES = {0: np.dtype(np.int8), 1: np.dtype(np.int16), 2: np.dtype(np.float32), 6: np.dtype(np.uint16)}


def _path(p) -> bytes:
    return os.fsencode(os.fspath(p))


def read_header(filename) -> dict:
    """MRCImage(filename, 0) + getHeader (@MRCImage/private/readHeader.m:43-186):
    the reference's field names, plus ``swapped`` and ``dataIndex``."""
    h = MrcHeader()
    _lib.call("emc_mrc_read_header", _path(filename), C.byref(h))
    nl = max(0, min(10, h.n_labels))
    return {
        "nX": h.nx, "nY": h.ny, "nZ": h.nz, "mode": h.mode,
        "nXStart": h.nxstart, "nYStart": h.nystart, "nZStart": h.nzstart,
        "mX": h.mx, "mY": h.my, "mZ": h.mz,
        "cellDimensionX": h.cell[0], "cellDimensionY": h.cell[1], "cellDimensionZ": h.cell[2],
        "cellAngleX": h.cell_angle[0], "cellAngleY": h.cell_angle[1], "cellAngleZ": h.cell_angle[2],
        "mapColumns": h.mapc, "mapRows": h.mapr, "mapSections": h.maps,
        "minDensity": h.dmin, "maxDensity": h.dmax, "meanDensity": h.dmean,
        "spaceGroup": h.space_group, "nSymmetryBytes": h.n_symmetry_bytes,
        "nBytesExtended": h.n_bytes_extended, "creatorID": h.creator_id,
        "nBytesPerSection": h.n_bytes_per_section, "serialEMType": h.serialem_type,
        "imodStamp": h.imod_stamp, "imodFlags": h.imod_flags,
        "idtype": h.idtype, "lens": h.lens, "ndl": h.nd1, "nd2": h.nd2, "vdl": h.vd1, "vd2": h.vd2,
        "tiltAngles": list(h.tilt_angles),
        "xOrigin": h.origin[0], "yOrigin": h.origin[1], "zOrigin": h.origin[2],
        "map": bytes(h.map), "machineStamp": bytes(h.machine_stamp), "densityRMS": h.rms,
        "nLabels": h.n_labels, "labels": [bytes(h.labels[i]) for i in range(nl)],
        "swapped": bool(h.swapped), "dataIndex": int(h.data_offset),
    }


def _range(r, n):
    """getVolume.m:45-118 index arguments: [] / -1 = all, scalar = one index,
    [a, b] = a..b (1-based, inclusive)."""
    if r is None:
        return 1, n
    r = np.atleast_1d(np.asarray(r, dtype=np.int64))
    if r.size == 0 or (r.size == 1 and r[0] == -1):
        return 1, n
    if r.size == 1:
        return int(r[0]), int(r[0])
    return int(r[0]), int(r[1])


class MRCImage:
    """@MRCImage opened on a file. ``MRCImage(name, 0)`` reads the header only
    (the form BH_multi_loadOrBuild.m:180 and BH_alignRaw3d_v2.m:1062 use);
    ``flgLoad`` 1 also reads the volume."""

    def __init__(self, filename, flgLoad: int = 1):
        self.filename = os.fspath(filename)
        self.header = read_header(self.filename)
        self.volume = None
        if flgLoad:
            self.volume = self.getVolume()

    def getHeader(self) -> dict:
        return dict(self.header)

    def getVolume(self, iRange=None, jRange=None, kRange=None) -> np.ndarray:
        """@MRCImage/getVolume.m:40-298 -> float32 (k, j, i) array."""
        h = self.header
        i0, i1 = _range(iRange, h["nX"])
        j0, j1 = _range(jRange, h["nY"])
        k0, k1 = _range(kRange, h["nZ"])
        if self.volume is not None and (i0, i1, j0, j1, k0, k1) == (1, h["nX"], 1, h["nY"], 1, h["nZ"]):
            return self.volume.copy()
        out = np.empty((max(k1 - k0 + 1, 0), max(j1 - j0 + 1, 0), max(i1 - i0 + 1, 0)), np.float32)
        _lib.call("emc_mrc_read_volume", _path(self.filename), (C.c_int * 6)(i0, i1, j0, j1, k0, k1),
                  out.ctypes.data_as(C.POINTER(C.c_float)))
        return out


def getVolume(m, iRange=None, jRange=None, kRange=None) -> np.ndarray:
    return m.getVolume(iRange, jRange, kRange)


def SAVE_IMG(volume, filename, pixelSize=None, origin=None):
    """SAVE_IMG(MRCImage(volume), filename[, pixelSize[, origin]])
    (@MRCImage/SAVE_IMG.m) for real volumes, written as mode 2. ``volume`` is
    a numpy (nz, ny, nx) array or a DeviceArray (gathered first). origin:
    None = 0, a scalar = -size/2 (SAVE_IMG.m's one-value form), or (x, y, z)."""
    if isinstance(volume, DeviceArray):
        volume = volume.get()
    v = np.ascontiguousarray(volume, dtype=np.float32)
    if np.iscomplexobj(volume):
        raise NotImplementedError("complex volumes (modes 3/4) are not written")
    if v.ndim == 2:
        v = v[None]
    if v.ndim != 3:
        raise ValueError("SAVE_IMG takes a 2-D or 3-D volume")
    dims = _lib.i3(v.shape[::-1])
    px = None
    if pixelSize is not None:
        p = np.broadcast_to(np.asarray(pixelSize, dtype=np.float64), (3,))
        px = _lib.d_arr(p)
    org, n_org = None, 0
    if origin is not None:
        o = np.atleast_1d(np.asarray(origin, dtype=np.float64))
        if o.size == 1:
            org, n_org = _lib.d_arr([0.0]), 1
        elif o.size == 3:
            org, n_org = _lib.d_arr(o), 3
        else:
            raise ValueError("origin takes one or three values")
    _lib.call("emc_mrc_write", _path(filename), v.ctypes.data_as(C.POINTER(C.c_float)), dims, px, org, n_org)
