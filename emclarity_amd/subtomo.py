"""Subtomogram extraction from a tomogram resident in HBM, and the tomogram
cache of the alignment loop.

The reference loads each tomogram once per alignment pass
(BH_multi_loadOrBuild, alignment/BH_alignRaw3d_v2.m:922-933) and, per
particle, checks the window (BH_isWindowValid, :1036-1048), cuts it
(:1074-1081) and pads it with a tapered edge (BH_padZeros3d 'singleTaper',
:1085-1086). Here the tomogram is uploaded once in its stored MRC type and
every particle's window is cut in one launch (emc_extract_subtomos).
"""
from __future__ import annotations

import ctypes as C
import math
import os
from collections import OrderedDict

import numpy as np

from . import _lib
from .device import Context, DeviceArray, default_context
from .mrc import MODE_DTYPES, MRCImage, read_header

_MODE_OF_DTYPE = {np.dtype(v): k for k, v in MODE_DTYPES.items()}


class ResidentTomogram:
    """A tomogram in device memory in its stored type (MRC mode 0 int8,
    1 int16, 2 float32, 6 uint16): int16 maps keep half the bytes of float32
    in HBM and in every window cut."""

    def __init__(self, raw: DeviceArray, mode: int, header: dict | None = None, filename: str | None = None):
        if mode not in MODE_DTYPES or raw.dtype != MODE_DTYPES[mode] or len(raw.shape) != 3:
            raise ValueError("raw must be a (nz, ny, nx) device array of the mode's type")
        self.raw = raw
        self.mode = mode
        self.header = header
        self.filename = filename

    @classmethod
    def from_file(cls, filename, ctx: Context | None = None, kRange=None) -> "ResidentTomogram":
        """Upload sections kRange = (k0, k1) (1-based, inclusive; None = all)
        of an MRC file (emc_mrc_upload)."""
        ctx = ctx or default_context()
        h = read_header(filename)
        if h["mode"] not in MODE_DTYPES:
            raise ValueError(f"{filename}: MRC mode {h['mode']} is not read (0, 1, 2, 6)")
        k0, k1 = (1, h["nZ"]) if kRange is None else (int(kRange[0]), int(kRange[1]))
        raw = DeviceArray((k1 - k0 + 1, h["nY"], h["nX"]), MODE_DTYPES[h["mode"]], ctx)
        _lib.call("emc_mrc_upload", ctx.handle, os.fsencode(os.fspath(filename)), k0, k1, raw.vp)
        return cls(raw, h["mode"], h, os.fspath(filename))

    @classmethod
    def from_numpy(cls, volume: np.ndarray, ctx: Context | None = None) -> "ResidentTomogram":
        v = np.ascontiguousarray(volume)
        if v.dtype not in _MODE_OF_DTYPE:
            v = v.astype(np.float32)
        return cls(DeviceArray.from_numpy(v, ctx=ctx), _MODE_OF_DTYPE[v.dtype])

    @property
    def dims(self):
        """MATLAB size [nX, nY, nZ]."""
        return self.raw.dims

    @property
    def nbytes(self) -> int:
        return self.raw.nbytes

    def to_float(self) -> DeviceArray:
        out = DeviceArray(self.raw.shape, np.float32, self.raw.ctx)
        n = int(np.prod(self.raw.shape, dtype=np.int64))
        _lib.call("emc_mrc_to_float", self.raw.ctx.handle, self.raw.vp, self.mode, C.c_int64(n), out.vp)
        return out

    def free(self):
        self.raw.free()


def extract_subtomos(tomo: ResidentTomogram, centres, window, mask_radius, out: DeviceArray | None = None,
                     taper: bool = True):
    """Cut every particle's window: BH_isWindowValid(tomo size, window,
    mask_radius, centre) + crop + BH_padZeros3d(.., 'singleTaper') (taper
    False: 'single', plain zeros, as transformations/BH_average3d.m:1345).

    centres: (n, 3) (x, y, z) in the reference's 1-based CENTER convention
    (positionList(:, 11:13) ./ samplingRate + binShift). Returns
    (windows (n, wz, wy, wx) float32 DeviceArray, valid (n,) bool — False
    where the reference says 'noUse' — and shiftVAL (n, 3))."""
    c = np.ascontiguousarray(np.asarray(centres, dtype=np.float64).reshape(-1, 3))
    n = c.shape[0]
    w = [int(x) for x in np.broadcast_to(np.asarray(window), (3,))]
    mr = np.broadcast_to(np.asarray(mask_radius, dtype=np.float64), (3,))
    shape = (n, w[2], w[1], w[0])
    if out is None:
        out = DeviceArray(shape, np.float32, tomo.raw.ctx)
    elif out.shape != shape or out.dtype != np.float32:
        raise ValueError(f"out must be float32 {shape}")
    valid = np.zeros(n, np.int32)
    shifts = np.zeros((n, 3), np.float64)
    pd = C.POINTER(C.c_double)
    _lib.call("emc_extract_subtomos", tomo.raw.ctx.handle, tomo.raw.vp, tomo.mode, _lib.i3(tomo.dims), n,
              c.ctypes.data_as(pd), _lib.i3(w), _lib.d_arr(mr), int(bool(taper)), out.vp,
              valid.ctypes.data_as(C.POINTER(C.c_int32)), shifts.ctypes.data_as(pd))
    return out, valid.astype(bool), shifts


class TomogramCache:
    """Tomograms kept resident between alignment passes, least recently used
    evicted first beyond ``budget_bytes`` (the on-disk cache/*.rec of
    BH_multi_loadOrBuild, held in HBM instead). Keyed by path, size and
    modification time, so a rewritten file is uploaded again."""

    def __init__(self, budget_bytes: int = 64 << 30, ctx: Context | None = None):
        self.budget = int(budget_bytes)
        self.ctx = ctx
        self._items: OrderedDict = OrderedDict()

    def _key(self, filename):
        st = os.stat(filename)
        return (os.path.abspath(os.fspath(filename)), st.st_size, st.st_mtime_ns)

    def get(self, filename) -> ResidentTomogram:
        key = self._key(filename)
        t = self._items.get(key)
        if t is not None:
            self._items.move_to_end(key)
            return t
        t = ResidentTomogram.from_file(filename, self.ctx)
        self._items[key] = t
        while sum(v.nbytes for v in self._items.values()) > self.budget and len(self._items) > 1:
            _, old = self._items.popitem(last=False)
            old.free()
        return t

    def __len__(self):
        return len(self._items)

    def clear(self):
        for t in self._items.values():
            t.free()
        self._items.clear()


def calc_rg(header: dict, rCoords) -> np.ndarray:
    """reconGeom of BH_multi_loadOrBuild.m:203-210 (local calc_rg): 2 x 3
    [[width, nY, thickness], round([-shiftX, dY, shiftZ])]."""
    r = np.asarray(rCoords, dtype=np.float64)
    oY = math.ceil((header["nY"] + 1) / 2)
    nY = r[2] - r[1] + 1
    dY = math.floor(r[1] + nY / 2) - oY
    v = np.array([-r[4], dY, r[5]])
    rounded = np.where(v >= 0, np.floor(v + 0.5), np.ceil(v - 0.5))
    return np.array([[r[0], nY, r[3]], rounded])


def BH_multi_loadOrBuild(tomoName: str, rCoords, mapBackIter: int, SAMPLING: int, gpuIDX: int,
                         reconScaling: float, flgLoad: int, recon: str | None = None, root=".",
                         cache: TomogramCache | None = None):
    """transformations/BH_multi_loadOrBuild.m:1-212 for a reconstruction that
    is already cached: resolves cache/<tomo>_<n>_bin<S>.rec, the reconGeom of
    the aligned stack's header, and loads the tomogram. flgLoad 0 returns the
    header-only MRCImage (:184-186); flgLoad 1 returns the tomogram resident
    in device memory (through ``cache`` when given) where the reference
    returns single(getVolume(m)) (:182-183). Reconstructing with IMOD's tilt
    (gpuIDX < 0 and no cached file) and binning a missing stack are outside
    this library and raise NotImplementedError."""
    do_recon = gpuIDX < 0
    parts = tomoName.split("_")
    base = "_".join(parts[:-1])
    number = int(float(parts[-1]))
    r = np.asarray(rCoords, dtype=np.float64) / SAMPLING
    r[:4] = np.fix(r[:4])
    rec = recon or f"cache/{base}_{number}_bin{SAMPLING}.rec"
    if SAMPLING > 1:
        stack = f"cache/{base}_ali{mapBackIter + 1}_bin{SAMPLING}.fixed"
    else:
        stack = f"aliStacks/{base}_ali{mapBackIter + 1}.fixed"
    rec_p, stack_p = os.path.join(root, rec), os.path.join(root, stack)
    if not os.path.exists(stack_p):
        raise NotImplementedError(f"{stack} is missing: binning or aligning stacks is not part of this library")
    if not os.path.exists(rec_p):
        if do_recon:
            raise NotImplementedError("reconstruction (IMOD tilt) is not part of this library")
        raise FileNotFoundError(f"An appropriate reconstruction was not found for {rec}")
    recon_geom = calc_rg(read_header(stack_p), r)
    if not flgLoad:
        return MRCImage(rec_p, 0), recon_geom
    tomo = cache.get(rec_p) if cache is not None else ResidentTomogram.from_file(rec_p)
    return tomo, recon_geom
