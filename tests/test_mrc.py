"""MRC files and subtomogram windows on the CPU: the oracle pinned against
the published MRC2014 / IMOD header layout (fields packed at their byte
offsets with struct, independent of the oracle's sequential reader), closed
forms for SAVE_IMG's statistics, hand-worked BH_isWindowValid cases; then
the library's host-only entry points (header read, (sub)volume read, write)
against the oracle, and BH_multi_loadOrBuild's cache resolution."""
import math
import os
import struct

import numpy as np
import pytest

from emclarity_amd import _lib
from emclarity_amd.mrc import MRCImage, SAVE_IMG, read_header
from emclarity_amd.sizing import BH_isWindowValid
from emclarity_amd.subtomo import BH_multi_loadOrBuild, calc_rg
from oracle import emc_oracle as O


def _lib_ok():
    try:
        _lib.load()
        return True
    except Exception:
        return False


needs_lib = pytest.mark.skipif(not _lib_ok(), reason="libemc_hip.so not built")


def _mrc_bytes(data, mode, bo="<", next_bytes=0, extra_fields=None):
    """An MRC file by the MRC2014 word table: words 1-10 ints at 0..36,
    cell 40, angles 52, mapc/r/s 64, dmin/max/mean 76, ispg 88 (int16 in
    IMOD's layout), nsymbt 90, next 92, creatorid 96, nint 128, nreal 130,
    imodStamp 152, imodFlags 156, origin 196, MAP 208, MACHST 212, RMS 216,
    NLABL 220, labels 224."""
    nz, ny, nx = data.shape
    b = bytearray(1024)
    struct.pack_into(bo + "10i", b, 0, nx, ny, nz, mode, 1, 2, 3, nx, ny, nz)
    struct.pack_into(bo + "6f", b, 40, 10.0 * nx, 10.0 * ny, 10.0 * nz, 90.0, 91.0, 92.0)
    struct.pack_into(bo + "3i", b, 64, 1, 2, 3)
    struct.pack_into(bo + "3f", b, 76, -1.5, 2.5, 0.25)
    struct.pack_into(bo + "hhi", b, 88, 1, 0, next_bytes)
    struct.pack_into(bo + "h", b, 96, 7)
    struct.pack_into(bo + "hh", b, 128, 5, 6)
    struct.pack_into(bo + "ii", b, 152, 1146047817, 1)
    struct.pack_into(bo + "6h", b, 160, 11, 12, 13, 14, 15, 16)
    struct.pack_into(bo + "6f", b, 172, 1, 2, 3, 4, 5, 6)
    struct.pack_into(bo + "3f", b, 196, -3.0, 4.0, 5.5)
    b[208:212] = b"MAP "
    b[212:216] = bytes([68, 68, 0, 0]) if bo == "<" else bytes([17, 17, 0, 0])
    struct.pack_into(bo + "f", b, 216, 0.75)
    struct.pack_into(bo + "i", b, 220, 2)
    b[224:224 + 9] = b"label one"
    b[304:304 + 9] = b"label two"
    t = {0: "i1", 1: "i2", 2: "f4", 6: "u2"}[mode]
    return bytes(b) + bytes(next_bytes) + \
        np.ascontiguousarray(data, dtype=np.dtype(bo + t) if t != "i1" else np.int8).tobytes()


def _data(mode, shape=(5, 6, 7), seed=0):
    rng = np.random.default_rng(seed)
    lo, hi = {0: (-128, 128), 1: (-32768, 32768), 6: (0, 65536)}.get(mode, (0, 1))
    if mode == 2:
        return rng.standard_normal(shape).astype(np.float32)
    d = rng.integers(lo, hi, size=shape)
    d.flat[0], d.flat[1] = lo, hi - 1     # the extremes of the type
    return d


@pytest.mark.parametrize("bo", ["<", ">"])
def test_oracle_header_matches_mrc2014_offsets(tmp_path, bo):
    p = tmp_path / "a.mrc"
    p.write_bytes(_mrc_bytes(_data(1), 1, bo, next_bytes=96))
    h = O.mrc_read_header(p)
    assert (h["nX"], h["nY"], h["nZ"], h["mode"]) == (7, 6, 5, 1)
    assert (h["nXStart"], h["nYStart"], h["nZStart"], h["mX"], h["mY"], h["mZ"]) == (1, 2, 3, 7, 6, 5)
    assert (h["cellDimensionX"], h["cellAngleZ"]) == (70.0, 92.0)
    assert (h["minDensity"], h["maxDensity"], h["meanDensity"]) == (-1.5, 2.5, 0.25)
    assert (h["spaceGroup"], h["nBytesExtended"], h["creatorID"]) == (1, 96, 7)
    assert (h["nBytesPerSection"], h["serialEMType"]) == (5, 6)
    assert (h["imodStamp"], h["imodFlags"]) == (1146047817, 1)
    assert (h["idtype"], h["lens"], h["ndl"], h["nd2"], h["vdl"], h["vd2"]) == (11, 12, 13, 14, 15, 16)
    assert h["tiltAngles"] == [1, 2, 3, 4, 5, 6]
    assert (h["xOrigin"], h["yOrigin"], h["zOrigin"]) == (-3.0, 4.0, 5.5)
    assert h["map"] == b"MAP " and h["densityRMS"] == 0.75 and h["nLabels"] == 2
    assert h["labels"][1].rstrip(b"\0") == b"label two"
    assert h["dataIndex"] == 1024 + 96
    assert h["byteOrder"] == bo
    # the IMOD stamp spells "IMOD" in little-endian byte order
    assert (1146047817).to_bytes(4, "little") == b"IMOD"


@pytest.mark.parametrize("mode", [0, 1, 2, 6])
@pytest.mark.parametrize("bo", ["<", ">"])
def test_oracle_volume_modes_and_ranges(tmp_path, mode, bo):
    d = _data(mode, seed=mode)
    p = tmp_path / "v.mrc"
    p.write_bytes(_mrc_bytes(d, mode, bo))
    v = O.mrc_get_volume(p)
    assert np.array_equal(v, d.astype(np.float64))
    sub = O.mrc_get_volume(p, (2, 6), (3, 3), (1, 4))
    assert np.array_equal(sub, d[0:4, 2:3, 1:6].astype(np.float64))


def test_oracle_save_statistics_closed_forms(tmp_path):
    nz, ny, nx = 4, 5, 6
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    v = (x * (y + 1) * (z + 1)).astype(np.float32)
    p = tmp_path / "s.mrc"
    O.mrc_save(v, p, pixel=2.0, origin=0)
    h = O.mrc_read_header(p)
    # std over x of x*(y+1)*(z+1) = sx*(y+1)*(z+1); over y -> sx*sy'*(z+1) with
    # sy' the std of 1..ny; over z -> sx*sy'*sz', sd of 0..n-1 = sqrt(n(n+1)/12)
    sd = lambda n: math.sqrt(n * (n + 1) / 12.0)
    assert h["densityRMS"] == pytest.approx(sd(nx) * sd(ny) * sd(nz), rel=1e-6)
    assert (h["minDensity"], h["maxDensity"]) == (0.0, float(v.max()))
    assert h["meanDensity"] == pytest.approx(float(v.mean()), rel=1e-6)
    assert (h["cellDimensionX"], h["cellDimensionZ"]) == (12.0, 8.0)
    assert (h["xOrigin"], h["yOrigin"], h["zOrigin"]) == (-3.0, -2.5, -2.0)    # -size/2, unscaled
    assert (h["spaceGroup"], h["nLabels"], h["imodStamp"], h["mode"]) == (1, 1, 1146047817, 2)
    assert h["machineStamp"] == bytes([68, 65, 0, 0]) and h["map"] == b"MAP "
    # a singleton axis ends the nested std at a scalar: rms 0
    O.mrc_save(v[:, :1, :], p)
    assert O.mrc_read_header(p)["densityRMS"] == 0.0
    O.mrc_save(v[0], p)
    assert O.mrc_read_header(p)["densityRMS"] == 0.0
    assert np.array_equal(O.mrc_get_volume(p)[0], v[0])


def test_window_valid_hand_cases():
    # W = 64: low corner 32, top corner 31; centre 100.3 keeps 68..131
    ind, pad, sh = O.window_valid((200, 200, 200), (64, 64, 64), (20, 20, 20), (100.3, 100.0, 150.75))
    assert ind[:, 0].tolist() == [68, 131] and pad[:, 0].tolist() == [0, 0]
    assert ind[:, 2].tolist() == [118, 181]
    assert np.allclose(sh, [0.3, 0.0, 0.75])
    # near the low edge: LOW = 10 - 32 = -22 -> pad 23, kept 1..41
    ind, pad, _ = O.window_valid((200, 200, 200), (64, 64, 64), (10, 10, 10), (10.5, 100, 100))
    assert ind[:, 0].tolist() == [1, 41] and pad[:, 0].tolist() == [23, 0]
    # near the top: TOP = 200 - (190 + 31) = -21 -> pad 21, kept 158..200
    ind, pad, _ = O.window_valid((200, 200, 200), (64, 64, 64), (10, 10, 10), (190.2, 100, 100))
    assert ind[:, 0].tolist() == [158, 200] and pad[:, 0].tolist() == [0, 21]
    # odd window W = 63: both corners 31
    ind, pad, _ = O.window_valid((200, 200, 200), (63, 63, 63), (10, 10, 10), (100, 100, 100))
    assert ind[:, 0].tolist() == [69, 131]
    # too little kept for the mask: 'noUse'
    ind, pad, _ = O.window_valid((200, 200, 200), (64, 64, 64), (20, 20, 20), (3.0, 100, 100))
    assert ind is None


def test_window_valid_host_mirror_matches_oracle():
    rng = np.random.default_rng(4)
    for _ in range(300):
        vs = rng.integers(40, 300, 3)
        ws = rng.integers(16, 90, 3)
        mr = rng.uniform(2, 20, 3)
        c = rng.uniform(-20, vs + 20)
        ind, pad, sh = O.window_valid(vs, ws, mr, c)
        hi, hp, hs = BH_isWindowValid(vs, ws, mr, c)
        assert np.allclose(sh, hs)
        if ind is None:
            assert isinstance(hi, str) and hi == "noUse"
        else:
            assert np.array_equal(ind, hi) and np.array_equal(pad, hp)


def test_oracle_extraction_pins():
    rng = np.random.default_rng(2)
    tomo = rng.standard_normal((60, 70, 80))
    w = O.extract_subtomo(tomo, (32, 32, 32), (5, 5, 5), (40.6, 30.2, 25.9))[0]
    # the window's centre voxel (0-based 16) is the tomogram at floor(centre)
    assert w[16, 16, 16] == tomo[24, 29, 39]
    # the interior (7 voxels from every edge) is an exact copy
    assert np.array_equal(w[7:-7, 7:-7, 7:-7], tomo[24 - 9:24 + 9, 29 - 9:29 + 9, 39 - 9:39 + 9])
    # edge voxels carry the taper 0.5 + 0.5 cos(6 pi / 7) on each edge axis
    t6 = 0.5 + 0.5 * math.cos(6 * math.pi / 7)
    assert w[0, 16, 16] == pytest.approx(tomo[24 - 16, 29, 39] * t6)
    assert w[0, 0, 16] == pytest.approx(tomo[24 - 16, 29 - 16, 39] * t6 * t6)
    # near the edge the pad is zero and the kept block starts tapered
    w, sh = O.extract_subtomo(tomo, (32, 32, 32), (5, 5, 5), (5.0, 30.0, 30.0))
    assert np.all(w[:, :, :12] == 0.0) and w[16, 16, 12] == pytest.approx(tomo[29, 29, 0] * t6)


def test_calc_rg_hand_case():
    # stack nY 928 -> oY 465; slices 10..909 -> nY 900, dY = floor(10 + 450) - 465
    rg = O.calc_rg(928, [1000, 10, 909, 300, 2.5, -3.5])
    assert rg.tolist() == [[1000, 900, 300], [-3, -5, -4]]


@needs_lib
@pytest.mark.parametrize("mode", [0, 1, 2, 6])
@pytest.mark.parametrize("bo", ["<", ">"])
def test_library_reads_match_oracle(tmp_path, mode, bo):
    d = _data(mode, shape=(9, 11, 13), seed=10 + mode)
    p = tmp_path / "v.mrc"
    p.write_bytes(_mrc_bytes(d, mode, bo, next_bytes=64))
    h, r = read_header(p), O.mrc_read_header(p)
    for k in ("nX", "nY", "nZ", "mode", "nXStart", "mZ", "cellDimensionY", "cellAngleY", "minDensity",
              "meanDensity", "spaceGroup", "nBytesExtended", "creatorID", "nBytesPerSection", "serialEMType",
              "imodStamp", "imodFlags", "vd2", "tiltAngles", "zOrigin", "map", "densityRMS", "nLabels",
              "dataIndex"):
        assert h[k] == r[k], k
    assert h["swapped"] == (bo == ">")
    m = MRCImage(p, 0)
    assert np.array_equal(m.getVolume(), O.mrc_get_volume(p))
    assert np.array_equal(m.getVolume([3, 10], 4, [2, 8]), O.mrc_get_volume(p, (3, 10), (4, 4), (2, 8)))
    with pytest.raises(_lib.EmcError):
        m.getVolume([0, 5])


@needs_lib
def test_library_write_matches_oracle(tmp_path):
    rng = np.random.default_rng(3)
    v = rng.standard_normal((7, 9, 10)).astype(np.float32)
    a, b = tmp_path / "a.mrc", tmp_path / "b.mrc"
    SAVE_IMG(v, a, pixelSize=[1.5, 2.0, 2.5], origin=[1, 2, 3])
    O.mrc_save(v, b, pixel=[1.5, 2.0, 2.5], origin=[1, 2, 3])
    ba, bb = a.read_bytes(), b.read_bytes()
    assert len(ba) == len(bb) == 1024 + v.nbytes
    assert ba[1024:] == bb[1024:] and ba[:216] == bb[:216] and ba[220:1024] == bb[220:1024]
    ra, rb = struct.unpack("<f", ba[216:220])[0], struct.unpack("<f", bb[216:220])[0]
    assert ra == pytest.approx(rb, rel=1e-6)
    assert np.array_equal(MRCImage(a).getVolume(), v)
    SAVE_IMG(v[:, :1, :], a, origin=0)
    assert read_header(a)["densityRMS"] == 0.0 and read_header(a)["xOrigin"] == -5.0


@needs_lib
def test_load_or_build_resolves_the_cache(tmp_path):
    os.makedirs(tmp_path / "aliStacks")
    os.makedirs(tmp_path / "cache")
    stack = np.zeros((3, 928, 16), np.float32)
    O.mrc_save(stack, tmp_path / "aliStacks" / "tomoA_ali1.fixed")
    rec = np.arange(4 * 5 * 6, dtype=np.float32).reshape(4, 5, 6)
    O.mrc_save(rec, tmp_path / "cache" / "tomoA_1_bin1.rec")
    rc = [1000, 10, 909, 300, 2.5, -3.5]
    m, rg = BH_multi_loadOrBuild("tomoA_1", rc, 0, 1, 0, 1.0, 0, root=str(tmp_path))
    assert np.array_equal(rg, O.calc_rg(928, rc))
    assert np.array_equal(rg, calc_rg({"nY": 928}, rc))
    assert np.array_equal(m.getVolume(), rec)
    with pytest.raises(NotImplementedError):
        BH_multi_loadOrBuild("tomoB_1", rc, 0, 1, -1, 1.0, 0, root=str(tmp_path))
    with pytest.raises(NotImplementedError):
        BH_multi_loadOrBuild("tomoA_1", rc, 0, 2, 0, 1.0, 0, root=str(tmp_path))
