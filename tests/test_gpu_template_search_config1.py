"""Template search at the workload's own geometry: a 512 x 512 x 256
tomogram, a 48^3 template, the reference chunker's 480^3 chunks with [2, 2,
1] iterations (BH_templateSearch3d_2.m:339-351, :650-982), orientations of
the 15-degree grid ([180, 15, 180, 15], 4,600 orientations).

* One 480^3 chunk against the FP64 oracle: the per-voxel max CC to 1e-4 of
  the map's maximum; the argmax orientation id exact at every voxel whose
  oracle top-2 margin exceeds twice that tolerance (a score error below the
  tolerance cannot reorder such a voxel's best two orientations).
* The chunked driver over the whole tomogram: chunk 0's valid area against
  the same oracle result, and every inserted template found at its centre
  with its own orientation id (the map index is the particle centre).
* Templates with edges above 32 and non-cubic (48^3, 40 x 33 x 64): the
  pruned template spectrum (vector and matrix-core z stage) against the
  oracle and against the padded volume's R2C."""
import numpy as np
import pytest

import tsearch_config as TC
from emclarity_amd import _lib
from emclarity_amd.device import DeviceArray
from emclarity_amd.geometry import eulerSearch
from emclarity_amd.sizing import BH_multi_padVal
from emclarity_amd.template_search import symmetric_pad, template_search
from oracle import emc_oracle as O

pytestmark = pytest.mark.gpu

ROWS = [0, 1, 2, 250, 2300, 4599]   # rows of the 4,600-orientation grid
RTOL = 1e-4


def _check(peak, ang, rpeak, rang, margin):
    scale = float(np.max(np.abs(rpeak)))
    err = float(np.max(np.abs(peak - rpeak)))
    assert err < RTOL * scale, (err, scale)
    sure = margin > 2 * RTOL * scale
    bad = int(np.count_nonzero((ang != rang) & sure))
    assert bad == 0, (bad, int(np.count_nonzero(sure)))
    return err


@pytest.fixture(scope="module")
def config1():
    pts = TC.grid_15()[ROWS]
    tomo, t, placed = TC.tomogram(7, pts)
    r = TC.chunking()
    assert list(r["chunkSize"]) == [480] * 3 and list(r["nIters"]) == [2, 2, 1]
    padded = symmetric_pad(tomo, r["tomoPre"], r["tomoPost"])
    sc = [int(v) for v in r["chunkSize"]]
    pad, vlow, trim = TC.chunk_offsets(sc, r["validAreaKeep"])
    chunk = np.ascontiguousarray(padded[0:sc[2], 0:sc[1], 0:sc[0]])
    del padded
    rpeak, rang, margin = O.template_search_chunk(chunk.astype(np.float64), t.astype(np.float64), pts,
                                                  pad[0:2], trim[0:2], return_margin=True)
    return dict(pts=pts, tomo=tomo, t=t, placed=placed, r=r, chunk=chunk, pad=pad, vlow=vlow,
                rpeak=rpeak, rang=rang, margin=margin)


def test_config1_chunk_matches_oracle(ctx, config1):
    c = config1
    sc = [int(v) for v in c["r"]["chunkSize"]]
    va = [int(v) for v in c["r"]["validAreaKeep"]]
    d_chunk = DeviceArray.from_numpy(c["chunk"])
    d_t = DeviceArray.from_numpy(c["t"])
    peak = DeviceArray.zeros((va[2], va[1], va[0]), np.float32)
    ang = DeviceArray.zeros((va[2], va[1], va[0]), np.int32)
    pts = np.ascontiguousarray(c["pts"])
    _lib.call("emc_template_search_chunk", ctx.handle, d_chunk.vp, _lib.i3(sc), d_t.vp, _lib.i3([48] * 3),
              _lib.i3(c["pad"][0]), _lib.i3(c["vlow"]), _lib.i3(va), pts.ctypes.data_as(_lib.C.POINTER(_lib.C.c_double)),
              len(pts), 1, 1, peak.vp, ang.vp)
    err = _check(peak.get(), ang.get(), c["rpeak"], c["rang"], c["margin"])
    print(f"config-1 chunk: max |score - oracle| {err:.3e} (map max {np.max(c['rpeak']):.2f})")


def test_config1_driver_recovers_every_template(ctx, config1):
    c = config1
    r = c["r"]
    peak, ang = template_search(c["tomo"], c["t"], c["pts"], r["chunkSize"], r["validAreaKeep"], r["nIters"],
                                r["tomoPre"], r["tomoPost"], ctx=ctx)
    assert peak.shape == c["tomo"].shape
    # chunk 0's valid area, in tomogram coordinates: [0, 276 - 102 + ...)
    pre = [int(v) for v in r["tomoPre"]]
    va = [int(v) for v in r["validAreaKeep"]]
    lo = [int(c["vlow"][k]) - pre[k] for k in range(3)]          # 0
    hi = [lo[k] + va[k] for k in range(3)]
    sub_p = peak[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
    sub_a = ang[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
    o = [pre[k] - int(c["vlow"][k]) for k in range(3)]           # oracle offset of tomogram voxel 0
    _check(sub_p, sub_a, c["rpeak"][o[2]:, o[1]:, o[0]:][:sub_p.shape[0], :sub_p.shape[1], :sub_p.shape[2]],
           c["rang"][o[2]:, o[1]:, o[0]:][:sub_p.shape[0], :sub_p.shape[1], :sub_p.shape[2]],
           c["margin"][o[2]:, o[1]:, o[0]:][:sub_p.shape[0], :sub_p.shape[1], :sub_p.shape[2]])
    # the score is normalised per chunk (std over its valid area), so the
    # local-maximum check stays inside the centre's own chunk
    missed = []
    for (x, y, z, k) in c["placed"]:
        lo3 = [(q // va[a]) * va[a] for a, q in enumerate((x, y, z))]
        hi3 = [lo3[a] + va[a] for a in range(3)]
        loc = peak[max(z - 2, lo3[2]):min(z + 3, hi3[2]), max(y - 2, lo3[1]):min(y + 3, hi3[1]),
                   max(x - 2, lo3[0]):min(x + 3, hi3[0])]
        if not (ang[z, y, x] == k + 1 and peak[z, y, x] == loc.max() and peak[z, y, x] > 8.0):
            missed.append((int(x), int(y), int(z), int(k), float(peak[z, y, x]), int(ang[z, y, x])))
    assert not missed, missed
    print(f"config-1 driver: {len(c['placed'])} inserted templates found at their centres with their ids")


def _blob_template(rng, dims):
    """(z, y, x) array of Gaussian blobs inside a box of edges dims (x, y, z)."""
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz) - nz // 2, np.arange(ny) - ny // 2, np.arange(nx) - nx // 2, indexing="ij")
    v = np.zeros((nz, ny, nx))
    for _ in range(12):
        c = rng.uniform(-0.3, 0.3, 3) * np.array(dims)
        s = rng.uniform(1.5, 3.0)
        v += rng.uniform(0.5, 1.0) * np.exp(-((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) / (2 * s * s))
    return v.astype(np.float32)


@pytest.mark.parametrize("tdims,cdims", [((48, 48, 48), (128, 128, 128)), ((40, 33, 64), (120, 112, 144))])
def test_large_templates_pruned_padded_and_oracle(ctx, tdims, cdims):
    rng = np.random.default_rng(sum(tdims))
    t = _blob_template(rng, tdims)
    grid = eulerSearch("C1", 180, 15, 180, 15)
    pts = np.ascontiguousarray(grid[[0, 5, 30, 400, 1777, 3000, 4599]])
    chunk = rng.standard_normal((cdims[2], cdims[1], cdims[0])).astype(np.float32)
    for k, cen in ((1, (44, 50, 60)), (4, (70, 60, 80))):
        p, th, ps = pts[k]
        rot = O.interp3d(t.astype(np.float64), [p, th, ps - p], [1, 1, 1], "Bah", "forward")
        sl = tuple(slice(cen[2 - a] - tdims[2 - a] // 2, cen[2 - a] - tdims[2 - a] // 2 + tdims[2 - a])
                   for a in range(3))
        chunk[sl] -= (1.5 * rot / np.std(t)).astype(np.float32)
    vd = tuple(int(c - t_ - 8) for c, t_ in zip(cdims, tdims))
    pad = BH_multi_padVal(tdims, cdims)
    vA = BH_multi_padVal(vd, cdims)
    trim = BH_multi_padVal(cdims, vd)
    d_chunk = DeviceArray.from_numpy(chunk)
    d_t = DeviceArray.from_numpy(t)
    dp = _lib.C.POINTER(_lib.C.c_double)
    out = []
    for pruned in (2, 1, 0):
        _lib.call("emc_tune_set", ctx.handle, b"tsearch_pruned", 0, pruned)
        peak = DeviceArray.zeros((vd[2], vd[1], vd[0]), np.float32)
        ang = DeviceArray.zeros((vd[2], vd[1], vd[0]), np.int32)
        _lib.call("emc_template_search_chunk", ctx.handle, d_chunk.vp, _lib.i3(cdims), d_t.vp, _lib.i3(tdims),
                  _lib.i3(pad[0]), _lib.i3(vA[0]), _lib.i3(vd), pts.ctypes.data_as(dp), len(pts), 1, 1,
                  peak.vp, ang.vp)
        out.append((peak.get(), ang.get()))
    _lib.call("emc_tune_set", ctx.handle, b"tsearch_pruned", 0, 2)   # the library default
    rpeak, rang, margin = O.template_search_chunk(chunk.astype(np.float64), t.astype(np.float64), pts, pad[0:2],
                                                  trim[0:2], return_margin=True)
    for p, a in out:
        _check(p, a, rpeak, rang, margin)
