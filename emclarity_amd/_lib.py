"""ctypes binding of libemc_hip.so (the C ABI declared in include/emc_hip.h).

The library must exist: there is no CPU fallback. ``load()`` raises
``RuntimeError`` with the build command when it is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libemc_hip.so"
HEADER_PATH = Path(__file__).resolve().parent.parent / "include" / "emc_hip.h"

EMC_OK = 0
STATUS_NAMES = {0: "EMC_OK", 1: "EMC_ERR_INVALID", 2: "EMC_ERR_HIP", 3: "EMC_ERR_FFT",
                4: "EMC_ERR_ALLOC", 5: "EMC_ERR_RANGE", 6: "EMC_ERR_NODEVICE", 7: "EMC_ERR_IO"}

_st = C.c_int
_vp = C.c_void_p
_i3 = C.POINTER(C.c_int)
_d = C.POINTER(C.c_double)
_f = C.POINTER(C.c_float)

# name -> (restype, argtypes); every entry point of include/emc_hip.h
SIGNATURES = {
    "emc_last_error": (C.c_char_p, []),
    "emc_version": (C.c_char_p, []),
    "emc_ctx_create": (_st, [C.c_int, C.POINTER(_vp)]),
    "emc_ctx_destroy": (_st, [_vp]),
    "emc_ctx_synchronize": (_st, [_vp]),
    "emc_ctx_wait": (_st, [_vp, _vp]),
    "emc_ctx_stream": (_vp, [_vp]),
    "emc_timing_enable": (_st, [_vp, C.c_int]),
    "emc_timing_reset": (_st, [_vp]),
    "emc_timing_read": (_st, [_vp, C.c_char_p, _d, C.POINTER(C.c_int64)]),
    "emc_tune_set": (_st, [_vp, C.c_char_p, C.c_int, C.c_int]),
    "emc_tune_get": (_st, [_vp, C.c_char_p, C.c_int, C.POINTER(C.c_int)]),
    "emc_malloc": (_st, [_vp, C.c_size_t, C.POINTER(_vp)]),
    "emc_free": (_st, [_vp, _vp]),
    "emc_memcpy_h2d": (_st, [_vp, _vp, _vp, C.c_size_t]),
    "emc_memcpy_h2d_async": (_st, [_vp, _vp, _vp, C.c_size_t]),
    "emc_memcpy_d2h": (_st, [_vp, _vp, _vp, C.c_size_t]),
    "emc_memcpy_d2d": (_st, [_vp, _vp, _vp, C.c_size_t]),
    "emc_memset_zero": (_st, [_vp, _vp, C.c_size_t]),
    "emc_define_matrix": (_st, [_d, C.c_int, C.c_int, _d]),
    "emc_euler_search": (_st, [C.c_char_p, C.c_double, C.c_double, C.c_double, C.c_double, _d,
                               C.c_int64, C.POINTER(C.c_int64)]),
    "emc_xform3d": (_st, [_vp, _vp, _i3, _f, _f, C.c_int, _vp]),
    "emc_interp3d": (_st, [_vp, _vp, _i3, _d, C.c_int, _d, _f, C.c_int, _vp]),
    "emc_sf3d": (_st, [_vp, C.c_int, C.c_int, _i3, C.c_int, _f, _f, _f, _f, _f, _f, _f, _f, _f, _f,
                       _vp]),
    "emc_weight_mask": (_st, [_vp, _i3, C.c_double, _d, C.c_int, _d, _d, _d, _vp]),
    "emc_weight_mask_batched": (_st, [_vp, _i3, C.c_double, C.c_int, _d, C.POINTER(C.c_int32), _d, _d, _d,
                                      _vp]),
    "emc_ctf2d": (_st, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                        C.c_float, C.c_float, C.c_float, C.c_float, C.c_int, _vp]),
    "emc_fft_plan_create": (_st, [_vp, _i3, C.c_int, C.POINTER(_vp)]),
    "emc_fft_plan_destroy": (_st, [_vp]),
    "emc_fft_r2c": (_st, [_vp, _vp, _vp]),
    "emc_fft_c2r": (_st, [_vp, _vp, _vp]),
    "emc_fft_c2c": (_st, [_vp, _vp, _vp, C.c_int]),
    "emc_bandpass3d": (_st, [_vp, _i3, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, _vp]),
    "emc_pad_zeros3d": (_st, [_vp, _vp, _i3, _i3, _i3, C.c_float, C.c_int, _vp]),
    "emc_ft_condition": (_st, [_vp, _vp, _i3, C.c_int, C.c_int, _vp, _vp]),
    "emc_band_limit_center_normalize": (_st, [_vp, _vp, _vp, _i3, _i3, _i3, _vp, _vp]),
    "emc_ref_ft1": (_st, [_vp, _vp, _vp, _vp, _i3, _i3, _i3, _vp]),
    "emc_xcf_translational": (_st, [_vp, _vp, _vp, _vp, _i3, _vp, _i3, _vp, _d, _d]),
    "emc_shell_plan_create": (_st, [_vp, C.c_int, C.c_double, C.c_int, C.POINTER(C.c_int64),
                                    C.POINTER(C.c_uint32), C.POINTER(_vp)]),
    "emc_shell_plan_destroy": (_st, [_vp, _vp]),
    "emc_xcf_rotational_shells": (_st, [_vp, _vp, _vp, _vp, _vp, _vp, _i3, _d]),
    "emc_xcf_rotational": (_st, [_vp, _vp, _vp, _vp, _vp, _i3, _d]),
    "emc_align_grid": (_st, [_vp, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _d, _i3,
                             _d, C.c_int, C.c_int, _d, _d]),
    "emc_align_grid_padded": (_st, [_vp, C.c_int, C.c_int, _i3, C.c_int, _d, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp,
                                    _vp, _vp, _vp, _vp, _d, _d, _i3, _d, C.c_int, C.c_int, _d, _d]),
    "emc_template_search_chunk": (_st, [_vp, _vp, _i3, _vp, _i3, _i3, _i3, _i3, _d, C.c_int, C.c_int,
                                        C.c_int, _vp, _vp]),
    "emc_align_pick": (_st, [_vp, C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "emc_align_refine": (_st, [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, C.c_int,
                               C.c_double, C.c_double, C.c_double, _vp]),
    "emc_mask3d_sphere": (_st, [_vp, _i3, _d, _vp]),
    "emc_mask_shape": (_st, [_vp, C.c_int, _i3, _d, _d, C.c_double, _vp]),
    "emc_tsearch_peaks": (_st, [_vp, _vp, _vp, _i3, _vp, C.c_int, _d, C.c_int, C.c_int, C.c_double, C.c_double,
                                C.c_int, C.c_int, _d, C.POINTER(C.c_int)]),
    "emc_tsearch_peaks_multi": (_st, [_vp, _vp, _vp, _i3, _vp, C.c_int, _d, C.c_int, C.c_int, C.c_double,
                                      C.c_double, C.c_int, C.c_int, C.c_int, C.c_double, _d, _d,
                                      C.POINTER(C.c_int)]),
    "emc_moving_average": (_st, [_vp, _vp, _i3, C.c_int, _vp]),
    "emc_local_normalize": (_st, [_vp, _vp, _i3, C.c_int, _vp]),
    "emc_tsearch_preprocess_chunk": (_st, [_vp, _vp, _i3, _d, C.c_int, C.c_int, C.c_float, _vp, _vp, _d]),
    "emc_median_filter3d": (_st, [_vp, _vp, _i3, C.c_int, _vp]),
    "emc_scale": (_st, [_vp, _vp, C.c_int64, C.c_double]),
    "emc_mrc_read_header": (_st, [C.c_char_p, _vp]),
    "emc_mrc_read_volume": (_st, [C.c_char_p, _i3, _f]),
    "emc_mrc_upload": (_st, [_vp, C.c_char_p, C.c_int, C.c_int, _vp]),
    "emc_mrc_to_float": (_st, [_vp, _vp, C.c_int, C.c_int64, _vp]),
    "emc_mrc_write": (_st, [C.c_char_p, _f, _i3, _d, _d, C.c_int]),
    "emc_average_accumulate": (_st, [_vp, C.c_int, _i3, _vp, _d, _d, C.POINTER(C.c_int32), _d, C.c_int, _d, _vp,
                                     _i3, _vp, _vp, _vp, _vp, _vp, C.POINTER(C.c_int32)]),
    "emc_extract_subtomos": (_st, [_vp, _vp, C.c_int, _i3, C.c_int, _d, _i3, _d, C.c_int, _vp,
                                   C.POINTER(C.c_int32), _d]),
    "emc_batch_reference_spectrum": (_st, [_vp, C.c_int, _vp, C.c_int, _vp]),
    "emc_batch_prepare_particle": (_st, [_vp, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "emc_batch_prepare_particles": (_st, [_vp, C.c_int, C.c_int, C.POINTER(C.c_void_p), _vp,
                                          C.POINTER(C.c_void_p), _vp, _vp, _vp]),
    "emc_batch_prepare_particles_masked": (_st, [_vp, C.c_int, C.c_int, C.POINTER(C.c_void_p), _vp, _vp,
                                                 C.POINTER(C.c_void_p), _vp, _vp, _vp]),
    "emc_align_batched": (_st, [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp, _d, C.c_int, C.c_int,
                                C.c_double, _vp, _vp]),
    "emc_align_eval_grad": (_st, [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, C.c_double,
                                  _vp]),
    "emc_align_eval_grad_noalloc": (_st, [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp,
                                          C.c_double, _vp]),
}

_LIB = None


class EmcError(RuntimeError):
    """A failing emc_* call: status code and the library's message."""

    def __init__(self, func, status, message):
        super().__init__(f"{func} -> {STATUS_NAMES.get(status, status)}: {message}")
        self.status = status
        self.func = func


def load(path: os.PathLike | None = None):
    """Load the library once and attach the signatures."""
    global _LIB
    if _LIB is not None:
        return _LIB
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: build it with `python -m emclarity_amd.csrc.build` "
            "(there is no CPU fallback)")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def check(func: str, status: int):
    if status != EMC_OK:
        msg = load().emc_last_error()
        raise EmcError(func, status, msg.decode() if msg else "")


def call(func: str, *args):
    lib = load()
    check(func, getattr(lib, func)(*args))


def i3(v):
    return (C.c_int * 3)(*[int(x) for x in v])


def d_arr(v):
    v = [float(x) for x in v]
    return (C.c_double * len(v))(*v)


def f_arr(v):
    v = [float(x) for x in v]
    return (C.c_float * len(v))(*v)
