"""ctypes binding of libchapsim_gpu.so (the C ABI of include/chapsim_gpu.h).

The library is the only compute path: nothing here falls back to the CPU.
Loading fails loudly when the shared library is missing; device work needs a
visible HIP device.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libchapsim_gpu.so")

CHS_OK, CHS_E_ARG, CHS_E_HIP, CHS_E_NOMEM, CHS_E_RANGE, CHS_E_NOTCONV = range(6)
CHS_FP64, CHS_MIXED = 0, 1
CHS_PROPS_TABLE, CHS_PROPS_SURROGATE = 0, 1

_dp = C.POINTER(C.c_double)


class ChsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"chapsim_gpu error {code}: {msg}")
        self.code = code


class GridDesc(C.Structure):
    _fields_ = [("ncl1", C.c_int), ("ncl2", C.c_int), ("ncl3", C.c_int), ("icase", C.c_int),
                ("flowdrvtp", C.c_int), ("precision", C.c_int),
                ("dxi", C.c_double), ("dzi", C.c_double), ("dxqi", C.c_double), ("dzqi", C.c_double),
                ("cvisc", C.c_double), ("volm", C.c_double)] + \
               [(n, _dp) for n in ("dyfi", "dyci", "wff", "wfb", "amph", "acph", "apph",
                                   "amvr", "acvr", "apvr", "ak1", "ak3")] + \
               [("nflow", C.c_int), ("cfgv", C.c_double)]


class ThermalFields(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("rhoh", "enthalpy", "temperature", "density", "density0", "density1",
                                          "viscosity", "viscosity0", "thermcondt", "rhs_energy0", "g", "g0",
                                          "wallflux")]


class IoFields(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("q", "g", "g0", "pr", "drhodtp", "divu", "rho_stg", "mu_stg", "explt0",
                                          "drhoi_stg")]


class IoParams(C.Structure):
    _fields_ = [("thermal", C.c_int), ("nflow", C.c_int), ("flowdrvtp", C.c_int), ("gravdir", C.c_int),
                ("cfgv", C.c_double), ("f_a", C.c_double), ("visimplicit", C.c_int)]


# every export of include/chapsim_gpu.h with its ctypes signature
_SIGS = {
    "chs_create": ([C.POINTER(GridDesc), C.c_void_p, C.POINTER(C.c_void_p)], C.c_int),
    "chs_destroy": ([C.c_void_p], C.c_int),
    "chs_upload_state": ([C.c_void_p, _dp, _dp, _dp], C.c_int),
    "chs_download_state": ([C.c_void_p, _dp, _dp, _dp], C.c_int),
    "chs_device_arrays": ([C.c_void_p] + [C.POINTER(C.c_void_p)] * 5, C.c_int),
    "chs_rk3_substep": ([C.c_void_p, C.c_int, C.c_double], C.c_int),
    "chs_poisson_fft99": ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "chs_set_refinement": ([C.c_void_p, C.c_double, C.c_int], C.c_int),
    "chs_poisson_stats": ([C.c_void_p, C.POINTER(C.c_int), _dp], C.c_int),
    "chs_set_cvisc": ([C.c_void_p, C.c_double], C.c_int),
    "chs_set_cn_solver": ([C.c_void_p, C.c_int], C.c_int),
    "chs_cn_info": ([C.c_void_p, C.c_int, C.c_double, C.POINTER(C.c_int), C.POINTER(C.c_int),
                     C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
    "chs_divgck": ([C.c_void_p, _dp], C.c_int),
    "chs_cfl": ([C.c_void_p, _dp], C.c_int),
    "chs_last_dpgrns": ([C.c_void_p, _dp], C.c_int),
    "chs_mean_xz": ([C.c_void_p, _dp], C.c_int),
    "chs_tdma_batched": ([C.c_void_p] * 5 + [C.c_int, C.c_int, C.c_longlong, C.c_longlong, C.c_int, C.c_void_p],
                         C.c_int),
    "chs_tdma_batched_noalloc": ([C.c_void_p] * 5 + [C.c_int, C.c_int, C.c_longlong, C.c_longlong, C.c_int,
                                                     C.c_void_p], C.c_int),
    "chs_tdma_batched_host": ([_dp] * 5 + [C.c_int, C.c_int, C.c_longlong, C.c_longlong, C.c_int], C.c_int),
    "chs_props_create": ([C.c_int] + [_dp] * 6 + [C.c_void_p, C.POINTER(C.c_void_p)], C.c_int),
    "chs_props_destroy": ([C.c_void_p], C.c_int),
    "chs_therm_prop_update": ([C.c_void_p, C.c_int, C.c_void_p, C.c_longlong, C.c_longlong, C.c_longlong]
                              + [C.c_void_p] * 5 + [C.POINTER(C.c_longlong)], C.c_int),
    "chs_props_rhoh_max": ([C.c_void_p, _dp], C.c_int),
    "chs_energy_substep": ([C.c_void_p, C.c_void_p, C.POINTER(ThermalFields), C.c_int, C.c_double, C.c_double],
                           C.c_int),
    "chs_io_prepare": ([C.c_void_p, C.POINTER(ThermalFields), C.POINTER(IoFields), C.POINTER(IoParams), C.c_int],
                       C.c_int),
    "chs_solverrk3_mom_io": ([C.c_void_p, C.POINTER(ThermalFields), C.POINTER(IoFields), C.POINTER(IoParams),
                              C.c_int, C.c_double], C.c_int),
    "chs_divgck_io": ([C.c_void_p, C.POINTER(IoFields), _dp, _dp], C.c_int),
    "chs_surrogate_train": ([C.c_void_p, C.c_void_p, C.c_longlong, C.c_int, C.c_int, C.c_double, C.c_ulonglong,
                             _dp], C.c_int),
    "chs_surrogate_certify": ([C.c_void_p, C.c_double, C.c_int, _dp], C.c_int),
    "chs_surrogate_grad": ([C.c_void_p, C.c_void_p, C.c_longlong, C.POINTER(C.c_float), _dp], C.c_int),
    "chs_surrogate_eval": ([C.c_void_p, C.c_void_p, C.c_longlong, C.c_void_p], C.c_int),
    "chs_surrogate_params": ([C.c_void_p, C.c_int, C.POINTER(C.c_float), C.c_int, _dp], C.c_int),
    "chs_profile_enable": ([C.c_void_p, C.c_int], C.c_int),
    "chs_profile_read": ([C.c_void_p, C.c_int, C.c_char_p, _dp, C.POINTER(C.c_long), C.POINTER(C.c_int)],
                         C.c_int),
    "chs_profile_reset": ([C.c_void_p], C.c_int),
    "chs_last_error": ([], C.c_char_p),
    "chs_surrogate_certified": ([C.c_void_p, _dp, _dp], C.c_int),
    "chs_props_set_mode": ([C.c_void_p, C.c_int], C.c_int),
    "chs_props_last_fallback": ([C.c_void_p, C.POINTER(C.c_longlong)], C.c_int),
    "chs_props_check": ([C.c_void_p], C.c_int),
    "chs_mean_xz_io": ([C.c_void_p, C.POINTER(ThermalFields), C.POINTER(IoFields), C.c_int, _dp], C.c_int),
    "chs_mean_t": ([C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), _dp], C.c_int),
    "chs_mean_t_reset": ([C.c_void_p, C.c_int], C.c_int),
    "chs_version": ([], C.c_char_p),
}

_lib = None


def load() -> C.CDLL:
    """Load the in-tree HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m chapsim1ww_amd.build`")
        lib = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
    return _lib


def exported_symbols():
    return sorted(_SIGS)


def check(code: int) -> None:
    if code != CHS_OK:
        raise ChsError(code, load().chs_last_error().decode())


def dptr(x) -> int:
    """Device address of a torch CUDA tensor (or an int address)."""
    if isinstance(x, int) or x is None:
        return x
    if not x.is_cuda:
        raise ValueError("expected a device tensor")
    if not x.is_contiguous():
        raise ValueError("expected a contiguous device tensor")
    return x.data_ptr()


def hptr(a: np.ndarray):
    if a is None:
        return None
    if a.dtype != np.float64 or not (a.flags.f_contiguous or a.flags.c_contiguous):
        raise ValueError("expected a contiguous float64 array")
    return a.ctypes.data_as(_dp)
