"""The C-ABI library loads and exports every entry point include/chapsim_gpu.h
declares; argument errors are reported before any device work (no GPU
needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from chapsim1ww_amd import native as nat
from chapsim1ww_amd.grid import ChannelGrid

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "chapsim_gpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(chs_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = nat.load()
    declared = _declared()
    assert declared, "no declarations found"
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(nat.exported_symbols())
    assert b"gfx950" in lib.chs_version()


def test_argument_errors_without_device():
    lib = nat.load()
    # n < 3 is rejected (the reference reads outside its arrays below that)
    rc = lib.chs_tdma_batched(1, 1, 1, 1, None, 2, 10, 1, 2, 0, None)
    assert rc == nat.CHS_E_ARG
    assert b"n must be" in lib.chs_last_error()
    # bad FFT length (prime factor 7) is rejected at creation, as FFTFAX STOPs
    g = ChannelGrid(16, 8, 16)
    from chapsim1ww_amd.solver import _desc
    d, keep = _desc(g, nat.CHS_FP64)
    d.ncl1 = 14
    h = C.c_void_p()
    assert lib.chs_create(C.byref(d), None, C.byref(h)) == nat.CHS_E_ARG
    assert b"prime factors" in lib.chs_last_error()
    d.ncl1 = 16
    # FLOWDRVTP 1 or 2 only; NFLOW 1 or 3 (the wall-normal component has no
    # right-hand side on its wall row)
    d.flowdrvtp = 3
    assert lib.chs_create(C.byref(d), None, C.byref(h)) == nat.CHS_E_ARG
    assert b"FLOWDRVTP" in lib.chs_last_error()
    d.flowdrvtp, d.nflow = 2, 2
    assert lib.chs_create(C.byref(d), None, C.byref(h)) == nat.CHS_E_ARG
    assert b"NFLOW" in lib.chs_last_error()
    assert lib.chs_rk3_substep(None, 1, 1e-3) == nat.CHS_E_ARG
    assert lib.chs_props_check(None) == nat.CHS_E_ARG


def test_grid_rejects_bad_fft_lengths():
    with pytest.raises(ValueError):
        ChannelGrid(14, 8, 16)
    with pytest.raises(ValueError):
        ChannelGrid(16, 8, 22)
