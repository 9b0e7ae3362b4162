"""Host-side mirror of the reference's tg-channel solver routines, driving the
device library through its C ABI.

    SOLVERRK3_MOM_tg(NS)        -> ChannelSolver.solverrk3_mom_tg(ns, dt)
    FFT99_POIS3D_periodicxz     -> ChannelSolver.fft99_pois3d(f, dph)
    DIVGCK_tg / CFL_tg          -> ChannelSolver.divgck() / .cfl()
    TDMAIJJ/TDMAIJI (non)CYC    -> tdmaijj_noncyc, tdmaiji_noncyc, tdmaijj_cyc, tdmaiji_cyc

Arrays on the host side are numpy arrays in the reference's Fortran layouts
(see ChannelGrid.shape_*); device arrays are torch tensors on the HIP device
(torch provides the allocations and the stream; all arithmetic is in the
library).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import native as nat
from .grid import ChannelGrid


def _desc(g: ChannelGrid, precision: int, flowdrvtp: int = 1, nflow: int = 1, cfgv: float = 0.0):
    keep = {k: np.ascontiguousarray(getattr(g, k), dtype=np.float64)
            for k in ("dyfi", "dyci", "wff", "wfb", "amph", "acph", "apph", "amvr", "acvr", "apvr", "ak1", "ak3")}
    d = nat.GridDesc()
    d.ncl1, d.ncl2, d.ncl3 = g.L, g.M, g.N
    d.icase, d.flowdrvtp, d.precision = 1, int(flowdrvtp), precision
    d.nflow, d.cfgv = int(nflow), float(cfgv)
    d.dxi, d.dzi, d.dxqi, d.dzqi, d.cvisc, d.volm = g.dxi, g.dzi, g.dxqi, g.dzqi, g.cvisc, g.volm
    for k, v in keep.items():
        setattr(d, k, nat.hptr(v))
    return d, keep


class ChannelSolver:
    """Device-resident tg channel state plus the per-substep solver core."""

    def __init__(self, grid: ChannelGrid, precision: str = "fp64", stream=None, flowdrvtp: int = 1, nflow: int = 1,
                 cfgv: float = 0.0):
        """flowdrvtp: FLOWDRVTP, 1 constant mass flux or 2 constant pressure
        gradient -cfgv/2; nflow: NFLOW, the streamwise direction 1 (x) or 3 (z)."""
        self.grid = grid
        self.lib = nat.load()
        self.precision = {"fp64": nat.CHS_FP64, "mixed": nat.CHS_MIXED}[precision]
        d, self._keep = _desc(grid, self.precision, flowdrvtp, nflow, cfgv)
        h = C.c_void_p()
        s = None if stream is None else C.c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream))
        nat.check(self.lib.chs_create(C.byref(d), s, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.chs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state ----------------------------------------------------------
    def upload_state(self, q, pr, convh0):
        g = self.grid
        q = None if q is None else np.asfortranarray(q, dtype=np.float64)
        pr = None if pr is None else np.asfortranarray(pr, dtype=np.float64)
        convh0 = None if convh0 is None else np.asfortranarray(convh0, dtype=np.float64)
        for a, shp in ((q, g.shape_q), (pr, g.shape_ghost), (convh0, g.shape_convh0)):
            if a is not None and a.shape != shp:
                raise ValueError(f"shape {a.shape} != {shp}")
        nat.check(self.lib.chs_upload_state(self.h, nat.hptr(q), nat.hptr(pr), nat.hptr(convh0)))

    def download_state(self):
        g = self.grid
        q, pr, c0 = g.zeros(g.shape_q), g.zeros(g.shape_ghost), g.zeros(g.shape_convh0)
        nat.check(self.lib.chs_download_state(self.h, nat.hptr(q), nat.hptr(pr), nat.hptr(c0)))
        return q, pr, c0

    def write_instant_vars_tg(self, iterg: int, phytime: float, dt: float, filepath: str = None):
        """WRT_INSTANT_VARS_TG of the device state: the four reference-named
        files DNS_perixz_INSTANT_T<ES15.9 time>_{U,V,W,P}.D under filepath
        (default filepath1 = '1_instant_D/').  Returns their paths."""
        from . import fields
        q, pr, _ = self.download_state()
        return fields.write_instant_vars_tg(q, pr, iterg, phytime, self.grid.ren, dt,
                                            fields.INSTANT_DIR if filepath is None else filepath)

    def restart_instant_vars_tg(self, phytime: float, filepath: str = None):
        """RESTART_INSTANT_VARS_TG(phytime) into the device state: Q and PR from
        the four files with the wall ghost rows of BC_WALL_Q_tg/BC_WALL_PR_tg,
        CONVH0 zero (the NS = 1 substep does not read it).  Returns (ITERG,
        phyTIME) of the files."""
        from . import fields
        g = self.grid
        q, pr, iterg, t = fields.restart_instant_vars_tg(g, phytime, fields.INSTANT_DIR if filepath is None
                                                         else filepath)
        self.upload_state(q, pr, g.zeros(g.shape_convh0))
        return iterg, t

    def device_arrays(self):
        ptrs = [C.c_void_p() for _ in range(5)]
        nat.check(self.lib.chs_device_arrays(self.h, *[C.byref(p) for p in ptrs]))
        return dict(zip(("q", "pr", "convh0", "dph", "rhsllphi"), (p.value for p in ptrs)))

    # ---- solver routines -------------------------------------------------
    def solverrk3_mom_tg(self, ns: int, dt: float):
        nat.check(self.lib.chs_rk3_substep(self.h, int(ns), float(dt)))

    def set_cvisc(self, cvisc: float):
        """CVISC = 1/REN from here on: the REINI -> REN switch at TLGRE
        (src/SOLVE.f90:97-100, BCAST_COMM_STEP :167)."""
        nat.check(self.lib.chs_set_cvisc(self.h, float(cvisc)))

    def fft99_pois3d(self, f_dev, dph_dev):
        """Poisson solve on device tensors: f (L*M*N, RHSLLPHI layout) -> dph (L*(M+2)*N, DPH layout)."""
        g = self.grid
        if f_dev.numel() != g.ncell or dph_dev.numel() != g.L * (g.M + 2) * g.N:
            raise ValueError("fft99_pois3d: wrong array sizes")
        nat.check(self.lib.chs_poisson_fft99(self.h, nat.dptr(f_dev), nat.dptr(dph_dev)))

    def set_refinement(self, rel_tol: float, max_iter: int):
        nat.check(self.lib.chs_set_refinement(self.h, rel_tol, max_iter))

    def poisson_stats(self):
        it, rel = C.c_int(), C.c_double()
        nat.check(self.lib.chs_poisson_stats(self.h, C.byref(it), C.byref(rel)))
        return it.value, rel.value

    def set_cn_solver(self, mode: str):
        """'auto' (truncated-inverse stencils where their bound holds; the
        y-marching RHS kernels where their chunks have 8 rows or more),
        'thomas', 'rowtile' ('auto' with the per-row RHS kernels only) or
        'marching' ('auto' with the y-marching RHS kernels wherever they fit)."""
        nat.check(self.lib.chs_set_cn_solver(self.h, {"auto": 0, "thomas": 1, "rowtile": 2, "marching": 3}[mode]))

    def cn_info(self, ns: int, dt: float):
        p, kx, kz, ky = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        nat.check(self.lib.chs_cn_info(self.h, ns, dt, C.byref(p), C.byref(kx), C.byref(kz), C.byref(ky)))
        return {"path": "stencil" if p.value else "thomas", "kx": kx.value, "kz": kz.value, "ky": ky.value}

    def divgck(self) -> float:
        v = C.c_double()
        nat.check(self.lib.chs_divgck(self.h, C.byref(v)))
        return v.value

    def cfl(self) -> float:
        v = C.c_double()
        nat.check(self.lib.chs_cfl(self.h, C.byref(v)))
        return v.value

    def mean_t(self, domain: str = "tg", tp: int = 1, pp_instn_sz: int = 1):
        """PP_MEAN_T_TG ("tg") or PP_MEAN_T_Xperiodic_IO ("io"): the last xz
        moments (mean_xz / HeatedChannel.mean_xz_io) into the time means.
        Returns (time means (NCL2, 13 or 37), NSTATIS)."""
        d = {"tg": 0, "io": 1}[domain]
        out = np.zeros((13 if d == 0 else 37) * self.grid.M)
        ns = C.c_int()
        nat.check(self.lib.chs_mean_t(self.h, d, int(tp), int(pp_instn_sz), C.byref(ns),
                                      out.ctypes.data_as(C.POINTER(C.c_double))))
        return out.reshape(self.grid.M, -1), ns.value

    def mean_t_reset(self, domain: str = "tg"):
        """PP_TMEAN_INI_ZERO_TG / _IO: NSTATIS = 0, time means zero."""
        nat.check(self.lib.chs_mean_t_reset(self.h, {"tg": 0, "io": 1}[domain]))

    def mean_xz(self):
        """xz-plane means of the device state, PP_MEAN_ZX_FLOW_TG
        (src/processing/PP_MEAN_ZX.f90:104-237): array (NCL2, 13) with columns
        U1xzL(1:4) = <u>, <v>, <w>, <p>; U2xzL(1:6) = <uu> <uv> <uw> <vv> <vw> <ww>;
        UPxzL(1:3) = <up> <vp> <wp> (cell-centred, the reference's scaling)."""
        out = np.zeros(13 * self.grid.M)
        nat.check(self.lib.chs_mean_xz(self.h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out.reshape(self.grid.M, 13)

    def last_dpgrns(self) -> float:
        v = C.c_double()
        nat.check(self.lib.chs_last_dpgrns(self.h, C.byref(v)))
        return v.value

    # ---- per-kernel timing ------------------------------------------------
    def profile(self, on: bool = True):
        nat.check(self.lib.chs_profile_enable(self.h, int(on)))

    def profile_reset(self):
        nat.check(self.lib.chs_profile_reset(self.h))

    def profile_read(self):
        cap = 64
        names = C.create_string_buffer(48 * cap)
        ms = (C.c_double * cap)()
        calls = (C.c_long * cap)()
        n = C.c_int()
        nat.check(self.lib.chs_profile_read(self.h, cap, names, ms, calls, C.byref(n)))
        out = {}
        for i in range(min(n.value, cap)):
            nm = names.raw[48 * i:48 * (i + 1)].split(b"\0", 1)[0].decode()
            out[nm] = (ms[i], calls[i])
        return out


# ---- the reference's four line solvers on device tensors -------------------
def tdma_batched(a, b, c, r, bc, n: int, nlines: int, elem_stride: int, line_stride: int, cyclic: bool,
                 stream=None, noalloc: bool = False):
    """Batched TDMAIJ?_(non)CYC line solves on device tensors (r is
    overwritten).  noalloc: use chs_tdma_batched_noalloc, which allocates
    nothing (workspace from the library's static per-GPU scratch array)."""
    lib = nat.load()
    s = None if stream is None else C.c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream))
    need = (nlines - 1) * line_stride + (n - 1) * elem_stride + 1 if nlines > 0 else 0
    for t in (a, b, c, r):
        if t.numel() < need:
            raise ValueError("tdma_batched: array smaller than the strides describe")
    if bc is not None and bc.numel() < 2 * nlines:
        raise ValueError("tdma_batched: bc needs 2*nlines values")
    fn = lib.chs_tdma_batched_noalloc if noalloc else lib.chs_tdma_batched
    nat.check(fn(nat.dptr(a), nat.dptr(b), nat.dptr(c), nat.dptr(r), None if bc is None else nat.dptr(bc), n, nlines,
                 elem_stride, line_stride, int(bool(cyclic)), s))


def tdmaijj_noncyc(A, B, Cc, R, BCJ, JS, JSZ, IS, ISZ, stream=None):
    """TDMAIJJ_nonCYC(A,B,C,R,BCJ,JS,JSZ,IS,ISZ): march along index 2 of (ISZ, JSZ) arrays."""
    tdma_batched(A, B, Cc, R, BCJ, JSZ, ISZ, ISZ, 1, False, stream)


def tdmaiji_noncyc(A, B, Cc, R, BCI, IS, ISZ, JS, JSZ, stream=None):
    """TDMAIJI_nonCYC(A,B,C,R,BCI,IS,ISZ,JS,JSZ): march along index 1 of (ISZ, JSZ) arrays."""
    tdma_batched(A, B, Cc, R, BCI, ISZ, JSZ, 1, ISZ, False, stream)


def tdmaijj_cyc(A, B, Cc, R, JS, JSZ, IS, ISZ, stream=None):
    """TDMAIJJ_CYC(A,B,C,R,JS,JSZ,IS,ISZ)."""
    tdma_batched(A, B, Cc, R, None, JSZ, ISZ, ISZ, 1, True, stream)


def tdmaiji_cyc(A, B, Cc, R, IS, ISZ, JS, JSZ, stream=None):
    """TDMAIJI_CYC(A,B,C,R,IS,ISZ,JS,JSZ)."""
    tdma_batched(A, B, Cc, R, None, ISZ, JSZ, 1, ISZ, True, stream)
