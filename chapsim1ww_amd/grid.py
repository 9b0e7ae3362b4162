"""Host-side mesh, metrics and line-solver coefficients of the channel.

This is the set-up the reference performs once on the host before its time
loop; the GPU library receives the resulting 1-D arrays through the C ABI
(``chs_create``), exactly as the reference's kernels read them from its
``mesh_info`` / ``init_info`` modules.

Reference routines restated here (channel, ``ICASE = ICHANL``, ``ISTR2 = 1``):

* ``COORDJR``               src/COORDJR.f90:28-38 (tanh stretching), :112-131 (DYFI/DYCI)
* ``AREA_INOUTLET``         src/COORDJR.f90:355-399 (VOLM_tg)
* ``MESH_YWEIGHT_FACTORS``  src/COORDJR.f90:427-470 (YCL2ND_WFF/WFB)
* ``CONSPARA``              src/CONSPARA.f90:15-110 (DX, DZ, DXI, DXQI, ...)
* ``TDMA_COEF``             src/TDMA_COEF.f90:30-116 (AMPH/ACPH/APPH, AMVR/ACVR/APVR)
* ``FFT99_ROOT``            src/possion3d_FFT99.f90:124-182 (modified wavenumbers AK1, AK3)
* ``CONS_RKCOEF``           src/CONS_RKCOEF.f90:18-54 (RK3 coefficients)

Scalar ``math`` calls are used on purpose, in the reference's operation
order, so that every coefficient is the same double the reference computes.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

ICHANL = 1
IBOX3P = 4

# RK3 coefficients: gamma, rho, alpha = gamma + rho (src/CONS_RKCOEF.f90:24-54)
TGAM = (1.0, 8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0)
TROH = (0.0, 0.0, -17.0 / 60.0, -5.0 / 12.0)
TALP = tuple(g + r for g, r in zip(TGAM, TROH))


def _check_fft_length(n: int, name: str) -> None:
    """FFT99 accepts lengths whose factors are 2, 3 and 5 (libs/fft99/FFTFAX.f90:78-84)."""
    if n < 4 or n % 2:
        raise ValueError(f"{name}={n}: FFT length must be even and >= 4")
    m = n
    for p in (2, 3, 5):
        while m % p == 0:
            m //= p
    if m != 1:
        raise ValueError(f"{name}={n}: FFT length may only have prime factors 2, 3 and 5")


@dataclass
class ChannelGrid:
    """Staggered channel grid, x and z periodic, y in [-1, 1] with walls."""

    L: int                      # NCL1 (x cells)
    M: int                      # NCL2 (y cells)
    N: int                      # NCL3 (z cells)
    hx: float = 2.0 * math.pi   # HX_tg
    hz: float = math.pi         # HZ
    str2: float = 2.0           # STR2, tanh stretching
    ren: float = 2800.0         # REN, bulk Reynolds number
    icase: int = ICHANL
    # derived
    ynd: np.ndarray = field(init=False, repr=False)
    ycc: np.ndarray = field(init=False, repr=False)
    dyfi: np.ndarray = field(init=False, repr=False)
    dyci: np.ndarray = field(init=False, repr=False)
    wff: np.ndarray = field(init=False, repr=False)
    wfb: np.ndarray = field(init=False, repr=False)
    amph: np.ndarray = field(init=False, repr=False)
    acph: np.ndarray = field(init=False, repr=False)
    apph: np.ndarray = field(init=False, repr=False)
    amvr: np.ndarray = field(init=False, repr=False)
    acvr: np.ndarray = field(init=False, repr=False)
    apvr: np.ndarray = field(init=False, repr=False)
    ak1: np.ndarray = field(init=False, repr=False)
    ak3: np.ndarray = field(init=False, repr=False)

    def __post_init__(self) -> None:
        if self.icase != ICHANL:
            raise ValueError("only the wall-bounded channel (ICASE=1) is supported")
        if self.M < 4:
            raise ValueError("M (NCL2) must be >= 4")
        _check_fft_length(self.L, "L")
        _check_fft_length(self.N, "N")
        L, M, N = self.L, self.M, self.N
        NND2 = M + 1
        s2 = self.str2
        ynd = [math.tanh(s2 * ((J - 1) / (NND2 - 1) - 0.5)) / math.tanh(s2 * 0.5)
               for J in range(1, NND2 + 1)]
        ycc = [math.tanh(s2 * (((J - 1) + 0.5) / (NND2 - 1) - 0.5)) / math.tanh(s2 * 0.5)
               for J in range(1, M + 1)]
        dyfi = [1.0 / (ynd[J] - ynd[J - 1]) for J in range(1, M + 1)]
        dyci = [0.0] * NND2
        for J in range(2, NND2):
            dyci[J - 1] = 1.0 / ((ynd[J] - ynd[J - 2]) * 0.5)
        dyci[0] = 2.0 / (ynd[1] - ynd[0])
        dyci[NND2 - 1] = 2.0 / (ynd[NND2 - 1] - ynd[M - 1])
        wff = [0.0] * NND2
        wfb = [0.0] * NND2
        for JC in range(2, M + 1):
            JM = JC - 1
            den = ycc[JC - 1] - ycc[JM - 1]
            wff[JC - 1] = (ynd[JC - 1] - ycc[JM - 1]) / den
            wfb[JC - 1] = (ycc[JC - 1] - ynd[JC - 1]) / den
        wfb[0], wff[0] = 1.0, 0.0
        wfb[NND2 - 1], wff[NND2 - 1] = 0.0, 1.0

        self.dx = self.hx / L
        self.dxi = 1.0 / self.dx
        self.dxqi = self.dxi * self.dxi
        self.dz = self.hz / N
        self.dzi = 1.0 / self.dz
        self.dzqi = self.dzi * self.dzi
        self.cvisc = 1.0 / self.ren
        # VOLM_tg: 1/RCCI1/DYFI accumulated in the loop order K, J, I of
        # AREA_INOUTLET (RCCI1 = 1); add.accumulate is a sequential sum.
        terms = np.tile(np.repeat(1.0 / np.array(dyfi, np.float64), L), N)
        self.volm = float(np.add.accumulate(terms)[-1])
        del terms

        amph = [0.0] * M
        apph = [0.0] * M
        for JC in range(1, M + 1):
            JP = JC + 1
            amph[JC - 1] = dyfi[JC - 1] * dyci[JC - 1]
            apph[JC - 1] = dyfi[JC - 1] * dyci[JP - 1]
            if JC == 1:
                amph[JC - 1] = dyfi[JC - 1] * dyci[JC - 1] * 0.5
            if JP == NND2:
                apph[JC - 1] = dyfi[JC - 1] * dyci[JP - 1] * 0.5
        acph = [-(amph[j] + apph[j]) for j in range(M)]
        acph[0] = amph[0] + acph[0]
        amph[0] = 0.0
        acph[M - 1] = apph[M - 1] + acph[M - 1]
        apph[M - 1] = 0.0

        amvr = np.zeros((3, NND2))
        acvr = np.zeros((3, NND2))
        apvr = np.zeros((3, NND2))
        for IDR in (1, 2, 3):
            d = IDR - 1
            if IDR in (1, 3):
                for JC in range(2, M):
                    JP = JC + 1
                    amvr[d, JC - 1] = dyfi[JC - 1] * dyci[JC - 1]
                    apvr[d, JC - 1] = dyfi[JC - 1] * dyci[JP - 1]
                    acvr[d, JC - 1] = -(amvr[d, JC - 1] + apvr[d, JC - 1])
                ucaj = 4.0 / (2.0 / dyci[1] + 1.0 / dyfi[0])
                amvr[d, 0] = 0.0
                apvr[d, 0] = ucaj * dyci[1]
                acvr[d, 0] = -ucaj * (dyci[1] + 2.0 * dyfi[0])
                ucaj = 4.0 / (2.0 / dyci[M - 1] + 1.0 / dyfi[M - 1])
                amvr[d, M - 1] = ucaj * dyci[M - 1]
                apvr[d, M - 1] = 0.0
                acvr[d, M - 1] = -ucaj * (2.0 * dyfi[M - 1] + dyci[M - 1])
            else:
                for JC in range(2, M + 1):
                    JM = JC - 1
                    amvr[d, JC - 1] = dyfi[JM - 1] * dyci[JC - 1]
                    apvr[d, JC - 1] = dyfi[JC - 1] * dyci[JC - 1]
                    acvr[d, JC - 1] = -(amvr[d, JC - 1] + apvr[d, JC - 1])
                amvr[d, 0], apvr[d, 0], acvr[d, 0] = 0.0, 0.0, 1.0
                amvr[d, NND2 - 1], apvr[d, NND2 - 1], acvr[d, NND2 - 1] = 0.0, 0.0, 1.0

        pi = 2.0 * math.asin(1.0)
        ak3 = [0.0] * N
        for K in range(1, N // 2 + 1):
            ak3[K - 1] = 2.0 * (1.0 - math.cos(((K - 1) * 2.0 * pi) / N)) * self.dzqi
        for K in range(N // 2 + 1, N + 1):
            ak3[K - 1] = 2.0 * (1.0 - math.cos((-(N - K + 1) * 2.0 * pi) / N)) * self.dzqi
        ak1 = [0.0] * L
        for I in range(1, L // 2 + 1):
            ak1[I - 1] = 2.0 * (1.0 - math.cos(((I - 1) * 2.0 * pi) / L)) * self.dxqi
        for I in range(L // 2 + 1, L + 1):
            ak1[I - 1] = 2.0 * (1.0 - math.cos((-(L - I + 1) * 2.0 * pi) / L)) * self.dxqi

        f64 = np.float64
        self.ynd = np.array(ynd, f64)
        self.ycc = np.array(ycc, f64)
        self.dyfi = np.array(dyfi, f64)
        self.dyci = np.array(dyci, f64)
        self.wff = np.array(wff, f64)
        self.wfb = np.array(wfb, f64)
        self.amph = np.array(amph, f64)
        self.acph = np.array(acph, f64)
        self.apph = np.array(apph, f64)
        self.amvr = np.ascontiguousarray(amvr)
        self.acvr = np.ascontiguousarray(acvr)
        self.apvr = np.ascontiguousarray(apvr)
        self.ak1 = np.array(ak1, f64)
        self.ak3 = np.array(ak3, f64)

    @property
    def ncell(self) -> int:
        return self.L * self.M * self.N

    # ---- shapes of the reference's state arrays (Fortran order) ----------
    @property
    def shape_q(self):       # Q_tg(NCL1, 0:N2DO+1, NCL3, 3)
        return (self.L, self.M + 2, self.N, 3)

    @property
    def shape_ghost(self):   # PR_tg / DPH_tg (NCL1, 0:N2DO+1, NCL3)
        return (self.L, self.M + 2, self.N)

    @property
    def shape_cell(self):    # RHS_tg / RHSLLPHI_tg (NCL1, N2DO, NCL3)
        return (self.L, self.M, self.N)

    @property
    def shape_convh0(self):  # CONVH0_tg (NCL1, N2DO, NCL3, 3)
        return (self.L, self.M, self.N, 3)

    def zeros(self, shape) -> np.ndarray:
        return np.zeros(shape, dtype=np.float64, order="F")
