"""ctypes front-end of the CPU oracle (oracle/chapsim_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
the cpu_baseline leg of bench.py, always as the checker or the timed CPU
baseline, never as part of the product path.

Also loads, when they exist, the reference's own FFT99 package
(libs/fft99/*.f90, oracle/_ref/libfft99_ref.so) and its four tridiagonal line
solvers (src/TDMAIJ{I,J}_{CYC,nonCYC}.f90, oracle/_ref/libtdma_ref.so), both
compiled from the reference sources with amdflang by oracle/Makefile, used to
pin the oracle's Fourier conventions and line solves and to write the golden
fixtures under tests/golden/.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libfft99_ref.so")
REF_TDMA_PATH = os.path.join(HERE, "_ref", "libtdma_ref.so")

_dp = C.POINTER(C.c_double)


def build(ref: bool = False) -> None:
    """Compile the oracle (and optionally the reference FFT99 checker)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if ref and os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def _ptr(a: np.ndarray):
    if a is None:
        return None
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.flags.c_contiguous), "need contiguous float64"
    return a.ctypes.data_as(_dp)


class _Grid(C.Structure):
    _fields_ = [("L", C.c_int), ("M", C.c_int), ("N", C.c_int), ("icase", C.c_int),
                ("dx", C.c_double), ("dz", C.c_double), ("dxi", C.c_double), ("dzi", C.c_double),
                ("dxqi", C.c_double), ("dzqi", C.c_double), ("cvisc", C.c_double), ("volm", C.c_double)] + \
               [(n, _dp) for n in ("ynd", "ycc", "dyfi", "dyci", "wff", "wfb", "amph", "acph", "apph",
                                   "amvr", "acvr", "apvr", "ak1", "ak3")]


class _State(C.Structure):
    _fields_ = [(n, _dp) for n in ("q", "pr", "convh0", "dph", "rhs", "rhsllphi", "qtmp")] + \
               [("dpgrns", C.c_double)]


_lib = None
_ref = None
_ref_tdma = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_grid_setup.argtypes = [C.POINTER(_Grid), C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_tdma_noncyc.argtypes = [_dp, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_long, C.c_long]
        L.orc_tdma_cyc.argtypes = [_dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_long, C.c_long]
        L.orc_cfft99.argtypes = [_dp, C.c_int, C.c_int]
        L.orc_fft99_fwd.argtypes = [_dp, _dp, C.c_int]
        L.orc_fft99_inv.argtypes = [_dp, _dp, C.c_int]
        L.orc_poisson_fft99.argtypes = [C.POINTER(_Grid), _dp, _dp, C.c_int]
        L.orc_poisson_residual.argtypes = [C.POINTER(_Grid), _dp, _dp, _dp]
        L.orc_poisson_residual.restype = C.c_double
        for name in ("orc_convection_x", "orc_convection_y", "orc_convection_z", "orc_bc_wall_q"):
            getattr(L, name).argtypes = [C.POINTER(_Grid), C.POINTER(_State)]
        L.orc_rhs_tg.argtypes = [C.POINTER(_Grid), C.POINTER(_State), C.c_int, C.c_int, C.c_double, C.c_int,
                                 C.c_int, C.c_double]
        L.orc_momfa_tg.argtypes = [C.POINTER(_Grid), C.POINTER(_State), C.c_int, C.c_int, C.c_double]
        for name in ("orc_divg_tg", "orc_veloupdt_tg", "orc_prcalc_tg"):
            getattr(L, name).argtypes = [C.POINTER(_Grid), C.POINTER(_State), C.c_int, C.c_double]
        L.orc_bc_wall_neumann.argtypes = [C.POINTER(_Grid), _dp]
        L.orc_divgck_tg.argtypes = [C.POINTER(_Grid), C.POINTER(_State)]
        L.orc_divgck_tg.restype = C.c_double
        L.orc_cfl_tg.argtypes = [C.POINTER(_Grid), C.POINTER(_State)]
        L.orc_cfl_tg.restype = C.c_double
        L.orc_mean_xz.argtypes = [C.POINTER(_Grid), C.POINTER(_State), _dp]
        L.orc_mean_xz.restype = None
        L.orc_substep_tg.argtypes = [C.POINTER(_Grid), C.POINTER(_State), C.c_int, C.c_double, C.c_int, C.c_int,
                                     C.c_double]
        L.orc_cubic_spline.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_sleval.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_int, _dp]
        L.orc_sleval.restype = C.c_int
        L.orc_therm_prop_update.argtypes = [C.c_int, _dp, C.c_long, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_therm_prop_update.restype = C.c_long
        L.orc_energy_substep.argtypes = [C.c_void_p, C.c_void_p, C.c_int, _dp, C.c_int, C.c_double, C.c_double,
                                         C.c_double]
        L.orc_energy_substep.restype = C.c_long
        L.orc_num_threads.restype = C.c_int
        L.orc_set_num_threads.argtypes = [C.c_int]
        for name in ("orc_io_prepare", "orc_io_substep"):
            getattr(L, name).argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int] + \
                ([C.c_double] if name == "orc_io_substep" else [])
            getattr(L, name).restype = None
        L.orc_io_momfa.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_double, _dp]
        L.orc_io_momfa.restype = None
        L.orc_io_terms.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_io_terms.restype = None
        L.orc_divgck_io.argtypes = [C.c_void_p, C.c_void_p, _dp]
        L.orc_divgck_io.restype = None
        _lib = L
    return _lib


def ref_fft99():
    """The reference FFT99 package, or None when oracle/_ref was not built."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        _ref = C.CDLL(REF_PATH)
    return _ref


def ref_tdma():
    """The reference's four line solvers, or None when oracle/_ref was not built."""
    global _ref_tdma
    if _ref_tdma is None and os.path.exists(REF_TDMA_PATH):
        _ref_tdma = C.CDLL(REF_TDMA_PATH)
    return _ref_tdma


# --------------------------------------------------------------------------
class OracleGrid:
    """The oracle's own restatement of COORDJR/TDMA_COEF/FFT99_ROOT."""

    def __init__(self, L, M, N, hx=2 * np.pi, hz=np.pi, str2=2.0, ren=2800.0):
        self.L, self.M, self.N = L, M, N
        self.arr = {
            "ynd": np.zeros(M + 1), "ycc": np.zeros(M), "dyfi": np.zeros(M), "dyci": np.zeros(M + 1),
            "wff": np.zeros(M + 1), "wfb": np.zeros(M + 1), "amph": np.zeros(M), "acph": np.zeros(M),
            "apph": np.zeros(M), "amvr": np.zeros(3 * (M + 1)), "acvr": np.zeros(3 * (M + 1)),
            "apvr": np.zeros(3 * (M + 1)), "ak1": np.zeros(L), "ak3": np.zeros(N),
        }
        g = _Grid()
        g.L, g.M, g.N, g.icase = L, M, N, 1
        for k, v in self.arr.items():
            setattr(g, k, _ptr(v))
        lib().orc_grid_setup(C.byref(g), hx, hz, str2, ren)
        self.g = g
        for k in ("dx", "dz", "dxi", "dzi", "dxqi", "dzqi", "cvisc", "volm"):
            setattr(self, k, getattr(g, k))
        for k, v in self.arr.items():
            setattr(self, k, v)


class OracleState:
    """Reference-layout state and work arrays for the tg substep."""

    def __init__(self, og: OracleGrid, q, pr, convh0):
        L, M, N = og.L, og.M, og.N
        f = lambda a: np.array(a, dtype=np.float64, order="F", copy=True)
        self.q, self.pr, self.convh0 = f(q), f(pr), f(convh0)
        self.dph = np.zeros((L, M + 2, N), order="F")
        self.rhs = np.zeros((L, M, N), order="F")
        self.rhsllphi = np.zeros((L, M, N), order="F")
        self.qtmp = np.zeros((L, M, N), order="F")
        s = _State()
        for k in ("q", "pr", "convh0", "dph", "rhs", "rhsllphi", "qtmp"):
            setattr(s, k, _ptr(getattr(self, k)))
        s.dpgrns = 0.0
        self.s = s
        self.og = og

    @property
    def dpgrns(self):
        return self.s.dpgrns

    def substep(self, ns: int, dt: float, flowdrvtp: int = 1, nflow: int = 1, cfgv: float = 0.0):
        """SOLVERRK3_MOM_tg(NS); flowdrvtp 1 (constant mass flux) or 2
        (constant pressure gradient -CFGV/2), streamwise direction nflow 1 or 3."""
        if nflow not in (1, 3) or flowdrvtp not in (1, 2):
            raise ValueError("nflow must be 1 or 3 and flowdrvtp 1 or 2")
        lib().orc_substep_tg(C.byref(self.og.g), C.byref(self.s), ns, dt, flowdrvtp, nflow, cfgv)

    def divgck(self) -> float:
        return lib().orc_divgck_tg(C.byref(self.og.g), C.byref(self.s))

    def cfl(self) -> float:
        return lib().orc_cfl_tg(C.byref(self.og.g), C.byref(self.s))

    def mean_xz(self):
        """PP_MEAN_ZX_FLOW_TG first and second moments, (M, 13)."""
        out = np.zeros(13 * self.og.M)
        lib().orc_mean_xz(C.byref(self.og.g), C.byref(self.s), _ptr(out))
        return out.reshape(self.og.M, 13)

    def stage(self, name: str, *args):
        fn = getattr(lib(), "orc_" + name)
        return fn(C.byref(self.og.g), C.byref(self.s), *args)


# --------------------------------------------------------------------------
def tdma_noncyc(a, b, c, r, bc, n, nlines, es, ls):
    lib().orc_tdma_noncyc(_ptr(a), _ptr(b), _ptr(c), _ptr(r), _ptr(bc), n, nlines, es, ls)


def tdma_cyc(a, b, c, r, n, nlines, es, ls):
    lib().orc_tdma_cyc(_ptr(a), _ptr(b), _ptr(c), _ptr(r), n, nlines, es, ls)


def fft99_fwd(x):
    n = x.shape[0]
    ab = np.zeros(n + 2)
    lib().orc_fft99_fwd(_ptr(np.ascontiguousarray(x, np.float64)), _ptr(ab), n)
    return ab


def fft99_inv(ab, n):
    x = np.zeros(n)
    lib().orc_fft99_inv(_ptr(np.ascontiguousarray(ab, np.float64)), _ptr(x), n)
    return x


def cfft99(z, isign):
    a = np.ascontiguousarray(np.asarray(z, np.complex128)).view(np.float64).copy()
    lib().orc_cfft99(_ptr(a), a.shape[0] // 2, isign)
    return a.view(np.complex128)


def poisson(og: OracleGrid, f, cyclic_y=False):
    dph = np.zeros((og.L, og.M + 2, og.N), order="F")
    lib().orc_poisson_fft99(C.byref(og.g), _ptr(np.asfortranarray(f)), _ptr(dph), int(cyclic_y))
    return dph


def poisson_residual(og: OracleGrid, dph, f):
    """max |L_h(dph) - f| with DPH ghosts set by the Neumann wall rule."""
    d = np.array(dph, order="F", copy=True)
    lib().orc_bc_wall_neumann(C.byref(og.g), _ptr(d))
    res = np.zeros((og.L, og.M, og.N), order="F")
    mx = lib().orc_poisson_residual(C.byref(og.g), _ptr(d), _ptr(np.asfortranarray(f)), _ptr(res))
    return mx, res


def cubic_spline(x, y):
    n = len(x)
    b, c, d = np.zeros(n), np.zeros(n), np.zeros(n)
    lib().orc_cubic_spline(n, _ptr(np.ascontiguousarray(x, np.float64)), _ptr(np.ascontiguousarray(y, np.float64)),
                           _ptr(b), _ptr(c), _ptr(d))
    return b, c, d


def build_prop_tab(scaled):
    """The 21n-value table of orc_therm_prop_update from scaled columns
    (NIST_SPLINE_H/T): [H T D M K CP | HT(b c d) | TD | TM | TK | TCP]."""
    H, T = np.asarray(scaled["H"], np.float64), np.asarray(scaled["T"], np.float64)
    cols = [H, T] + [np.asarray(scaled[k], np.float64) for k in ("D", "M", "K", "CP")]
    parts = list(cols)
    parts += list(cubic_spline(H, T))
    for k in ("D", "M", "K", "CP"):
        parts += list(cubic_spline(T, scaled[k]))
    return np.ascontiguousarray(np.concatenate(parts)), len(H)


def therm_prop_update(tab, n, h, T, rho, mu, kap, cp=None):
    return lib().orc_therm_prop_update(n, _ptr(tab), h.size, _ptr(h), _ptr(T), _ptr(rho), _ptr(mu), _ptr(kap), _ptr(cp))


class _Thermal(C.Structure):
    _fields_ = [(n, _dp) for n in ("rhoh", "h", "T", "rho", "rho0", "rho1", "mu", "mu0", "kap", "rhs0", "g", "g0",
                                   "wallflux")]


THERMAL_FIELDS = ("rhoh", "h", "T", "rho", "rho0", "rho1", "mu", "mu0", "kap", "rhs0")


def energy_substep(og: "OracleGrid", st: dict, tab, n, ns, dt, cthecd, rhohmax):
    """SOLVERRK3_ENG_IO on the numpy arrays of st (Fortran order, updated in
    place): rhoh h T rho rho0 rho1 mu mu0 kap (L, M+2, N), rhs0 (L, M, N),
    g g0 (L, M+2, N, 3), wallflux (L, 2).  Returns the out-of-range count."""
    for k in THERMAL_FIELDS + ("g", "g0", "wallflux"):
        a = st[k]
        assert a.dtype == np.float64 and a.flags.f_contiguous, k
    t = _Thermal(*[_ptr(st[k]) for k in THERMAL_FIELDS + ("g", "g0", "wallflux")])
    return lib().orc_energy_substep(C.byref(og.g), C.byref(t), n, _ptr(tab), ns, dt, cthecd, rhohmax)


def rhoh_max(scaled):
    """RHOHmax of NIST_SPLINE_H (src/Flow_State.f90:960-990): max of D(H) * H
    over H_I = H_1 + (I-1)(H_n - H_1)/(3n), I = 2..3n, D from the H -> D spline."""
    H = np.ascontiguousarray(scaled["H"], np.float64)
    D = np.ascontiguousarray(scaled["D"], np.float64)
    n = len(H)
    b, c, d = cubic_spline(H, D)
    best = -1e-14
    e = np.zeros(1)
    for I in range(2, 3 * n + 1):
        hv = H[0] + float(I - 1) * (H[-1] - H[0]) / float(3 * n)
        assert lib().orc_sleval(n, _ptr(H), _ptr(D), _ptr(b), _ptr(c), _ptr(d), hv, 1, _ptr(e)) == 0
        if e[0] * hv > best:
            best = e[0] * hv
    return best


def thermal_case(og: "OracleGrid", scaled, seed=0, qw=0.1):
    """A consistent random thermal state on og's grid: h smooth + noise inside
    the table, T, rho, mu, kappa from the oracle's table lookup, rho*h,
    histories, random mass fluxes with zero wall-normal flux at the walls."""
    L, M, N = og.L, og.M, og.N
    rng = np.random.default_rng(seed)
    tab, n = build_prop_tab(scaled)
    y = np.concatenate([[og.ynd[0]], og.ycc, [og.ynd[-1]]])
    x = np.arange(L) / L * 2 * np.pi
    z = np.arange(N) / N * 2 * np.pi
    T0 = 1.0 + 0.25 * (1 - y[None, :, None] ** 2) + 0.03 * np.sin(x)[:, None, None] * np.cos(z)[None, None, :]
    T0 = T0 + 0.005 * rng.standard_normal((L, M + 2, N))
    h = np.interp(T0.ravel(), scaled["T"], scaled["H"]).reshape(T0.shape)
    f = lambda a: np.asfortranarray(a, dtype=np.float64)
    h = f(h)
    out = [np.zeros(h.size) for _ in range(5)]
    assert therm_prop_update(tab, n, np.ascontiguousarray(h.ravel(order="F")), *out) == 0
    T, rho, mu, kap, _ = (f(o.reshape(h.shape, order="F")) for o in out)
    st = {"h": h, "T": T, "rho": rho, "mu": mu, "kap": kap, "rhoh": f(h * rho),
          "rho0": f(rho * (1 + 1e-3 * rng.standard_normal(rho.shape))), "rho1": f(rho.copy()),
          "mu0": f(mu.copy()), "rhs0": f(1e-3 * rng.standard_normal((L, M, N)))}
    g = 1.0 + 0.2 * rng.standard_normal((L, M + 2, N, 3))
    g[:, :, :, 1] *= 0.1
    g[:, 1, :, 1] = 0.0          # wall faces: no wall-normal mass flux
    g[:, M + 1, :, 1] = 0.0
    g0 = g * (1 + 0.01 * rng.standard_normal(g.shape))
    g0[:, 1, :, 1] = 0.0
    g0[:, M + 1, :, 1] = 0.0
    st["g"], st["g0"] = f(g), f(g0)
    st["wallflux"] = f(np.stack([np.full(L, qw), np.full(L, qw)], 1))
    return st, tab, n


class _IO(C.Structure):
    _fields_ = [(n, _dp) for n in ("q", "g", "g0", "pr", "rho", "rho0", "mu", "mu0", "drhodtp", "divu", "rho_stg",
                                   "mu_stg", "explt0", "dph", "hx", "hy", "hz", "rhsllphi")] + \
               [("dpgrns", C.c_double), ("drhoi_stg", _dp), ("im", _dp), ("ad", _dp)]


class _IOParams(C.Structure):
    _fields_ = [("thermal", C.c_int), ("nflow", C.c_int), ("flowdrvtp", C.c_int), ("gravdir", C.c_int),
                ("cfgv", C.c_double), ("f_a", C.c_double), ("visimplicit", C.c_int)]


IO_FIELDS = ("q", "pr", "drhodtp", "divu", "rho_stg", "mu_stg", "explt0", "drhoi_stg")


class OracleIO:
    """The io momentum substep (SOLVERRK3_MOM_IO, explicit viscous terms) and
    the property staging before it, on the numpy arrays of st (Fortran order,
    updated in place).  st holds g g0 rho rho0 mu mu0 (shared with the energy
    substep's dict) and q pr drhodtp divu rho_stg mu_stg explt0 drhoi_stg.
    visimplicit = 1 selects the RK-CN viscous branch (MOMFA_io)."""

    def __init__(self, og: "OracleGrid", st: dict, thermal=1, nflow=1, flowdrvtp=1, gravdir=0, cfgv=0.0, f_a=0.0,
                 visimplicit=0):
        L, M, N = og.L, og.M, og.N
        for k in ("g", "g0", "rho", "rho0", "mu", "mu0") + IO_FIELDS:
            a = st[k]
            assert a.dtype == np.float64 and a.flags.f_contiguous, k
        self.og, self.st = og, st
        self.work = {"dph": np.zeros((L, M + 2, N), order="F")}
        for k in ("hx", "hy", "hz", "rhsllphi"):
            self.work[k] = np.zeros((L, M, N), order="F")
        for k in ("im", "ad"):
            self.work[k] = np.zeros((L, M, N, 3), order="F")
        self.p = _IOParams(thermal, nflow, flowdrvtp, gravdir, cfgv, f_a, visimplicit)
        self.s = _IO(*[_ptr(st[k]) for k in ("q", "g", "g0", "pr", "rho", "rho0", "mu", "mu0", "drhodtp", "divu",
                                               "rho_stg", "mu_stg", "explt0")],
                     *[_ptr(self.work[k]) for k in ("dph", "hx", "hy", "hz", "rhsllphi")], 0.0,
                     _ptr(st["drhoi_stg"]), _ptr(self.work["im"]), _ptr(self.work["ad"]))

    @property
    def dpgrns(self):
        return self.s.dpgrns

    def prepare(self, ns: int):
        """DENSITY_TIME_DERIVATION, DENSITY_Staggered, MU_Staggered, VELOCITY_CALC_io."""
        lib().orc_io_prepare(C.byref(self.og.g), C.byref(self.s), C.byref(self.p), ns)

    def terms(self):
        """Convection, DIVG_U and explicit viscous terms -> work hx, hy, hz."""
        lib().orc_io_terms(C.byref(self.og.g), C.byref(self.s))
        return self.work["hx"], self.work["hy"], self.work["hz"]

    def momfa(self, ns: int, idr: int, dt: float, r):
        """MOMFA_io(NS, IDR) on the right-hand side r (L, M, N, Fortran order,
        overwritten by the increment); G(IDR) += increment."""
        assert r.dtype == np.float64 and r.flags.f_contiguous
        lib().orc_io_momfa(C.byref(self.og.g), C.byref(self.s), ns, idr, dt, _ptr(r))

    def substep(self, ns: int, dt: float):
        lib().orc_io_substep(C.byref(self.og.g), C.byref(self.s), C.byref(self.p), ns, dt)

    def divgck(self):
        """DIVGCK_io: (max|div G + DrhoDtP|, max|div Q|)."""
        out = np.zeros(2)
        lib().orc_divgck_io(C.byref(self.og.g), C.byref(self.s), _ptr(out))
        return float(out[0]), float(out[1])


def io_case(og: "OracleGrid", scaled, seed=0, qw=0.1):
    """thermal_case plus the io momentum state: a smooth pressure, small
    explicit-term histories, and the staged arrays initialised as the
    reference allocates them (RHO_STG = MU_STG = 1, DrhoDtP = 0, Q = 0)."""
    st, tab, n = thermal_case(og, scaled, seed=seed, qw=qw)
    L, M, N = og.L, og.M, og.N
    rng = np.random.default_rng(seed + 1000)
    f = lambda a: np.asfortranarray(a, dtype=np.float64)
    x = np.arange(L) / L * 2 * np.pi
    z = np.arange(N) / N * 2 * np.pi
    pr = 0.01 * np.sin(x)[:, None, None] * np.cos(2 * z)[None, None, :] * np.ones((1, M + 2, 1))
    st["pr"] = f(pr + 1e-3 * rng.standard_normal((L, M + 2, N)))
    st["q"] = f(np.zeros((L, M + 2, N, 3)))
    st["drhodtp"] = f(np.zeros((L, M + 2, N)))
    st["divu"] = f(np.zeros((L, M + 2, N)))
    st["rho_stg"] = f(np.ones((L, M + 2, N, 3)))
    st["mu_stg"] = f(np.ones((L, M + 2, N, 3)))
    st["explt0"] = f(1e-2 * rng.standard_normal((L, M, N, 3)))
    st["drhoi_stg"] = f(np.zeros((L, M + 2, N, 3)))
    return st, tab, n


def num_threads() -> int:
    return lib().orc_num_threads()


def set_num_threads(n: int) -> None:
    """OpenMP threads of the oracle's parallel loops from here on."""
    lib().orc_set_num_threads(int(n))


# ---- reference FFT99 package (Fortran, by-reference arguments) -----------
def ref_fft99_transform(data, n, isign):
    """FFT99 on LOT vectors stored as columns of data (n+2, LOT), INC=1, JUMP=n+2."""
    R = ref_fft99()
    a = np.array(data, dtype=np.float64, order="F", copy=True)
    lot = a.shape[1]
    ifax = np.zeros(13, np.int32)
    trigs = np.zeros(3 * n // 2 + 1)
    ip = lambda v: C.byref(C.c_int(v))
    R.fftfax_(ip(n), ifax.ctypes.data_as(C.POINTER(C.c_int)), _ptr(trigs))
    work = np.zeros((n + 1) * lot + 16)
    R.fft99_(_ptr(a), _ptr(work), _ptr(trigs), ifax.ctypes.data_as(C.POINTER(C.c_int)),
             ip(1), ip(n + 2), ip(n), ip(lot), ip(isign))
    return a


def ref_cfft99_transform(z, isign):
    """CFFT99 on LOT complex vectors stored as columns of z (n, LOT), INC=1, JUMP=n."""
    R = ref_fft99()
    a = np.array(z, dtype=np.complex128, order="F", copy=True)
    n, lot = a.shape
    ifax = np.zeros(13, np.int32)
    trigs = np.zeros(2 * n)
    ip = lambda v: C.byref(C.c_int(v))
    R.cftfax_(ip(n), ifax.ctypes.data_as(C.POINTER(C.c_int)), _ptr(trigs))
    work = np.zeros(4 * n * lot + 16)
    R.cfft99_(a.ctypes.data_as(_dp), _ptr(work), _ptr(trigs), ifax.ctypes.data_as(C.POINTER(C.c_int)),
              ip(1), ip(n), ip(n), ip(lot), ip(isign))
    return a


# ---- reference line solvers (Fortran, by-reference arguments) --------------
def ref_tdma_solve(a, b, c, r, bc, march, cyclic, lo=(1, 1)):
    """One call of the reference solver for arrays of shape (ISZ, JSZ) with
    Fortran lower bounds lo = (IS, JS):
      march "IJJ": TDMAIJJ_nonCYC(A,B,C,R,BCJ,JS,JSZ,IS,ISZ) (BCJ(ISZ,2)) or
                   TDMAIJJ_CYC(A,B,C,R,JS,JSZ,IS,ISZ)     (src/TDMAIJJ_*.f90:67)
      march "IJI": TDMAIJI_nonCYC(A,B,C,R,BCI,IS,ISZ,JS,JSZ) (BCI(JSZ,2)) or
                   TDMAIJI_CYC(A,B,C,R,IS,ISZ,JS,JSZ)     (src/TDMAIJI_*.f90:67)
    Returns the solution (R after the call)."""
    R = ref_tdma()
    f = lambda v: np.array(v, dtype=np.float64, order="F", copy=True)
    A, B, Cc, X = f(a), f(b), f(c), f(r)
    isz, jsz = A.shape
    ip = lambda v: C.byref(C.c_int(v))
    IS, JS = lo
    if march == "IJJ":
        if cyclic:
            R.tdmaijj_cyc_(_ptr(A), _ptr(B), _ptr(Cc), _ptr(X), ip(JS), ip(jsz), ip(IS), ip(isz))
        else:
            BC = f(bc if bc is not None else np.zeros((isz, 2)))
            R.tdmaijj_noncyc_(_ptr(A), _ptr(B), _ptr(Cc), _ptr(X), _ptr(BC), ip(JS), ip(jsz), ip(IS), ip(isz))
    else:
        if cyclic:
            R.tdmaiji_cyc_(_ptr(A), _ptr(B), _ptr(Cc), _ptr(X), ip(IS), ip(isz), ip(JS), ip(jsz))
        else:
            BC = f(bc if bc is not None else np.zeros((jsz, 2)))
            R.tdmaiji_noncyc_(_ptr(A), _ptr(B), _ptr(Cc), _ptr(X), _ptr(BC), ip(IS), ip(isz), ip(JS), ip(jsz))
    return X


def mean_xz_io(og: "OracleGrid", st: dict, thermal: int = 1):
    """PP_MEAN_ZX_FLOW_Xperiodic_IO moments of the io state dict (keys q, g,
    pr, T, rho, h, mu in the reference layouts): (M, 37), columns as
    orc_mean_xz_io documents them."""
    out = np.zeros(37 * og.M)
    f = lambda k: _ptr(np.asfortranarray(st[k], dtype=np.float64))
    keep = {k: np.asfortranarray(st[k], dtype=np.float64) for k in ("q", "g", "pr", "T", "rho", "h", "mu")}
    lib().orc_mean_xz_io(C.byref(og.g), *[_ptr(keep[k]) for k in ("q", "g", "pr", "T", "rho", "h", "mu")],
                         int(thermal), _ptr(out))
    return out.reshape(og.M, 37)


def mean_t_update(acc, val, nstatis, tp=1, pp_instn_sz=1):
    """PP_MEAN_T_TG / PP_MEAN_T_Xperiodic_IO (src/processing/PP_MEAN_T.f90:322-370,
    436-470): NSTATIS += 1, then acc = COE3 acc + COE2 val with
    (COE2, COE3) = (1/NSTATIS, (NSTATIS-1)/NSTATIS) for TP = 1 (asymptotic)
    or (1/pp_instn_sz, 1) for TP = 2 (summing).  Returns (acc, NSTATIS)."""
    nstatis += 1
    if tp == 1:
        coe2, coe3 = 1.0 / float(nstatis), float(nstatis - 1) / float(nstatis)
    elif tp == 2:
        coe2, coe3 = 1.0 / float(pp_instn_sz), 1.0
    else:
        raise ValueError("TP must be 1 or 2")
    return coe3 * acc + coe2 * val, nstatis
