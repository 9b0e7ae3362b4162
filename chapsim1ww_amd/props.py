"""Thermophysical-property tables and the device lookup of THERM_PROP_UPDATE.

Host side of the property path (src/THERM_PROP_UPDATE.f90, src/Flow_State.f90):

* ``synthetic_table``: the heated-channel table of this build's workload, 2048
  points over T in [0.8, 1.6] with c_p = 1 + 4 exp(-((T-1.1)/0.05)^2),
  h = int c_p dT, rho = 3 / (3 + 2 tanh((T-1.1)/0.05)), mu = T^0.7,
  kappa = T^0.8, in the reference's table columns P, H, T, D, M, K, CP, B
  (P and B are dummies).
* ``write_table`` / ``read_table``: the reference's text format (a comment
  line, the number of rows, then rows P H T D M K CP B; NIST_TABLE_SCALING,
  src/Flow_State.f90:378-395).
* ``scale_table``: NIST_TABLE_SCALING (src/Flow_State.f90:398-432): the
  reference state at T0 by bisection and linear interpolation
  (BISECTION_NIST_T, :881-912), then H -> (H - H0)/(T0 CP0), T/T0, D/D0,
  M/M0, K/K0, CP/CP0, B/B0.
* ``PropertyTable``: the device handle (chs_props_create) and
  ``therm_prop_update`` on torch device tensors.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import native as nat

COLUMNS = ("P", "H", "T", "D", "M", "K", "CP", "B")


def synthetic_table(n: int = 2048, t_lo: float = 0.8, t_hi: float = 1.6, t_pc: float = 1.1,
                    width: float = 0.05, peak: float = 4.0) -> dict:
    """Dimensional table columns (numpy arrays of length n), sorted by H."""
    T = np.linspace(t_lo, t_hi, n)
    s = (T - t_pc) / width
    cp = 1.0 + peak * np.exp(-s * s)
    # h = int_{t_lo}^{T} c_p dT in closed form
    erf = np.vectorize(math.erf)
    h = (T - t_lo) + peak * width * math.sqrt(math.pi) / 2.0 * (erf(s) - math.erf((t_lo - t_pc) / width))
    rho = 3.0 / (3.0 + 2.0 * np.tanh(s))
    return {"P": np.ones(n), "H": h, "T": T, "D": rho, "M": T ** 0.7, "K": T ** 0.8, "CP": cp, "B": np.ones(n)}


def write_table(path: str, cols: dict) -> None:
    n = len(cols["H"])
    with open(path, "w") as fh:
        fh.write("# P H T D M K CP B\n")
        fh.write(f"{n}\n")
        for i in range(n):
            fh.write(" ".join(f"{cols[c][i]:.17e}" for c in COLUMNS) + "\n")


def read_table(path: str) -> dict:
    with open(path) as fh:
        fh.readline()
        n = int(fh.readline().split()[0])
        rows = np.array([[float(v) for v in fh.readline().split()[:8]] for _ in range(n)])
    return {c: rows[:, i].copy() for i, c in enumerate(COLUMNS)}


def _bisection_t(T: np.ndarray, t0: float):
    """BISECTION_NIST_T: 1-based (IE, IS) bracketing t0."""
    i_s, i_e = 1, len(T)
    while i_e - i_s > 1:
        im = i_s + (i_e - i_s) // 2
        vs = T[i_s - 1] - t0
        vm = T[im - 1] - t0
        if vs * vm > 0.0:
            i_s = im
        else:
            i_e = im
    return i_e, i_s


def scale_table(cols: dict, t0: float = 1.0) -> dict:
    """NIST_TABLE_SCALING: nondimensional columns H, T, D, M, K, CP, B plus the
    reference values H0, D0, K0, M0, CP0, B0."""
    T = cols["T"]
    ie, i_s = _bisection_t(T, t0)
    sfs = (T[ie - 1] - t0) / (T[ie - 1] - T[i_s - 1])
    sfe = 1.0 - sfs
    ref = {k: sfe * cols[k][ie - 1] + sfs * cols[k][i_s - 1] for k in ("H", "D", "K", "M", "CP", "B")}
    out = {
        "H": (cols["H"] - ref["H"]) / t0 / ref["CP"],
        "T": cols["T"] / t0,
        "D": cols["D"] / ref["D"],
        "M": cols["M"] / ref["M"],
        "K": cols["K"] / ref["K"],
        "CP": cols["CP"] / ref["CP"],
        "B": cols["B"] / ref["B"],
    }
    out["ref"] = ref
    return out


def snapshot_enthalpy(u: np.ndarray, scaled: dict, n: int, seed: int = 0, t_lo: float = 0.82,
                      t_hi: float = 1.58, uniform_share: float = 0.5) -> np.ndarray:
    """Enthalpy samples drawn from a channel-solver snapshot: the streamwise
    velocity u of the snapshot becomes a heated-wall temperature field
    T = t_hi - (t_hi - t_lo) u / max u (slow fluid near the walls is hot),
    mapped to h through the scaled table; cells are drawn at random.  The
    snapshot leaves the hot near-wall tail thinly sampled, so a share of the
    n samples (uniform_share) is drawn uniformly over the snapshot's h range."""
    uu = np.asarray(u, np.float64).ravel(order="F")
    umax = np.abs(uu).max()
    T = np.clip(t_hi - (t_hi - t_lo) * np.abs(uu) / umax, t_lo, t_hi)
    rng = np.random.default_rng(seed)
    nu = int(round(n * uniform_share))
    Ts = T[rng.integers(0, T.size, n - nu)]
    hs = np.interp(Ts, scaled["T"], scaled["H"])
    lo, hi = np.interp([T.min(), T.max()], scaled["T"], scaled["H"])
    out = np.concatenate([hs, rng.uniform(lo, hi, nu)])
    return out[rng.permutation(n)]


class PropertyTable:
    """Device property table (chs_props): splines H -> T and T -> D, M, K, CP."""

    def __init__(self, scaled: dict, stream=None):
        self.lib = nat.load()
        cols = [np.ascontiguousarray(scaled[k], dtype=np.float64) for k in ("H", "T", "D", "M", "K", "CP")]
        self._keep = cols
        self.n = len(cols[0])
        self.h_range = (float(cols[0][0]), float(cols[0][-1]))
        p = C.c_void_p()
        s = None if stream is None else C.c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream))
        nat.check(self.lib.chs_props_create(self.n, *[nat.hptr(c) for c in cols], s, C.byref(p)))
        self.p = p

    def set_mode(self, mode: str):
        """The property update inside the energy substep: "table" or
        "surrogate" (the certified network, the table elsewhere)."""
        m = {"table": nat.CHS_PROPS_TABLE, "surrogate": nat.CHS_PROPS_SURROGATE}[mode]
        nat.check(self.lib.chs_props_set_mode(self.p, m))

    def check(self):
        """The range check of the last energy substep with this table: raises
        ChsError (CHS_E_RANGE) if a lookup left the table and the error was
        not reported yet (the energy substep reads its counts one call later)."""
        nat.check(self.lib.chs_props_check(self.p))

    def last_fallback(self) -> int:
        """Cells of the last energy substep the surrogate handed to the table."""
        v = C.c_longlong()
        nat.check(self.lib.chs_props_last_fallback(self.p, C.byref(v)))
        return v.value

    def rhoh_max(self) -> float:
        """RHOHmax of the table (the limiter of the energy update)."""
        v = C.c_double()
        nat.check(self.lib.chs_props_rhoh_max(self.p, C.byref(v)))
        return v.value

    def close(self):
        if getattr(self, "p", None):
            self.lib.chs_props_destroy(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- learned surrogate ---------------------------------------------------
    NPAR = 4613

    def train(self, h_samples, steps: int, batch: int = 4096, lr: float = 3e-3, seed: int = 1) -> float:
        """Train the surrogate on device enthalpy samples (labels from this
        table) with Adam; returns the final loss over all samples."""
        loss = C.c_double()
        nat.check(self.lib.chs_surrogate_train(self.p, nat.dptr(h_samples), h_samples.numel(), int(steps), int(batch),
                                               float(lr), int(seed), C.byref(loss)))
        return loss.value

    def certify(self, tol: float = 1e-4, nbins: int = 4096) -> float:
        """Bins of the trained range where the network is within tol of the
        table (checked on 33 points per bin with a factor-2 margin); returns
        the certified fraction."""
        f = C.c_double()
        nat.check(self.lib.chs_surrogate_certify(self.p, float(tol), int(nbins), C.byref(f)))
        return f.value

    def certified(self):
        """(fraction of bins certified for all five outputs, for T, rho, mu,
        kappa alone) of the last certification."""
        a, b = C.c_double(), C.c_double()
        nat.check(self.lib.chs_surrogate_certified(self.p, C.byref(a), C.byref(b)))
        return a.value, b.value

    def grad(self, h):
        """Loss and gradient (float32, NPAR) of the surrogate over exactly the
        samples h (device tensor), labels from the table; no update."""
        g = np.zeros(self.NPAR, np.float32)
        loss = C.c_double()
        nat.check(self.lib.chs_surrogate_grad(self.p, nat.dptr(h), h.numel(),
                                              g.ctypes.data_as(C.POINTER(C.c_float)), C.byref(loss)))
        return loss.value, g

    def eval(self, h, out):
        """Raw network outputs (no fallback) into out (device, FP64, 5 x n)."""
        nat.check(self.lib.chs_surrogate_eval(self.p, nat.dptr(h), h.numel(), nat.dptr(out)))

    def params(self):
        par = np.zeros(self.NPAR, np.float32)
        norm = np.zeros(14)
        nat.check(self.lib.chs_surrogate_params(self.p, 0, par.ctypes.data_as(C.POINTER(C.c_float)), self.NPAR,
                                                norm.ctypes.data_as(C.POINTER(C.c_double))))
        return par, norm

    def set_params(self, par, norm):
        par = np.ascontiguousarray(par, np.float32)
        norm = np.ascontiguousarray(norm, np.float64)
        nat.check(self.lib.chs_surrogate_params(self.p, 1, par.ctypes.data_as(C.POINTER(C.c_float)), self.NPAR,
                                                norm.ctypes.data_as(C.POINTER(C.c_double))))

    def update(self, h, T, rho, mu, kap, cp=None, mode: str = "table", rows=None) -> int:
        """THERM_PROP_UPDATE on device tensors.  rows = (nrow, rowlen, stride)
        (default: one flat run of h.numel() cells).  Returns the number of
        cells the surrogate handed to the table."""
        nrow, rowlen, stride = rows if rows is not None else (1, h.numel(), h.numel())
        need = (nrow - 1) * stride + rowlen if nrow > 0 else 0
        for t in (h, T, rho, mu, kap) + (() if cp is None else (cp,)):
            if t.numel() < need:
                raise ValueError("therm_prop_update: array smaller than the rows describe")
        nf = C.c_longlong()
        m = {"table": nat.CHS_PROPS_TABLE, "surrogate": nat.CHS_PROPS_SURROGATE}[mode]
        nat.check(self.lib.chs_therm_prop_update(self.p, m, nat.dptr(h), nrow, rowlen, stride, nat.dptr(T),
                                                 nat.dptr(rho), nat.dptr(mu), nat.dptr(kap),
                                                 None if cp is None else nat.dptr(cp), C.byref(nf)))
        return nf.value
