"""Host-side initial fields and I/O helpers.

The reference seeds its turbulence generator with the compiler's
``RANDOM_NUMBER`` stream, re-seeded per cell (src/INITIAL_FL_FLD_TG.f90:204-231,
src/processing/randomgen.f90:57-104); that stream differs between Fortran
compilers, so it cannot be reproduced.  The same recipe is applied here with a
portable counter-based generator, so the device solver, the CPU oracle and any
host driver all start from the same bits:

1. u, v, w = VPER * U(-1, 1) (times SVPER in the quarter next to each wall),
   v = 0 on the bottom wall face          (RANDOM_FL_FLD_TG :1-55, INIFIELD_FLOW_tg)
2. subtract the xz-mean of every component (INITIAL_MEAN_TG)
3. u += 1.5 (1 - y^2)                      (LAMPOISLPROF, src/LAMPOISLPROF.f90:39-52)
4. v = 0 on the wall faces, u /= bulk velocity (BULK_VELOCITY_TG)
5. p = 0, wall ghost values from BC_WALL_Q_tg / BC_WALL_PR_tg.
"""
from __future__ import annotations

import numpy as np

from .grid import ChannelGrid

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser (Steele, Lea, Flood 2014) on uint64 arrays."""
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform_pm1(seed: int, shape, stream: int = 0) -> np.ndarray:
    """U(-1, 1) in Fortran order, value (i, j, k, ...) = hash(seed, stream, linear index)."""
    n = int(np.prod(shape))
    idx = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = _splitmix64(np.uint64(seed) * np.uint64(0x100000001B3) + np.uint64(stream))
        z = _splitmix64(idx ^ key)
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)  # [0, 1)
    return (2.0 * u - 1.0).reshape(shape, order="F")


def bulk_velocity(g: ChannelGrid, q: np.ndarray) -> float:
    """BULK_VELOCITY_TG (src/INITIAL_FL_FLD_TG.f90:136-163): volume-weighted mean of u."""
    u = q[:, 1:g.M + 1, :, 0]
    s = float(np.sum(u * (1.0 / g.dyfi)[None, :, None]))
    s /= g.dzi
    area = float(np.sum(1.0 / g.dyfi)) * g.N / g.dzi
    return s / (area * g.L)


def apply_wall_bc(g: ChannelGrid, q: np.ndarray = None, pr: np.ndarray = None) -> None:
    """BC_WALL_Q_tg and BC_WALL_PR_tg (src/BC_WALL.f90:185-305) in place."""
    M = g.M
    if q is not None:
        q[:, 0, :, :] = 0.0
        q[:, M + 1, :, :] = 0.0
        q[:, 1, :, 1] = 0.0
    if pr is not None:
        pr[:, 0, :] = pr[:, 1, :]
        pr[:, M + 1, :] = pr[:, M, :]


def init_tg_fields(g: ChannelGrid, seed: int = 1973, vper: float = 0.1, svper: float = 1.0):
    """Seeded laminar-plus-noise channel state: returns (Q, PR, CONVH0) in the
    reference layouts Q_tg(L, 0:M+1, N, 3), PR_tg(L, 0:M+1, N), CONVH0_tg(L, M, N, 3)."""
    L, M, N = g.L, g.M, g.N
    q = g.zeros(g.shape_q)
    amp = np.where((1.0 - np.abs(g.ycc)) < 0.25, vper * svper, vper)
    for c in range(3):
        r = uniform_pm1(seed, (L, M, N), stream=c + 1)
        q[:, 1:M + 1, :, c] = amp[None, :, None] * r
    q[:, 1, :, 1] = 0.0
    for c in range(3):
        mean = q[:, 1:M + 1, :, c].mean(axis=(0, 2))
        q[:, 1:M + 1, :, c] -= mean[None, :, None]
    vini = np.where((g.ycc < 1.0) & (g.ycc > -1.0), (1.0 - g.ycc ** 2) * 1.5, 0.0)
    q[:, 1:M + 1, :, 0] += vini[None, :, None]
    q[:, 1, :, 1] = 0.0
    ub = bulk_velocity(g, q)
    q[:, 1:M + 1, :, 0] /= ub
    pr = g.zeros(g.shape_ghost)
    apply_wall_bc(g, q, pr)
    convh0 = g.zeros(g.shape_convh0)
    return q, pr, convh0


def poisson_rhs(g: ChannelGrid, seed: int = 7) -> np.ndarray:
    """Synthetic Poisson right-hand side: N(0,1) per cell (Box-Muller on the
    counter generator) with its volume-weighted mean removed, so the (0,0)
    Neumann mode is compatible (sum f / DYFI = 0)."""
    shape = g.shape_cell
    u1 = 0.5 * (uniform_pm1(seed, shape, stream=11) + 1.0)
    u2 = 0.5 * (uniform_pm1(seed, shape, stream=12) + 1.0)
    f = np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)
    w = (1.0 / g.dyfi)[None, :, None]
    mean = float(np.sum(f * w) / (np.sum(w) * g.L * g.N))
    f -= mean
    return np.asfortranarray(f)


# ---------------------------------------------------------------------------
# Instant-field files DNS_perixz_INSTANT_T<time>_{U,V,W,P}.D
# (WRT_INSTANT_VARS_TG, src/processing/WRT_VARS_TG.f90:1-91; read back by
# RESTART_INSTANT_VARS_TG and READ_3D_VARS, src/processing/RESTART_tg.f90:1-46,
# src/processing/READ_INTP_3D_VARS.f90:1-90): a 40-byte header of 4 int32
# (NCL1, NCL2, NCL3, ITERG) and 3 float64 (phyTIME, REN, DT), then the
# interior field (NCL1, NCL2, NCL3) as Fortran-order float64, native (little)
# endian.  The time in the name is Fortran's ES15.9 edit of phyTIME.
# ---------------------------------------------------------------------------
_HDR = np.dtype([("n", "<i4", 4), ("r", "<f8", 3)])
INSTANT_DIR = "1_instant_D/"          # filepath1 (src/modules.f90:125)
_COMPONENTS = ("U", "V", "W", "P")    # NDV + 1 files, in the reference's order


def es15_9(x: float) -> str:
    """Fortran's ES15.9 edit of a time: d.dddddddddE+ee (15 characters), as
    WRITE(PNTIM,'(1ES15.9)') phyTIME produces for 0 <= t < 1e100.  A negative
    time or one of 1e100 or more does not fit 15 characters (Fortran writes
    asterisks), so it is refused."""
    x = float(x)
    if not np.isfinite(x) or x < 0.0 or x >= 1e100:
        raise ValueError(f"ES15.9 cannot represent the time {x!r} in 15 characters")
    if x == 0.0:
        return "0.000000000E+00"
    mant, exp = f"{x:.9E}".split("E")
    e = int(exp)
    if e < -99:
        raise ValueError(f"ES15.9 exponent of {x!r} needs three digits")
    return f"{mant}E{'+' if e >= 0 else '-'}{abs(e):02d}"


def instant_filename(component: str, phytime: float, filepath: str = INSTANT_DIR) -> str:
    """TRIM(filepath1)//'DNS_perixz_INSTANT_T'//TRIM(PNTIM)//'_<U|V|W|P>.D'."""
    if component not in _COMPONENTS:
        raise ValueError(f"component must be one of {_COMPONENTS}")
    return f"{filepath}DNS_perixz_INSTANT_T{es15_9(phytime)}_{component}.D"


def write_instant(path: str, field: np.ndarray, iterg: int, time: float, ren: float, dt: float) -> None:
    """One instant file: header then the (NCL1, NCL2, NCL3) interior field."""
    L, M, N = field.shape
    hdr = np.zeros((), dtype=_HDR)
    hdr["n"] = (L, M, N, iterg)
    hdr["r"] = (time, ren, dt)
    with open(path, "wb") as fh:
        fh.write(hdr.tobytes())
        fh.write(np.asfortranarray(field, dtype="<f8").tobytes(order="F"))


def read_instant(path: str):
    """Returns (field (NCL1, NCL2, NCL3), ITERG, phyTIME, REN, DT) of one file."""
    with open(path, "rb") as fh:
        raw = fh.read(_HDR.itemsize)
        if len(raw) != _HDR.itemsize:
            raise ValueError(f"{path}: shorter than the 40-byte header")
        hdr = np.frombuffer(raw, dtype=_HDR)[0]
        L, M, N, it = (int(x) for x in hdr["n"])
        if min(L, M, N) <= 0:
            raise ValueError(f"{path}: bad dimensions {L}x{M}x{N} in the header")
        data = np.frombuffer(fh.read(8 * L * M * N), dtype="<f8")
    if data.size != L * M * N:
        raise ValueError(f"{path}: {data.size} values after the header, expected {L * M * N}")
    field = data.reshape((L, M, N), order="F").copy(order="F")
    t, ren, dt = (float(x) for x in hdr["r"])
    return field, it, t, ren, dt


def write_instant_vars_tg(q: np.ndarray, pr: np.ndarray, iterg: int, phytime: float, ren: float, dt: float,
                          filepath: str = INSTANT_DIR):
    """WRT_INSTANT_VARS_TG: the four files of the tg state Q_tg(:, 0:M+1, :, 1:3)
    and PR_tg(:, 0:M+1, :) (interior rows j = 1..M written).  Returns the paths."""
    M = q.shape[1] - 2
    paths = []
    for c, comp in enumerate(_COMPONENTS):
        fld = q[:, 1:M + 1, :, c] if c < 3 else pr[:, 1:M + 1, :]
        path = instant_filename(comp, phytime, filepath)
        write_instant(path, fld, iterg, phytime, ren, dt)
        paths.append(path)
    return paths


def restart_instant_vars_tg(g: ChannelGrid, phytime: float, filepath: str = INSTANT_DIR):
    """RESTART_INSTANT_VARS_TG(TFM) followed by FLOWSTART's wall conditions
    (src/FLOWSTART.f90:27-33): reads the four files of time phytime into
    Q_tg and PR_tg interiors, then BC_WALL_Q_tg / BC_WALL_PR_tg fill the ghost
    rows.  Returns (Q, PR, ITERG, phyTIME).  The files' dimensions must be the
    grid's (the reference reads them with the grid's sizes)."""
    q = g.zeros(g.shape_q)
    pr = g.zeros(g.shape_ghost)
    M = g.M
    iterg, t = None, None
    for c, comp in enumerate(_COMPONENTS):
        path = instant_filename(comp, phytime, filepath)
        fld, it, t, _, _ = read_instant(path)
        if fld.shape != (g.L, g.M, g.N):
            raise ValueError(f"{path}: field {fld.shape} does not match the grid {(g.L, g.M, g.N)}")
        iterg = it
        if c < 3:
            q[:, 1:M + 1, :, c] = fld
        else:
            pr[:, 1:M + 1, :] = fld
    apply_wall_bc(g, q, pr)
    return q, pr, iterg, t
