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
# (src/processing/WRT_VARS_TG.f90:1-91, src/processing/RESTART_tg.f90:1-46):
# header 4 x int32 (NCL1, NCL2, NCL3, ITERG) + 3 x float64 (time, Re, dt), then
# the interior field (NCL1, NCL2, NCL3) as Fortran-order float64.
# ---------------------------------------------------------------------------
_HDR = np.dtype([("n", "<i4", 4), ("r", "<f8", 3)])


def write_instant(path: str, field: np.ndarray, iterg: int, time: float, ren: float, dt: float) -> None:
    L, M, N = field.shape
    hdr = np.zeros((), dtype=_HDR)
    hdr["n"] = (L, M, N, iterg)
    hdr["r"] = (time, ren, dt)
    with open(path, "wb") as fh:
        fh.write(hdr.tobytes())
        fh.write(np.asfortranarray(field, dtype="<f8").tobytes(order="F"))


def read_instant(path: str):
    with open(path, "rb") as fh:
        hdr = np.frombuffer(fh.read(_HDR.itemsize), dtype=_HDR)[0]
        L, M, N, it = (int(x) for x in hdr["n"])
        data = np.frombuffer(fh.read(8 * L * M * N), dtype="<f8")
    field = data.reshape((L, M, N), order="F").copy(order="F")
    t, ren, dt = (float(x) for x in hdr["r"])
    return field, it, t, ren, dt
