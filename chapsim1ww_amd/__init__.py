"""MI355X-native per-RK3-substep solver core of CHAPSim (tg channel).

Device work runs in the HIP library libchapsim_gpu.so (C ABI in
include/chapsim_gpu.h); this package holds the host-side set-up (mesh,
coefficients, initial fields) and a thin binding that mirrors the reference's
routine names.
"""
from .grid import ChannelGrid, TGAM, TROH, TALP  # noqa: F401
from . import fields  # noqa: F401

__all__ = ["ChannelGrid", "TGAM", "TROH", "TALP", "fields", "ChannelSolver", "tdma_batched"]


def __getattr__(name):
    # the device binding is imported lazily so that host-only users (and the
    # CPU test suite) do not need a GPU; it raises if the library is missing
    if name in ("ChannelSolver", "tdma_batched", "tdmaijj_noncyc", "tdmaiji_noncyc", "tdmaijj_cyc", "tdmaiji_cyc"):
        from . import solver
        return getattr(solver, name)
    raise AttributeError(name)
