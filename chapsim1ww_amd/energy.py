"""Host mirror of the heated channel's energy substep, SOLVERRK3_ENG_IO(NS)
(src/SOLVERRK3_ENG_IO.f90:1-77), on device-resident thermal fields.

    ThermalState(solver, table)            device arrays of the thermal state
    ThermalState.upload(fields) / download()
    ThermalState.solverrk3_eng_io(ns, dt, cthecd)

Fields use the reference layouts (Fortran order): RHOH, ENTHALPY,
TEMPERATURE, DENSITY, DENSITY0, DENSITY1, VISCOUSITY, VISCOUSITY0,
THERMCONDT (NCL1, 0:NCL2+1, NCL3); RHS_ENERGY0 (NCL1, NCL2, NCL3); G_io and
G0_io (NCL1, 0:NCL2+1, NCL3, 3); WALLFLUX (NCL1, 2).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import native as nat

NAMES = ("rhoh", "enthalpy", "temperature", "density", "density0", "density1", "viscosity", "viscosity0",
         "thermcondt", "rhs_energy0", "g", "g0", "wallflux")


class ThermalState:
    """Device thermal fields on the grid of a ChannelSolver, with the
    property table (props.PropertyTable) that closes them."""

    def __init__(self, solver, table):
        import torch
        self.solver, self.table = solver, table
        g = solver.grid
        L, M, N = g.L, g.M, g.N
        self.shapes = {k: (L, M + 2, N) for k in NAMES[:9]}
        self.shapes["rhs_energy0"] = (L, M, N)
        self.shapes["g"] = self.shapes["g0"] = (L, M + 2, N, 3)
        self.shapes["wallflux"] = (L, 2)
        self.dev = {k: torch.zeros(int(np.prod(s)), dtype=torch.float64, device="cuda") for k, s in self.shapes.items()}

    def upload(self, fields: dict):
        import torch
        for k, v in fields.items():
            a = np.asfortranarray(v, dtype=np.float64)
            if a.shape != self.shapes[k]:
                raise ValueError(f"{k}: shape {a.shape} != {self.shapes[k]}")
            self.dev[k].copy_(torch.from_numpy(a.ravel(order="F").copy()))

    def download(self) -> dict:
        return {k: t.cpu().numpy().reshape(self.shapes[k], order="F") for k, t in self.dev.items()}

    def solverrk3_eng_io(self, ns: int, dt: float, cthecd: float):
        """SOLVERRK3_ENG_IO(NS) on the device, queued on the solver's stream.
        A property lookup outside the table is reported by the next call with
        the same table (before it does any work) or by table.check()."""
        f = nat.ThermalFields(*[nat.dptr(self.dev[k]) for k in NAMES])
        nat.check(self.solver.lib.chs_energy_substep(self.solver.h, self.table.p, C.byref(f), int(ns), float(dt),
                                                     float(cthecd)))
