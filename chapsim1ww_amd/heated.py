"""Host mirror of the heated periodic channel's substep on device-resident
fields: the energy equation (SOLVERRK3_ENG_IO), the property staging and the
io momentum equation (SOLVERRK3_MOM_IO, explicit viscous terms or, with
visimplicit = 1, the RK-CN branch), in the order
SOLVE runs them for each RK3 substep (src/SOLVE.f90:119-133).

    HeatedChannel(solver, table, thermal=1, nflow=1, flowdrvtp=1, gravdir=0,
                  cfgv=0.0, f_a=0.0, visimplicit=0)
    HeatedChannel.upload(fields) / download()
    HeatedChannel.solverrk3_eng_io(ns, dt, cthecd)    SOLVERRK3_ENG_IO(NS)
    HeatedChannel.io_prepare(ns)                      DENSITY_TIME_DERIVATION,
                                                      DENSITY_Staggered,
                                                      MU_Staggered, VELOCITY_CALC_io
    HeatedChannel.solverrk3_mom_io(ns, dt)            SOLVERRK3_MOM_IO(NS)
    HeatedChannel.substep(ns, dt, cthecd)             the three in SOLVE's order
    HeatedChannel.divgck_io()                         DIVGCK_io
    HeatedChannel.mean_xz_io()                        PP_MEAN_ZX_FLOW_Xperiodic_IO

Fields use the reference layouts (Fortran order): the thermal fields of
energy.ThermalState plus Q_io, PR_io, DrhoDtP, DivU_io (NCL1, 0:NCL2+1,
NCL3[, 3]), RHO_STG, MU_STG, DRHOI_STG (NCL1, 0:NCL2+1, NCL3, 3), EXPLT0_io
(NCL1, NCL2, NCL3, 3).  G_io and G0_io are shared by both equations.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import native as nat
from .energy import NAMES as THERMAL_NAMES
from .energy import ThermalState

IO_NAMES = ("q", "pr", "drhodtp", "divu", "rho_stg", "mu_stg", "explt0", "drhoi_stg")


class HeatedChannel(ThermalState):
    """Device fields of the heated channel on the grid of a ChannelSolver, with
    the property table (props.PropertyTable) that closes them."""

    def __init__(self, solver, table, thermal=1, nflow=1, flowdrvtp=1, gravdir=0, cfgv=0.0, f_a=0.0,
                 visimplicit=0):
        import torch
        super().__init__(solver, table)
        g = solver.grid
        L, M, N = g.L, g.M, g.N
        for k in ("q", "rho_stg", "mu_stg", "drhoi_stg"):
            self.shapes[k] = (L, M + 2, N, 3)
        for k in ("pr", "drhodtp", "divu"):
            self.shapes[k] = (L, M + 2, N)
        self.shapes["explt0"] = (L, M, N, 3)
        for k in IO_NAMES:
            self.dev[k] = torch.zeros(int(np.prod(self.shapes[k])), dtype=torch.float64, device="cuda")
        self.dev["rho_stg"].fill_(1.0)
        self.dev["mu_stg"].fill_(1.0)
        self.params = nat.IoParams(int(thermal), int(nflow), int(flowdrvtp), int(gravdir), float(cfgv), float(f_a),
                                   int(visimplicit))

    def _tf(self):
        return nat.ThermalFields(*[nat.dptr(self.dev[k]) for k in THERMAL_NAMES])

    def _iof(self):
        return nat.IoFields(*[nat.dptr(self.dev[k]) for k in ("q", "g", "g0", "pr", "drhodtp", "divu", "rho_stg",
                                                               "mu_stg", "explt0", "drhoi_stg")])

    def io_prepare(self, ns: int):
        tf, f = self._tf(), self._iof()
        nat.check(self.solver.lib.chs_io_prepare(self.solver.h, C.byref(tf), C.byref(f), C.byref(self.params),
                                                 int(ns)))

    def solverrk3_mom_io(self, ns: int, dt: float):
        tf, f = self._tf(), self._iof()
        nat.check(self.solver.lib.chs_solverrk3_mom_io(self.solver.h, C.byref(tf), C.byref(f), C.byref(self.params),
                                                       int(ns), float(dt)))

    def substep(self, ns: int, dt: float, cthecd: float):
        """One RK3 substep of the heated channel in SOLVE's order."""
        if self.params.thermal:
            self.solverrk3_eng_io(ns, dt, cthecd)
        self.io_prepare(ns)
        self.solverrk3_mom_io(ns, dt)

    def mean_xz_io(self):
        """PP_MEAN_ZX_FLOW_Xperiodic_IO moments of the device fields, (NCL2, 37):
        <u v w p>, <g_m>, <u_m p>, <u_m u_n>, <u_m g_n>, T1 D1 H1 M1 T2 D2 H2
        DH PH, <h u_m>, <h g_m> (columns as chs_mean_xz_io documents them)."""
        out = np.zeros(37 * self.solver.grid.M)
        tf, f = self._tf(), self._iof()
        nat.check(self.solver.lib.chs_mean_xz_io(self.solver.h, C.byref(tf), C.byref(f), int(self.params.thermal),
                                                 out.ctypes.data_as(C.POINTER(C.c_double))))
        return out.reshape(self.solver.grid.M, 37)

    def divgck_io(self):
        """(max|div G + DrhoDtP|, max|div Q|)."""
        a, b = C.c_double(), C.c_double()
        f = self._iof()
        nat.check(self.solver.lib.chs_divgck_io(self.solver.h, C.byref(f), C.byref(a), C.byref(b)))
        return a.value, b.value
