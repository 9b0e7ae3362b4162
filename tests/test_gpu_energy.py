"""Energy substep SOLVERRK3_ENG_IO on the device against the oracle.

The stencil, the RK3 update and the property cascade are evaluated without
fused multiply-adds on both sides, so every field must agree bit for bit."""
import numpy as np
import pytest

import oracle as O
from chapsim1ww_amd import props
from chapsim1ww_amd.grid import ChannelGrid

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

CTHECD = 1.0 / 2800.0
MAP = {"rhoh": "rhoh", "h": "enthalpy", "T": "temperature", "rho": "density", "rho0": "density0",
       "rho1": "density1", "mu": "viscosity", "mu0": "viscosity0", "kap": "thermcondt", "rhs0": "rhs_energy0",
       "g": "g", "g0": "g0", "wallflux": "wallflux"}


@pytest.fixture(scope="module")
def scaled():
    return props.scale_table(props.synthetic_table(), t0=1.0)


def test_rhoh_max_equals_oracle(scaled):
    t = props.PropertyTable(scaled)
    assert t.rhoh_max() == O.rhoh_max(scaled)
    t.close()


@pytest.mark.parametrize("dims,mode", [((16, 12, 8), "marching"), ((64, 48, 32), "marching"),
                                       ((64, 48, 32), "rowtile"), ((64, 48, 32), "auto"), ((24, 12, 8), "auto")])
def test_energy_substeps_bitwise(scaled, dims, mode):
    """SOLVERRK3_ENG_IO against the oracle, every field bit for bit: the
    marching RHS kernel (L a power of two), the row-tile one (mode 'rowtile',
    and L = 24)."""
    from chapsim1ww_amd.energy import ThermalState
    from chapsim1ww_amd.solver import ChannelSolver
    print(f"energy {dims} {mode}", flush=True)
    g = ChannelGrid(*dims)
    og = O.OracleGrid(*dims)
    st, tab, n = O.thermal_case(og, scaled, seed=sum(dims))
    s = ChannelSolver(g)
    s.set_cn_solver(mode)
    pt = props.PropertyTable(scaled)
    ts = ThermalState(s, pt)
    ts.upload({MAP[k]: v for k, v in st.items()})
    rmax = O.rhoh_max(scaled)
    for ns in (1, 2, 3):
        assert O.energy_substep(og, st, tab, n, ns, 2e-3, CTHECD, rmax) == 0
        ts.solverrk3_eng_io(ns, 2e-3, CTHECD)
        got = ts.download()
        for k in ("rhoh", "h", "T", "rho", "rho0", "rho1", "mu", "mu0", "kap", "rhs0"):
            a, b = got[MAP[k]], st[k]
            assert np.array_equal(a, b), (ns, k, np.abs(a - b).max())
    pt.close()
    s.close()


def test_energy_range_error_reported_once_without_sync(scaled):
    """A lookup outside the table inside chs_energy_substep is counted on the
    device and read one call later: chs_props_check reports it (once), and so
    does the next energy substep, which then leaves every field untouched."""
    from chapsim1ww_amd.energy import ThermalState
    from chapsim1ww_amd.native import CHS_E_RANGE, ChsError
    from chapsim1ww_amd.solver import ChannelSolver
    dims = (16, 12, 8)
    L, M, N = dims
    print(f"energy range error {dims}", flush=True)
    g = ChannelGrid(*dims)
    og = O.OracleGrid(*dims)
    st, tab, n = O.thermal_case(og, scaled, seed=5)
    s = ChannelSolver(g)
    pt = props.PropertyTable(scaled)
    ts = ThermalState(s, pt)
    ts.upload({MAP[k]: v for k, v in st.items()})
    ts.solverrk3_eng_io(1, 2e-3, CTHECD)
    pt.check()   # in range
    # the interior cell of largest |RHOH|: with a density far below the
    # table's, h = RHOH / rho leaves the table
    rhoh = ts.download()["rhoh"]
    inner = np.abs(rhoh[:, 1:M + 1, :])
    i, j, k = np.unravel_index(np.argmax(inner), inner.shape)
    assert inner[i, j, k] > 1e-3
    q = int(i + L * (j + 1 + (M + 2) * k))
    ts.dev["density"][q] = 1e-12
    ts.solverrk3_eng_io(2, 2e-3, CTHECD)   # returns: the count is read later
    with pytest.raises(ChsError) as ei:
        pt.check()
    assert ei.value.code == CHS_E_RANGE and "1 property lookups" in str(ei.value), str(ei.value)
    pt.check()   # reported once
    ts.dev["density"][q] = 1e-12
    ts.solverrk3_eng_io(3, 2e-3, CTHECD)
    torch.cuda.synchronize()
    before = ts.download()
    with pytest.raises(ChsError) as ei:
        ts.solverrk3_eng_io(1, 2e-3, CTHECD)
    assert ei.value.code == CHS_E_RANGE
    after = ts.download()
    for key in before:
        assert np.array_equal(before[key], after[key]), key
    pt.close()
    s.close()
