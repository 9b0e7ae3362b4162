"""Energy substep of the heated channel (SOLVERRK3_ENG_IO) in the oracle,
pinned by the discrete energy balance, a fixed point and RHOHmax."""
import numpy as np
import pytest

import oracle as O
from chapsim1ww_amd import props

CTHECD = 1.0 / 2800.0


@pytest.fixture(scope="module")
def scaled():
    return props.scale_table(props.synthetic_table(), t0=1.0)


def test_energy_balance(scaled):
    """Periodic x, z and no wall-normal mass flux at the walls: the volume sum
    of RHS_ENERGY (weights 1/DYFI) is the net wall heat input
    CTHECD * N * sum_i (q_bottom - q_top)."""
    og = O.OracleGrid(16, 12, 8)
    st, tab, n = O.thermal_case(og, scaled, seed=1)
    st["wallflux"][:, 0] = np.linspace(0.05, 0.15, 16)
    st["wallflux"][:, 1] = 0.02
    bad = O.energy_substep(og, st, tab, n, 1, 1e-3, CTHECD, O.rhoh_max(scaled))
    assert bad == 0
    rhs = st["rhs0"]
    total = (rhs / og.dyfi[None, :, None]).sum()
    expect = CTHECD * og.N * (st["wallflux"][:, 0] - st["wallflux"][:, 1]).sum()
    scale = np.abs(rhs / og.dyfi[None, :, None]).sum()
    assert abs(total - expect) <= 1e-12 * scale


def test_planted_wall_sign_is_caught(scaled):
    """The same balance with the sign of the bottom wall flux flipped in the
    expected value fails (the check sees a sign slip)."""
    og = O.OracleGrid(16, 12, 8)
    st, tab, n = O.thermal_case(og, scaled, seed=2)
    O.energy_substep(og, st, tab, n, 1, 1e-3, CTHECD, O.rhoh_max(scaled))
    total = (st["rhs0"] / og.dyfi[None, :, None]).sum()
    wrong = CTHECD * og.N * (-st["wallflux"][:, 0] - st["wallflux"][:, 1]).sum()
    assert abs(total - wrong) > 1e-6 * abs(wrong)


def test_quiescent_state_is_a_fixed_point(scaled):
    """No flow, uniform temperature, no wall flux: RHOH is unchanged and the
    property cascade reproduces the state."""
    og = O.OracleGrid(8, 8, 4)
    st, tab, n = O.thermal_case(og, scaled, seed=3, qw=0.0)
    Tu = 1.3
    hu = float(np.interp(Tu, scaled["T"], scaled["H"]))
    out = [np.zeros(1) for _ in range(5)]
    O.therm_prop_update(tab, n, np.array([hu]), *out)
    for k, v in (("h", hu), ("T", out[0][0]), ("rho", out[1][0]), ("mu", out[2][0]), ("kap", out[3][0])):
        st[k][...] = v
    st["rhoh"][...] = hu * out[1][0]
    st["g"][...] = 0.0
    st["g0"][...] = 0.0
    st["rhs0"][...] = 0.0
    before = {k: st[k].copy() for k in ("rhoh", "h", "T", "rho", "mu", "kap")}
    assert O.energy_substep(og, st, tab, n, 2, 1e-3, CTHECD, O.rhoh_max(scaled)) == 0
    assert np.array_equal(st["rhoh"][:, 1:-1], before["rhoh"][:, 1:-1])
    for k in ("h", "T", "rho", "mu", "kap"):
        assert np.abs(st[k] - before[k]).max() <= 1e-13 * np.abs(before[k]).max(), k
    assert np.abs(st["rhs0"]).max() == 0.0


def test_rhoh_max_matches_a_dense_search(scaled):
    """RHOHmax against the maximum of rho*h over the table knots (the
    3n-point spline search stops 1/(3n) of the range short of the last knot,
    so the two agree to the variation of rho*h over that last stretch)."""
    r = O.rhoh_max(scaled)
    prod = scaled["D"] * scaled["H"]
    knots = prod.max()
    assert abs(r - knots) <= 1e-3 * abs(knots)
    # the maximum sits at the hot end of this table: below the last knot
    assert np.argmax(prod) == len(prod) - 1 and r < knots
