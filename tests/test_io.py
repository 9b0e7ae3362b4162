"""The io momentum substep of the heated channel (SOLVERRK3_MOM_IO with
explicit viscous terms) and the property staging before it, in the oracle.

Pinned by what the discretisation fixes independently of the restatement:
closed forms from rest (gravity and the pressure-gradient drive), the
discrete Laplacian for parallel flows, the tg convection operator (its own
pinned restatement) when rho = 1, conservation of the bulk mass flux under
the constant-flux drive, and the projection (div G + DrhoDtP = 0)."""
import numpy as np
import pytest

import oracle as O
from chapsim1ww_amd import props

TGAM = (1.0, 8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0)
TROH = (0.0, 0.0, -17.0 / 60.0, -5.0 / 12.0)
CTHECD = 1.0 / 2800.0


@pytest.fixture(scope="module")
def scaled():
    return props.scale_table(props.synthetic_table(), t0=1.0)


def _uniform_case(og, rho=1.0, mu=1.0):
    """Quiescent io state with uniform properties (ghost rows included)."""
    L, M, N = og.L, og.M, og.N
    f = lambda a: np.asfortranarray(a, dtype=np.float64)
    sh, sh3 = (L, M + 2, N), (L, M + 2, N, 3)
    st = {"g": f(np.zeros(sh3)), "g0": f(np.zeros(sh3)), "rho": f(np.full(sh, rho)), "rho0": f(np.full(sh, rho)),
          "mu": f(np.full(sh, mu)), "mu0": f(np.full(sh, mu)), "q": f(np.zeros(sh3)), "pr": f(np.zeros(sh)),
          "drhodtp": f(np.zeros(sh)), "divu": f(np.zeros(sh)), "rho_stg": f(np.ones(sh3)),
          "mu_stg": f(np.ones(sh3)), "explt0": f(np.zeros((L, M, N, 3))), "drhoi_stg": f(np.zeros(sh3))}
    return st


def _lap_y(og, u):
    """CV-free discrete d2/dy2 of a cell-centred (L, M, N) field with zero
    wall ghosts: DYFI(j) [(u(j+1)-u(j)) DYCI(j+1) - (u(j)-u(j-1)) DYCI(j)]."""
    up = np.concatenate([np.zeros_like(u[:, :1]), u, np.zeros_like(u[:, :1])], axis=1)
    dyci = og.dyci
    fl = (up[:, 1:, :] - up[:, :-1, :]) * dyci[None, :, None]          # faces 1..M+1
    return og.dyfi[None, :, None] * (fl[:, 1:, :] - fl[:, :-1, :])


def _lap_per(u, axis, di):
    return (np.roll(u, -1, axis) - 2 * u + np.roll(u, 1, axis)) * di * di


@pytest.mark.parametrize("gravdir", [1, 3])
def test_rest_with_gravity_and_pressure_drive(gravdir):
    """From rest with uniform density: RHS = rho alpha dt F_A along the
    gravity direction plus alpha dt CFGV / 2 along x (DPGRNS = -CFGV alpha dt
    / 2); the field stays divergence-free, so phi = 0 and p is unchanged."""
    og = O.OracleGrid(16, 12, 8)
    st = _uniform_case(og, rho=1.25)
    io = O.OracleIO(og, st, thermal=1, nflow=1, flowdrvtp=2, gravdir=gravdir, cfgv=0.3, f_a=-0.7)
    dt, ns = 1e-2, 1
    io.prepare(ns)
    np.testing.assert_allclose(st["rho_stg"], 1.25, rtol=0, atol=1e-15)
    io.substep(ns, dt)
    a = (TGAM[ns] + TROH[ns]) * dt
    gx = 0.5 * 0.3 * a + (1.25 * -0.7 * a if gravdir == 1 else 0.0)
    gz = 1.25 * -0.7 * a if gravdir == 3 else 0.0
    G = st["g"]
    np.testing.assert_allclose(G[:, 1:-1, :, 0], gx, rtol=1e-14, atol=1e-17)
    np.testing.assert_allclose(G[:, 1:-1, :, 2], gz, rtol=1e-14, atol=1e-17)
    assert np.abs(G[:, :, :, 1]).max() == 0.0
    assert np.abs(G[:, 0, :, :]).max() == 0.0 and np.abs(G[:, -1, :, :]).max() == 0.0
    assert np.abs(st["pr"]).max() == 0.0
    assert io.dpgrns == pytest.approx(-0.5 * 0.3 * a, rel=1e-15)


@pytest.mark.parametrize("comp", ["u", "w"])
def test_parallel_flows_follow_the_discrete_laplacian(comp):
    """u(y, z) with v = w = 0, or w(x, y) with u = v = 0, is divergence-free
    and carries no convection; with rho = mu = 1 the substep is
    G_new = G + gamma dt CV (d2/dy2 + d2/dx_j2) G - DPGRNS (x only)."""
    og = O.OracleGrid(16, 12, 8, ren=500.0)
    L, M, N = og.L, og.M, og.N
    rng = np.random.default_rng(7)
    y = np.asarray(og.ycc)
    z = np.arange(N) / N * 2 * np.pi
    x = np.arange(L) / L * 2 * np.pi
    st = _uniform_case(og)
    if comp == "u":
        f = (1 - y[None, :, None] ** 2) * (1 + 0.3 * np.cos(z)[None, None, :]) * np.ones((L, 1, 1))
        f = f + 0.01 * rng.standard_normal((1, M, N))      # any y-z dependence, uniform in x
        st["g"][:, 1:-1, :, 0] = f
    else:
        f = 0.2 * (1 - y[None, :, None] ** 2) * np.sin(2 * x)[:, None, None] * np.ones((1, 1, N))
        f = f + 0.01 * rng.standard_normal((L, M, 1))      # any x-y dependence, uniform in z
        st["g"][:, 1:-1, :, 2] = f
    st["g0"][...] = st["g"]
    io = O.OracleIO(og, st, thermal=1, nflow=1, flowdrvtp=1)
    dt, ns = 2e-3, 1
    io.prepare(ns)
    io.substep(ns, dt)
    cv = 1.0 / 500.0
    G = st["g"]
    if comp == "u":
        r = TGAM[ns] * dt * cv * (_lap_y(og, f) + _lap_per(f, 2, og.dzi))
        dpg = (r / og.dyfi[None, :, None]).sum() / og.volm
        np.testing.assert_allclose(G[:, 1:-1, :, 0], f + r - dpg, rtol=1e-13, atol=1e-15)
        assert np.abs(G[:, :, :, 2]).max() == 0.0
        assert io.dpgrns == pytest.approx(dpg, rel=1e-12)
    else:
        r = TGAM[ns] * dt * cv * (_lap_y(og, f) + _lap_per(f, 0, og.dxi))
        np.testing.assert_allclose(G[:, 1:-1, :, 2], f + r, rtol=1e-13, atol=1e-15)
        assert np.abs(G[:, :, :, 0]).max() == 0.0
    assert np.abs(G[:, :, :, 1]).max() == 0.0


def test_convection_equals_tg_operator_when_rho_is_one():
    """With rho = 1 the mass flux is the velocity and the io convection terms
    are the tg ones (CONVECTION_*_tg, pinned by their own tests)."""
    from chapsim1ww_amd import fields
    from chapsim1ww_amd.grid import ChannelGrid
    dims = (16, 12, 8)
    og = O.OracleGrid(*dims, ren=1e300)
    q, pr, c0 = fields.init_tg_fields(ChannelGrid(*dims), seed=5)
    ref = O.OracleState(og, q, pr, c0)
    for name in ("convection_x", "convection_y", "convection_z"):
        ref.stage(name)
    st = _uniform_case(og)
    st["g"][...] = q
    io = O.OracleIO(og, st, thermal=0)
    io.prepare(1)
    np.testing.assert_array_equal(st["q"], q)
    hx, hy, hz = io.terms()
    np.testing.assert_allclose(hx, ref.qtmp, rtol=0, atol=1e-13 * np.abs(ref.qtmp).max())
    np.testing.assert_allclose(hy, ref.dph[:, 1:-1, :], rtol=0, atol=1e-13 * np.abs(ref.dph).max())
    np.testing.assert_allclose(hz, ref.rhsllphi, rtol=0, atol=1e-13 * np.abs(ref.rhsllphi).max())


def test_heated_substeps_conserve_mass_flux_and_project(scaled):
    """Energy substep, staging and io momentum substep in the reference's
    order for NS = 1, 2, 3: the bulk mass flux sum G_x / DYFI is held by the
    constant-flux drive, and div G + DrhoDtP vanishes after each projection."""
    og = O.OracleGrid(16, 12, 8)
    st, tab, n = O.io_case(og, scaled, seed=11)
    io = O.OracleIO(og, st, thermal=1, nflow=1, flowdrvtp=1, gravdir=1, f_a=-0.05)
    rmax = O.rhoh_max(scaled)
    flux = lambda: (st["g"][:, 1:-1, :, 0] / og.dyfi[None, :, None]).sum()
    f0 = None
    for ns in (1, 2, 3):
        assert O.energy_substep(og, st, tab, n, ns, 1e-3, CTHECD, rmax) == 0
        io.prepare(ns)
        if f0 is None:
            f0 = flux()
        io.substep(ns, 1e-3)
        assert abs(flux() - f0) <= 1e-13 * abs(f0)
        dg, _ = io.divgck()
        assert dg < 1e-10
        assert np.abs(st["g"][:, 0]).max() == 0.0 and np.abs(st["g"][:, -1]).max() == 0.0
        assert np.abs(st["g"][:, 1, :, 1]).max() == 0.0


def test_staggered_properties_closed_forms():
    """DENSITY_Staggered / MU_Staggered on fields linear in x: face values
    are the two-point averages (x faces), the four-point average (x-z edge),
    the y-weighted averages, and the wall rows take the ghost-row values."""
    og = O.OracleGrid(8, 6, 4)
    L, M, N = og.L, og.M, og.N
    st = _uniform_case(og)
    i = np.arange(L, dtype=float)
    rho = 1.0 + 0.125 * i[:, None, None] * np.ones((1, M + 2, N))
    st["rho"][...] = rho
    st["rho0"][...] = rho
    st["mu"][...] = 2.0 * rho
    st["mu0"][...] = 2.0 * rho
    io = O.OracleIO(og, st, thermal=1)
    io.prepare(1)
    rim = np.roll(rho, 1, axis=0)
    np.testing.assert_allclose(st["rho_stg"][:, 1:-1, :, 0], 0.5 * (rho + rim)[:, 1:-1], rtol=1e-15)
    np.testing.assert_allclose(st["rho_stg"][:, 1:-1, :, 2], rho[:, 1:-1], rtol=1e-15)
    wsum = np.asarray(og.wff)[1:M] + np.asarray(og.wfb)[1:M]
    np.testing.assert_allclose(st["rho_stg"][:, 2:M + 1, :, 1], rho[:, 2:M + 1] * wsum[None, :, None], rtol=1e-14)
    np.testing.assert_allclose(st["rho_stg"][:, 1, :, 1], rho[:, 0], rtol=1e-15)
    np.testing.assert_allclose(st["mu_stg"][:, 1:-1, :, 1], (rho + rim)[:, 1:-1], rtol=1e-15)
    np.testing.assert_allclose(st["mu_stg"][:, 2:M + 1, :, 0], (rho + rim)[:, 2:M + 1] * wsum[None, :, None],
                               rtol=1e-14)
    np.testing.assert_allclose(st["mu_stg"][:, 1, :, 0], 2.0 * rho[:, 0], rtol=1e-15)
    np.testing.assert_allclose(st["mu_stg"][:, M + 1, :, 2], 2.0 * rho[:, M + 1], rtol=1e-15)


def _ly_matrix(og, comp, F):
    """The y operator of the RK-CN factor for u or w with unit properties:
    1 - F L_y, L_y the TDMA_COEF stencil (AMVR/ACVR/APVR of the tg solver)
    whose wall rows use AMVR1 / APVRN for the ghost-free wall value."""
    M = og.M
    c = 0 if comp == "u" else 2
    am, ac, ap = (np.asarray(og.arr[k]).reshape(3, M + 1)[c] for k in ("amvr", "acvr", "apvr"))
    A = np.zeros((M, M))
    for j in range(M):
        A[j, j] = 1.0 - F * ac[j]
        if j > 0:
            A[j, j - 1] = -F * am[j]
        if j < M - 1:
            A[j, j + 1] = -F * ap[j]
    return A


@pytest.mark.parametrize("comp", ["u", "w"])
def test_cn_parallel_flows_closed_form(comp):
    """RK-CN branch, unit properties, u(y, z) (or w(x, y)) alone: no
    convection, no explicit stress part, no density correction; the right-hand
    side alpha dt CV (d2/dy2 + d2/dz2) u - DPGRNS is factorised as
    (1 - F Ly)^-1 (1 - F Lz)^-1 with F = alpha dt CV / 2, Lz the periodic
    second difference and Ly the tridiagonal wall operator."""
    og = O.OracleGrid(16, 12, 8, ren=500.0)
    L, M, N = og.L, og.M, og.N
    rng = np.random.default_rng(3)
    y = np.asarray(og.ycc)
    st = _uniform_case(og)
    if comp == "u":
        f = (1 - y[None, :, None] ** 2) * np.ones((1, 1, N)) + 0.01 * rng.standard_normal((1, M, N))
        f = f * np.ones((L, 1, 1))
        st["g"][:, 1:-1, :, 0] = f
    else:
        f = 0.1 * (1 - y[None, :, None] ** 2) * np.ones((L, 1, 1)) + 0.01 * rng.standard_normal((L, M, 1))
        f = f * np.ones((1, 1, N))
        st["g"][:, 1:-1, :, 2] = f
    st["g0"][...] = st["g"]
    io = O.OracleIO(og, st, thermal=1, nflow=1, flowdrvtp=1, visimplicit=1)
    dt, ns = 2e-3, 1
    io.prepare(ns)
    assert np.abs(st["drhoi_stg"]).max() == 0.0
    io.substep(ns, dt)
    cv, a = 1.0 / 500.0, TGAM[ns] + TROH[ns]
    F = 0.5 * a * dt * cv
    if comp == "u":
        r = a * dt * cv * (_lap_y(og, f) + _lap_per(f, 2, og.dzi))
        r = r - (r / og.dyfi[None, :, None]).sum() / og.volm
        per, dper, axis = N, og.dzi, 2
    else:
        r = a * dt * cv * (_lap_y(og, f) + _lap_per(f, 0, og.dxi))
        per, dper, axis = L, og.dxi, 0
    # periodic factor (1 - F d2/dx_j2) along the varying periodic direction
    P = np.eye(per) * (1 + 2 * F * dper ** 2) - F * dper ** 2 * (np.roll(np.eye(per), 1, 0) + np.roll(np.eye(per), -1, 0))
    r = np.moveaxis(np.linalg.solve(P, np.moveaxis(r, axis, 0).reshape(per, -1)).reshape(np.moveaxis(r, axis, 0).shape), 0, axis)
    Ay = _ly_matrix(og, comp, F)
    r = np.moveaxis(np.linalg.solve(Ay, np.moveaxis(r, 1, 0).reshape(M, -1)).reshape(np.moveaxis(r, 1, 0).shape), 0, 1)
    G = st["g"]
    c = 0 if comp == "u" else 2
    np.testing.assert_allclose(G[:, 1:-1, :, c], f + r, rtol=1e-12, atol=1e-15)
    assert np.abs(G[:, :, :, 1]).max() == 0.0


def test_cn_density_correction_closed_form():
    """DENSITY_implicit_add with rho = 2, rho0 = 1: RHO_STG = 3/2 and the
    one-substep-old face density 1, so DRHOI_STG = 2/3 - 1 on the x and z
    faces and the interior y faces, 0 on the wall faces."""
    og = O.OracleGrid(8, 6, 4)
    M = og.M
    st = _uniform_case(og)
    st["rho"][...] = 2.0
    st["rho0"][...] = 1.0
    io = O.OracleIO(og, st, thermal=1, visimplicit=1)
    io.prepare(1)
    d = st["drhoi_stg"]
    np.testing.assert_allclose(d[:, :, :, 0], 1.0 / 1.5 - 1.0, rtol=1e-15)
    np.testing.assert_allclose(d[:, :, :, 2], 1.0 / 1.5 - 1.0, rtol=1e-15)
    np.testing.assert_allclose(d[:, 2:M + 1, :, 1], 1.0 / 1.5 - 1.0, rtol=1e-14)
    assert np.abs(d[:, 1, :, 1]).max() == 0.0 and np.abs(d[:, M + 1, :, 1]).max() == 0.0


def test_cn_heated_substeps_project(scaled):
    """The RK-CN branch in SOLVE's order: div G + DrhoDtP = 0 after every
    projection.  (DPGRNS is taken before the factorised solves, whose
    variable coefficients do not preserve the weighted volume sum, so the bulk
    flux drifts at the 1e-11 level in this branch; it is not asserted.)"""
    og = O.OracleGrid(16, 12, 8)
    st, tab, n = O.io_case(og, scaled, seed=12)
    io = O.OracleIO(og, st, thermal=1, nflow=1, flowdrvtp=1, visimplicit=1)
    rmax = O.rhoh_max(scaled)
    flux = lambda: (st["g"][:, 1:-1, :, 0] / og.dyfi[None, :, None]).sum()
    f0 = None
    for ns in (1, 2, 3):
        assert O.energy_substep(og, st, tab, n, ns, 1e-3, CTHECD, rmax) == 0
        io.prepare(ns)
        if f0 is None:
            f0 = flux()
        io.substep(ns, 1e-3)
        assert abs(flux() - f0) <= 1e-9 * abs(f0)
        assert io.divgck()[0] < 1e-10


def test_planted_mistakes_in_the_closed_forms_are_caught():
    """The closed-form checks see plausible slips: the CN wall row without
    the AMVR1 term, and DPGRNS with the wrong sign."""
    og = O.OracleGrid(16, 12, 8, ren=500.0)
    L, M, N = og.L, og.M, og.N
    y = np.asarray(og.ycc)
    f = (1 - y[None, :, None] ** 2) * np.ones((L, 1, N))
    cv, dt, ns = 1.0 / 500.0, 2e-3, 1
    a = TGAM[ns] + TROH[ns]
    F = 0.5 * a * dt * cv
    r = a * dt * cv * _lap_y(og, f)
    r = r - (r / og.dyfi[None, :, None]).sum() / og.volm
    st = _uniform_case(og)
    st["g"][:, 1:-1, :, 0] = f
    st["g0"][...] = st["g"]
    io = O.OracleIO(og, st, thermal=1, nflow=1, flowdrvtp=1, visimplicit=1)
    io.prepare(ns)
    io.substep(ns, dt)
    got = st["g"][:, 1:-1, :, 0]
    solve = lambda A, rr: np.moveaxis(np.linalg.solve(A, np.moveaxis(rr, 1, 0).reshape(M, -1))
                                      .reshape(np.moveaxis(rr, 1, 0).shape), 0, 1)
    Ay = _ly_matrix(og, "u", F)
    assert np.abs(got - (f + solve(Ay, r))).max() < 1e-13       # the right form holds
    Ay_bad = Ay.copy()
    am = np.asarray(og.arr["amvr"]).reshape(3, M + 1)[0]
    ucaj = 4.0 / (2.0 / og.dyci[1] + 1.0 / og.dyfi[0])
    Ay_bad[0, 0] -= F * ucaj * 2.0 * og.dyfi[0]                  # drop the wall term of row 1
    assert am[0] == 0.0
    assert np.abs(got - (f + solve(Ay_bad, r))).max() > 1e-11    # 100x the right form's residual
    # the explicit branch with the mass-flux correction's sign flipped
    st = _uniform_case(og)
    st["g"][:, 1:-1, :, 0] = f
    st["g0"][...] = st["g"]
    io = O.OracleIO(og, st, thermal=1, nflow=1, flowdrvtp=1)
    io.prepare(ns)
    io.substep(ns, dt)
    rx = TGAM[ns] * dt * cv * _lap_y(og, f)
    dpg = (rx / og.dyfi[None, :, None]).sum() / og.volm
    got = st["g"][:, 1:-1, :, 0]
    assert np.abs(got - (f + rx - dpg)).max() < 1e-13
    assert np.abs(got - (f + rx + dpg)).max() > 1e-11
