"""Pins the CPU oracle (oracle/chapsim_oracle.c) before it is trusted as the
checker of the HIP library.

Each oracle function is checked against something that does not share its
code: the reference's own FFT99 package compiled from its sources
(oracle/_ref), numpy library routines, dense linear solves of operators built
from their finite-volume definitions, closed forms and invariants.  The
planted-mistake tests at the end show that a plausible error in the oracle
(a dropped term, a wrong index, a transposed operand) fails these checks.
"""
import math

import numpy as np
import pytest

import oracle as O
from chapsim1ww_amd.grid import ChannelGrid
from chapsim1ww_amd import fields

RNG = np.random.default_rng(1234)


# --------------------------------------------------------------------------
# Fourier transforms: conventions of libs/fft99 (FFT99.f90:98-105, CFFT99.f90:182-197)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("n", [8, 12, 16, 30, 64, 256])
def test_fft99_matches_numpy(n):
    x = RNG.standard_normal(n)
    ab = O.fft99_fwd(x)
    c = np.fft.rfft(x) / n
    assert np.allclose(ab[0::2], c.real, atol=1e-15 * n)
    assert np.allclose(ab[1::2], c.imag, atol=1e-15 * n)
    assert ab[1] == 0.0 and ab[n + 1] == 0.0
    back = O.fft99_inv(ab, n)
    assert np.allclose(back, x, atol=1e-13)
    z = RNG.standard_normal(n) + 1j * RNG.standard_normal(n)
    assert np.allclose(O.cfft99(z, -1), np.fft.fft(z), atol=1e-13 * n)
    assert np.allclose(O.cfft99(z, +1), np.fft.ifft(z) * n, atol=1e-13 * n)


def test_fft99_known_modes():
    n = 16
    j = np.arange(n)
    for k in range(1, n // 2):
        ab = O.fft99_fwd(np.cos(2 * np.pi * k * j / n))
        expect = np.zeros(n + 2)
        expect[2 * k] = 0.5
        assert np.allclose(ab, expect, atol=1e-15)
        ab = O.fft99_fwd(np.sin(2 * np.pi * k * j / n))
        expect = np.zeros(n + 2)
        expect[2 * k + 1] = -0.5
        assert np.allclose(ab, expect, atol=1e-15)


@pytest.mark.parametrize("n", [8, 12, 16, 30, 64, 256])
def test_fft99_matches_reference_package(n):
    # loaded inside the test, so that collection never maps the compiled
    # reference package (it is built only where the reference tree exists)
    if O.ref_fft99() is None:
        pytest.skip("oracle/_ref not built (no reference tree)")
    lot = 3
    x = RNG.standard_normal((n, lot))
    data = np.zeros((n + 2, lot))
    data[1:n + 1] = x
    data[0] = x[n - 1]
    data[n + 1] = x[0]
    ref = O.ref_fft99_transform(data, n, -1)
    mine = np.stack([O.fft99_fwd(x[:, l]) for l in range(lot)], 1)
    assert np.abs(ref - mine).max() < 1e-14
    # inverse, including non-zero b_0 / b_{n/2}, which FFT99 ignores
    ab = RNG.standard_normal((n + 2, lot))
    ref = O.ref_fft99_transform(ab, n, +1)
    mine = np.stack([O.fft99_inv(ab[:, l], n) for l in range(lot)], 1)
    assert np.abs(ref[1:n + 1] - mine).max() < 1e-12
    z = RNG.standard_normal((n, lot)) + 1j * RNG.standard_normal((n, lot))
    for s in (-1, 1):
        ref = O.ref_cfft99_transform(z, s)
        mine = np.stack([O.cfft99(z[:, l], s) for l in range(lot)], 1)
        assert np.abs(ref - mine).max() < 1e-12


# --------------------------------------------------------------------------
# Tridiagonal solvers (src/TDMAIJ{I,J}_{CYC,nonCYC}.f90) against dense solves
# --------------------------------------------------------------------------
def _dense(a, b, c, cyclic):
    n = len(b)
    A = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    if cyclic:
        A[0, n - 1] += a[0]
        A[n - 1, 0] += c[n - 1]
    return A


@pytest.mark.parametrize("shape", [(4, 7), (3, 12), (5, 3), (2, 192)])
@pytest.mark.parametrize("cyclic", [False, True])
@pytest.mark.parametrize("march", ["IJJ", "IJI"])
def test_tdma_dense(shape, cyclic, march):
    isz, jsz = shape
    a = RNG.uniform(-0.9, -0.1, shape)
    c = RNG.uniform(-0.9, -0.1, shape)
    b = 2.0 + np.abs(RNG.standard_normal(shape))
    d = RNG.standard_normal(shape)
    A, B, Cc, R = (np.asfortranarray(v) for v in (a, b, c, d.copy()))
    if march == "IJJ":
        n, nl, es, ls = jsz, isz, isz, 1
    else:
        n, nl, es, ls = isz, jsz, 1, isz
    if n < 3:
        pytest.skip("the reference needs at least 3 rows")
    bc = np.asfortranarray(RNG.standard_normal((nl, 2))) if not cyclic else None
    if cyclic:
        O.tdma_cyc(A, B, Cc, R, n, nl, es, ls)
    else:
        O.tdma_noncyc(A, B, Cc, R, bc, n, nl, es, ls)
    for l in range(nl):
        sl = (l, slice(None)) if march == "IJJ" else (slice(None), l)
        al, bl, cl, dl = a[sl], b[sl], c[sl], d[sl].copy()
        if not cyclic:
            dl[0] -= al[0] * bc[l, 0]
            dl[-1] -= cl[-1] * bc[l, 1]
        x = np.linalg.solve(_dense(al, bl, cl, cyclic), dl)
        assert np.allclose(R[sl], x, rtol=1e-13, atol=1e-14)


# --------------------------------------------------------------------------
# Grid and coefficients: package host code == oracle, plus closed forms
# --------------------------------------------------------------------------
@pytest.mark.parametrize("dims", [(8, 6, 8), (64, 64, 64), (128, 128, 128)])
def test_grid_bitwise_and_closed_forms(dims):
    g = ChannelGrid(*dims)
    og = O.OracleGrid(*dims)
    for k in ("ynd", "ycc", "dyfi", "dyci", "wff", "wfb", "amph", "acph", "apph", "ak1", "ak3"):
        assert np.array_equal(getattr(g, k), og.arr[k]), k
    for k in ("amvr", "acvr", "apvr"):
        assert np.array_equal(getattr(g, k).ravel(), og.arr[k]), k
    assert g.volm == og.volm
    L, M, N = dims
    assert g.ynd[0] == pytest.approx(-1.0, abs=1e-15) and g.ynd[-1] == pytest.approx(1.0, abs=1e-15)
    assert np.sum(1.0 / g.dyfi) == pytest.approx(2.0, abs=1e-13)
    # Neumann Poisson operator: every row sums to zero (constants in the kernel)
    assert np.allclose(g.amph + g.acph + g.apph, 0.0, atol=1e-9 * np.abs(g.acph).max())
    # flux form: column sums weighted by 1/DYFI vanish (discrete Gauss theorem)
    Ly = np.diag(g.acph) + np.diag(g.amph[1:], -1) + np.diag(g.apph[:-1], 1)
    assert np.abs((1.0 / g.dyfi) @ Ly).max() < 1e-9 * np.abs(g.acph).max()
    # modified wavenumbers are the exact eigenvalues of the periodic 2nd difference
    k = np.arange(L)
    assert np.allclose(g.ak1, 4.0 * np.sin(np.pi * k / L) ** 2 * g.dxqi, rtol=1e-12, atol=1e-12)
    k = np.arange(N)
    assert np.allclose(g.ak3, 4.0 * np.sin(np.pi * k / N) ** 2 * g.dzqi, rtol=1e-12, atol=1e-12)
    # volume of the box
    # (sequential summation of L*M*N terms, as the reference does: ~n*eps)
    assert g.volm / (g.dxi * g.dzi) == pytest.approx(2 * math.pi * math.pi * 2.0, rel=1e-10)


# --------------------------------------------------------------------------
# Poisson solve (FFT99_POIS3D_periodicxz)
# --------------------------------------------------------------------------
def _dense_poisson(g, f):
    """Discrete operator built from its finite-volume definition (x, z periodic
    second differences; y flux form with zero-gradient walls), solved densely;
    the free constant fixed by the reference's pin: xz-mean of phi at j=1 is 0."""
    L, M, N = g.L, g.M, g.N
    n = L * M * N
    idx = lambda i, j, k: (i % L) + L * (j + M * (k % N))
    A = np.zeros((n, n))
    dy = np.diff(g.ynd)                 # cell widths
    # centre-to-centre distance across node J+1: half the span of the two
    # neighbouring cells (DYCI's definition, src/COORDJR.f90:116-118)
    dc = 0.5 * (g.ynd[2:] - g.ynd[:-2])
    for k in range(N):
        for j in range(M):
            for i in range(L):
                r = idx(i, j, k)
                for di, dk, w in ((1, 0, g.dxqi), (-1, 0, g.dxqi), (0, 1, g.dzqi), (0, -1, g.dzqi)):
                    A[r, idx(i + di, j, k + dk)] += w
                    A[r, r] -= w
                if j + 1 < M:   # flux through the upper face
                    w = 1.0 / dy[j] / dc[j]
                    A[r, idx(i, j + 1, k)] += w
                    A[r, r] -= w
                if j > 0:
                    w = 1.0 / dy[j] / dc[j - 1]
                    A[r, idx(i, j - 1, k)] += w
                    A[r, r] -= w
    phi = np.linalg.lstsq(A, f.ravel(order="F"), rcond=None)[0].reshape((L, M, N), order="F")
    return phi - phi[:, 0, :].mean()


def test_poisson_dense_8x6x8():
    g = ChannelGrid(8, 6, 8)
    og = O.OracleGrid(8, 6, 8)
    f = fields.poisson_rhs(g, seed=3)
    dph = O.poisson(og, f)
    ref = _dense_poisson(g, f)
    assert np.abs(dph[:, 1:-1, :] - ref).max() < 1e-12 * np.abs(ref).max()
    assert np.all(dph[:, 0, :] == 0.0) and np.all(dph[:, -1, :] == 0.0)


@pytest.mark.parametrize("dims", [(16, 12, 16), (64, 64, 64), (12, 10, 20)])
def test_poisson_residual(dims):
    g = ChannelGrid(*dims)
    og = O.OracleGrid(*dims)
    f = fields.poisson_rhs(g, seed=5)
    dph = O.poisson(og, f)
    mx, _ = O.poisson_residual(og, dph, f)
    assert mx < 1e-12 * np.abs(f).max() * max(1.0, np.abs(g.acph).max() / 1e3)
    assert abs(dph[:, 1, :].mean()) < 1e-13 * np.abs(dph).max()


def test_poisson_manufactured_mode():
    """TEST_POISSON-style manufactured case (src/SOLVERRK3_MOM_IO.f90:171-202):
    f = -4 sin x cos x = -2 sin 2x is an eigenmode of the x operator."""
    g = ChannelGrid(32, 8, 16)
    og = O.OracleGrid(32, 8, 16)
    x = (np.arange(g.L) + 0.0) * g.dx
    f = np.zeros(g.shape_cell, order="F")
    f[:, :, :] = (-4.0 * np.sin(x) * np.cos(x))[:, None, None]
    dph = O.poisson(og, f)
    lam = -g.ak1[2]
    assert np.allclose(dph[:, 1:-1, :], f / lam, atol=1e-13)


# --------------------------------------------------------------------------
# tg substep: invariants and the projection
# --------------------------------------------------------------------------
def _laminar(g):
    q = g.zeros(g.shape_q)
    u = 1.5 * (1.0 - g.ycc ** 2)
    q[:, 1:g.M + 1, :, 0] = u[None, :, None]
    pr = g.zeros(g.shape_ghost)
    return q, pr, g.zeros(g.shape_convh0)


def test_convection_of_parallel_flow_vanishes():
    g = ChannelGrid(16, 12, 8)
    og = O.OracleGrid(16, 12, 8)
    st = O.OracleState(og, *_laminar(g))
    for name in ("convection_x", "convection_y", "convection_z"):
        st.stage(name)
    assert np.abs(st.qtmp).max() == 0.0
    assert np.abs(st.dph).max() == 0.0
    assert np.abs(st.rhsllphi).max() == 0.0


def test_convection_conservation():
    """Divergence form: volume sums of H_x and H_z vanish on the periodic/wall box."""
    g = ChannelGrid(16, 12, 8)
    og = O.OracleGrid(16, 12, 8)
    q, pr, c0 = fields.init_tg_fields(g, seed=11, vper=0.3)
    st = O.OracleState(og, q, pr, c0)
    for name in ("convection_x", "convection_y", "convection_z"):
        st.stage(name)
    w = (1.0 / g.dyfi)[None, :, None]
    scale = np.abs(st.qtmp).max()
    assert abs(np.sum(st.qtmp * w)) < 1e-12 * scale * g.ncell
    assert abs(np.sum(st.rhsllphi * w)) < 1e-12 * scale * g.ncell


@pytest.mark.parametrize("dims", [(16, 12, 16), (32, 24, 16)])
def test_substep_projection_and_mass_flux(dims):
    g = ChannelGrid(*dims)
    og = O.OracleGrid(*dims)
    q, pr, c0 = fields.init_tg_fields(g, seed=1973)
    st = O.OracleState(og, q, pr, c0)
    dt = 0.5 / st.cfl()
    ub0 = fields.bulk_velocity(g, st.q)
    for ns in (0, 1, 2, 3):
        st.substep(ns, dt)
        assert st.divgck() < 1e-10
        # wall values
        assert np.all(st.q[:, 0, :, :] == 0) and np.all(st.q[:, -1, :, :] == 0) and np.all(st.q[:, 1, :, 1] == 0)
    assert fields.bulk_velocity(g, st.q) == pytest.approx(ub0, rel=1e-6)


def test_laminar_stays_parallel():
    g = ChannelGrid(8, 16, 8)
    og = O.OracleGrid(8, 16, 8)
    st = O.OracleState(og, *_laminar(g))
    for ns in (1, 2, 3):
        st.substep(ns, 1e-3)
    u = st.q[:, 1:-1, :, 0]
    assert np.abs(u - u[:1, :, :1]).max() < 1e-14
    assert np.abs(st.q[..., 1:]).max() < 1e-14


# --------------------------------------------------------------------------
# Spline tables (CUBIC_SPLINE / NIST_SLEVAL_*)
# --------------------------------------------------------------------------
def test_spline_reproduces_cubics():
    x = np.sort(RNG.uniform(0.0, 2.0, 40))
    y = 0.3 - 1.2 * x + 0.7 * x ** 2 - 0.25 * x ** 3
    b, c, d = O.cubic_spline(x, y)
    u = RNG.uniform(x[0] + 1e-9, x[-1] - 1e-9, 500)
    for uu in u:
        e = np.zeros(1)
        st = O.lib().orc_sleval(len(x), O._ptr(x), O._ptr(y), O._ptr(b), O._ptr(c), O._ptr(d), uu, 0, O._ptr(e))
        assert st == 0
        assert e[0] == pytest.approx(0.3 - 1.2 * uu + 0.7 * uu ** 2 - 0.25 * uu ** 3, abs=1e-12)
    e = np.zeros(1)
    assert O.lib().orc_sleval(len(x), O._ptr(x), O._ptr(y), O._ptr(b), O._ptr(c), O._ptr(d), x[0] - 1, 0, O._ptr(e)) == 1


# --------------------------------------------------------------------------
# Planted mistakes: a wrong oracle must fail the checks above
# --------------------------------------------------------------------------
def test_planted_mistake_in_poisson_operator_is_caught():
    """Swap DYCI(j) and DYCI(j+1) in the y operator (a plausible index slip):
    the dense finite-volume check must fail."""
    g = ChannelGrid(8, 6, 8)
    og = O.OracleGrid(8, 6, 8)
    og.arr["apph"][:-1], og.arr["amph"][1:] = (og.arr["dyfi"][:-1] * og.arr["dyci"][:-2],
                                               og.arr["dyfi"][1:] * og.arr["dyci"][2:])
    f = fields.poisson_rhs(g, seed=3)
    dph = O.poisson(og, f)
    ref = _dense_poisson(g, f)
    assert np.abs(dph[:, 1:-1, :] - ref).max() > 1e-6 * np.abs(ref).max()


def test_planted_mistake_in_tdma_is_caught():
    """Drop the lower-diagonal term of one row: the dense solve disagrees."""
    n = 7
    a = RNG.uniform(-0.9, -0.1, n)
    b = 2.0 + np.abs(RNG.standard_normal(n))
    c = RNG.uniform(-0.9, -0.1, n)
    d = RNG.standard_normal(n)
    a_bad = a.copy()
    a_bad[3] = 0.0
    r = d.copy()
    O.tdma_noncyc(a_bad, b, c, r, None, n, 1, 1, n)
    x = np.linalg.solve(_dense(a, b, c, False), d)
    assert np.abs(r - x).max() > 1e-3


# --------------------------------------------------------------------------
# xz-mean statistics (PP_MEAN_ZX_FLOW_TG first and second moments)
# --------------------------------------------------------------------------
def test_mean_xz_against_numpy():
    """The oracle's moments against a vectorised numpy evaluation of the
    cell-centred definitions, and the laminar profile's known mean."""
    g = ChannelGrid(16, 12, 8)
    og = O.OracleGrid(16, 12, 8)
    q, pr, c0 = fields.init_tg_fields(g, seed=4)
    pr = np.asfortranarray(RNG.standard_normal(pr.shape))
    st = O.OracleState(og, q, pr, c0)
    out = st.mean_xz()
    u = q[:, 1:-1, :, 0] + np.roll(q[:, 1:-1, :, 0], -1, axis=0)
    v = q[:, 1:-1, :, 1] + q[:, 2:, :, 1]
    w = q[:, 1:-1, :, 2] + np.roll(q[:, 1:-1, :, 2], -1, axis=2)
    p = pr[:, 1:-1, :]
    m = lambda a: a.mean(axis=(0, 2))
    ref = np.stack([m(u) / 2, m(v) / 2, m(w) / 2, m(p), m(u * u) / 4, m(u * v) / 4, m(u * w) / 4, m(v * v) / 4,
                    m(v * w) / 4, m(w * w) / 4, m(u * p) / 2, m(v * p) / 2, m(w * p) / 2], 1)
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-14)
    # bulk velocity of the rescaled initial field: sum over rows of mean u / DYFI / 2 = 1
    assert np.sum(out[:, 0] / og.dyfi) / 2.0 == pytest.approx(1.0, rel=1e-12)
