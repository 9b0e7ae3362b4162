"""The io xz-plane moments (PP_MEAN_ZX_FLOW_Xperiodic_IO,
src/processing/PP_MEAN_ZX.f90:299-800) and the time means (PP_MEAN_T_*,
src/processing/PP_MEAN_T.f90:322-470) of the oracle against closed forms:
single Fourier modes in x and z on top of y profiles, whose cell-centred
averages and products have exact xz means."""
import numpy as np
import pytest

import oracle as O

L, M, N = 16, 6, 8


def _state(rng, v_shift=1):
    th, ph = 2 * np.pi / L, 2 * np.pi / N
    i = np.arange(L)[:, None, None]
    k = np.arange(N)[None, None, :]
    prof = {n: rng.standard_normal(M + 2)[None, :, None] for n in "abcdpeqtsrhmn"}
    q = np.zeros((L, M + 2, N, 3))
    q[..., 0] = prof["a"] + prof["b"] * np.cos(th * i)
    q[..., 1] = prof["c"] + 0 * i * k
    q[..., 2] = prof["d"] * np.sin(ph * k) + 0 * i
    pr = prof["p"] + prof["e"] * np.cos(th * i) + 0 * k
    T = prof["t"] + prof["s"] * np.cos(ph * k) + 0 * i
    rho = prof["r"] + 0 * i * k
    h = prof["h"] + prof["q"] * np.cos(th * i) + 0 * k
    mu = prof["m"] + 0 * i * k
    st = {"q": q, "g": 2.0 * q, "pr": pr, "T": T, "rho": rho, "h": h, "mu": mu}
    st = {k_: np.asfortranarray(v) for k_, v in st.items()}
    P = {n: v[0, :, 0] for n, v in prof.items()}
    return st, P


def _expected(P):
    th, ph = 2 * np.pi / L, 2 * np.pi / N
    cx, cz = np.cos(th / 2) ** 2, np.cos(ph / 2) ** 2
    J = slice(1, M + 1)
    a, b, d, p, e = P["a"][J], P["b"][J], P["d"][J], P["p"][J], P["e"][J]
    vc = 0.5 * (P["c"][1:M + 1] + P["c"][2:M + 2])
    t, s, r, hh, qq, m = P["t"][J], P["s"][J], P["r"][J], P["h"][J], P["q"][J], P["m"][J]
    ex = np.zeros((M, 37))
    U1 = [a, vc, 0 * a]
    ex[:, 0:3] = np.stack(U1, 1)
    ex[:, 3] = p
    ex[:, 4:7] = 2.0 * ex[:, 0:3]
    ex[:, 7] = a * p + b * e * cx / 2
    ex[:, 8] = vc * p
    ex[:, 9] = 0.0
    uu = {(0, 0): a * a + b * b * cx / 2, (0, 1): a * vc, (0, 2): 0 * a, (1, 1): vc * vc, (1, 2): 0 * a,
          (2, 2): d * d * cz / 2}
    for l, key in enumerate(((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))):
        ex[:, 10 + l] = uu[key]
        ex[:, 16 + l] = 2.0 * uu[key]
    ex[:, 22:31] = np.stack([t, r, hh, m, t * t + s * s / 2, r * r, hh * hh + qq * qq / 2, hh * r,
                             hh * p + qq * e / 2], 1)
    ex[:, 31] = hh * a + qq * b * cx / 2
    ex[:, 32] = hh * vc
    ex[:, 33] = 0.0
    ex[:, 34:37] = 2.0 * ex[:, 31:34]
    return ex


@pytest.fixture(scope="module")
def og():
    return O.OracleGrid(L, M, N)


def test_mean_xz_io_closed_forms(og):
    rng = np.random.default_rng(4)
    st, P = _state(rng)
    got = O.mean_xz_io(og, st)
    ex = _expected(P)
    assert got.shape == (M, 37)
    assert np.abs(got - ex).max() < 1e-14 * max(1.0, np.abs(ex).max())
    st0 = dict(st)
    assert np.all(O.mean_xz_io(og, st0, thermal=0)[:, 22:] == 0.0)


def test_planted_centring_mistake_is_caught(og):
    """v's cell average taken with row j-1 instead of j+1 (the pairing the
    wall-normal staggering fixes) changes <v> and every product with v."""
    rng = np.random.default_rng(4)
    st, P = _state(rng)
    q = st["q"].copy(order="F")
    q[:, 1:, :, 1] = st["q"][:, :-1, :, 1]    # v shifted one row up: v(j+1) now holds v(j)
    st["q"] = q
    got = O.mean_xz_io(og, st)
    ex = _expected(P)
    assert np.abs(got[:, 1] - ex[:, 1]).max() > 1e-3


def test_time_mean_recurrences():
    rng = np.random.default_rng(2)
    vals = [rng.standard_normal(5) for _ in range(7)]
    acc, ns = np.zeros(5), 0
    for v in vals:
        acc, ns = O.mean_t_update(acc, v, ns, tp=1)
    # asymptotic averaging is the running arithmetic mean
    assert ns == 7 and np.abs(acc - np.mean(vals, 0)).max() < 1e-15
    acc, ns = np.zeros(5), 0
    for v in vals:
        acc, ns = O.mean_t_update(acc, v, ns, tp=2, pp_instn_sz=7)
    assert np.abs(acc - np.sum(vals, 0) / 7).max() < 1e-15
