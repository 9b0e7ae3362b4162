"""The oracle against the golden fixtures written from the reference's own
routines (tests/golden/make_fixtures.py: the four TDMAIJ?_(non)CYC line
solvers and the FFT99/CFFT99 package, compiled from the reference sources),
and, where oracle/_ref is built, against those routines directly on fresh
random systems."""
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TDMA = np.load(os.path.join(GOLD, "tdma_ref.npz"))
FFT = np.load(os.path.join(GOLD, "fft99_ref.npz"))
TDMA_KEYS = sorted({k.rsplit("_", 1)[0] for k in TDMA.files})


def _case(key):
    return {n: np.array(TDMA[f"{key}_{n}"], order="F") for n in ("a", "b", "c", "r", "bc", "x")}


def _layout(key, shape):
    isz, jsz = shape
    if key.startswith("IJJ"):
        return jsz, isz, isz, 1       # n, nlines, elem_stride, line_stride
    return isz, jsz, 1, isz


def _oracle_solve(key, d, bc=None):
    n, nl, es, ls = _layout(key, d["a"].shape)
    R = np.array(d["r"], order="F")
    if "_cyc_" in key:
        O.tdma_cyc(d["a"], d["b"], d["c"], R, n, nl, es, ls)
    else:
        O.tdma_noncyc(d["a"], d["b"], d["c"], R, np.array(d["bc"] if bc is None else bc, order="F"), n, nl, es, ls)
    return R


def _dense_check(key, d):
    """The fixture itself solves the dense system (corner entries for the
    cyclic solvers, BC(line,1:2) moved to the right-hand side otherwise)."""
    n, nl, es, ls = _layout(key, d["a"].shape)
    for l in range(nl):
        sl = (l, slice(None)) if key.startswith("IJJ") else (slice(None), l)
        a, b, c, r, x = d["a"][sl], d["b"][sl], d["c"][sl], d["r"][sl].copy(), d["x"][sl]
        A = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
        if "_cyc_" in key:
            A[0, n - 1] += a[0]
            A[n - 1, 0] += c[n - 1]
        else:
            r[0] -= a[0] * d["bc"][l, 0]
            r[-1] -= c[-1] * d["bc"][l, 1]
        assert np.abs(A @ x - r).max() < 1e-13


@pytest.mark.parametrize("key", TDMA_KEYS)
def test_oracle_tdma_matches_reference_fixture(key):
    d = _case(key)
    _dense_check(key, d)
    x = _oracle_solve(key, d)
    err = np.abs(x - d["x"]).max() / np.abs(d["x"]).max()
    assert err < 1e-14, err


def test_planted_bc_swap_is_caught():
    """BC(line,1) and BC(line,2) exchanged in the oracle's input: the fixture
    comparison fails."""
    key = next(k for k in TDMA_KEYS if "_noncyc_" in k)
    d = _case(key)
    x = _oracle_solve(key, d, bc=d["bc"][:, ::-1])
    assert np.abs(x - d["x"]).max() > 1e-3


@pytest.mark.parametrize("n", [8, 12, 16, 30])
def test_oracle_fft99_matches_reference_fixture(n):
    x, ab = FFT[f"fft99_fwd_{n}_x"], FFT[f"fft99_fwd_{n}_ab"]
    mine = np.stack([O.fft99_fwd(x[:, l]) for l in range(x.shape[1])], 1)
    assert np.abs(mine - ab).max() < 1e-14
    ab, xr = FFT[f"fft99_inv_{n}_ab"], FFT[f"fft99_inv_{n}_x"]
    mine = np.stack([O.fft99_inv(ab[:, l], n) for l in range(ab.shape[1])], 1)
    assert np.abs(mine - xr).max() < 1e-12
    z = FFT[f"cfft99_{n}_z"]
    for s, tag in ((-1, "m"), (1, "p")):
        mine = np.stack([O.cfft99(z[:, l], s) for l in range(z.shape[1])], 1)
        assert np.abs(mine - FFT[f"cfft99_{n}_{tag}"]).max() < 1e-12
    # the fixture is the reference's definition (FFT99.f90:98-105): with the
    # 1/N of the forward transform, a_k + i b_k = numpy's rfft / N
    F = np.fft.rfft(x, axis=0) / n
    ab = FFT[f"fft99_fwd_{n}_ab"]
    assert np.abs(ab[0::2] - F.real).max() < 1e-14
    assert np.abs(ab[1::2] - F.imag).max() < 1e-14


@pytest.mark.parametrize("march", ["IJJ", "IJI"])
@pytest.mark.parametrize("cyclic", [0, 1])
@pytest.mark.parametrize("shape", [(5, 9), (2, 64), (33, 4)])
def test_oracle_tdma_matches_compiled_reference(march, cyclic, shape):
    """Fresh random systems through the compiled reference routines (only
    where oracle/_ref was built from the reference tree)."""
    if O.ref_tdma() is None:
        pytest.skip("oracle/_ref not built (no reference tree)")
    rng = np.random.default_rng(hash((march, cyclic, shape)) % 2**32)
    isz, jsz = shape
    n, nl = (jsz, isz) if march == "IJJ" else (isz, jsz)
    if n < 3:
        pytest.skip("the reference needs at least 3 rows")
    a = rng.uniform(-0.9, -0.1, shape)
    c = rng.uniform(-0.9, -0.1, shape)
    b = 2.0 + np.abs(rng.standard_normal(shape))
    r = rng.standard_normal(shape)
    bc = rng.standard_normal((nl, 2))
    ref = O.ref_tdma_solve(a, b, c, r, None if cyclic else bc, march, cyclic)
    key = f"{march}_{'cyc' if cyclic else 'noncyc'}_x"
    d = {"a": np.asfortranarray(a), "b": np.asfortranarray(b), "c": np.asfortranarray(c),
         "r": np.asfortranarray(r), "bc": np.asfortranarray(bc)}
    x = _oracle_solve(key, d)
    assert np.abs(x - ref).max() / np.abs(ref).max() < 1e-14
