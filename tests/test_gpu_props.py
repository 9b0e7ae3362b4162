"""THERM_PROP_UPDATE on the device (table lookup) against the oracle.

The lookup is integer interval search plus cubic polynomials evaluated
without fused multiply-adds, so the device result must equal the oracle's
bit for bit."""
import numpy as np
import pytest

import oracle as O
from chapsim1ww_amd import props

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scaled():
    return props.scale_table(props.synthetic_table(), t0=1.0)


@pytest.fixture(scope="module")
def ptab(scaled):
    t = props.PropertyTable(scaled)
    yield t
    t.close()


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _oracle(scaled, h, init):
    tab, n = O.build_prop_tab(scaled)
    out = [np.full_like(h, init) for _ in range(5)]
    bad = O.therm_prop_update(tab, n, h, *out)
    return bad, out


@pytest.mark.parametrize("ncell", [1, 1000, 1 << 20])
def test_table_lookup_bitwise(scaled, ptab, ncell):
    print(f"props table ncell={ncell}", flush=True)
    rng = np.random.default_rng(ncell)
    lo, hi = ptab.h_range
    h = rng.uniform(lo, hi, ncell)
    h[::97] = 0.0                     # |h| < 1e-10: left unchanged
    bad, ref = _oracle(scaled, h, -3.0)
    assert bad == 0
    dh = _dev(h)
    outs = [torch.full((ncell,), -3.0, dtype=torch.float64, device="cuda") for _ in range(5)]
    T, rho, mu, kap, cp = outs
    nf = ptab.update(dh, T, rho, mu, kap, cp)
    torch.cuda.synchronize()
    assert nf == 0
    for name, a, b in zip(("T", "rho", "mu", "kap", "cp"), outs, ref):
        got = a.cpu().numpy()
        assert np.array_equal(got, b), (name, np.abs(got - b).max())


def test_table_lookup_at_knots_and_edges(scaled, ptab):
    """Values at and next to every knot (the interval search's boundaries)."""
    H = scaled["H"]
    h = np.concatenate([H[1:-1], np.nextafter(H[1:-1], np.inf), np.nextafter(H[1:-1], -np.inf),
                        [np.nextafter(H[0], np.inf), np.nextafter(H[-1], -np.inf)]])
    bad, ref = _oracle(scaled, h, 0.0)
    # one ulp above the first knot, S_HT gives T = T_1 exactly: out of range
    # for the T splines, as in the reference (and the oracle)
    assert bad == 1
    outs = [torch.zeros(h.size, dtype=torch.float64, device="cuda") for _ in range(5)]
    from chapsim1ww_amd.native import CHS_E_RANGE, ChsError
    with pytest.raises(ChsError) as ei:
        ptab.update(_dev(h), *outs)
    assert ei.value.code == CHS_E_RANGE and "1 cells" in str(ei.value)
    for a, b in zip(outs, ref):
        assert np.array_equal(a.cpu().numpy(), b)


def test_out_of_range_is_reported_and_left_unchanged(scaled, ptab):
    from chapsim1ww_amd.native import CHS_E_RANGE, ChsError
    lo, hi = ptab.h_range
    h = np.array([0.2, lo - 0.5, hi + 0.5, lo, np.nan, 0.0, -0.1])
    bad, ref = _oracle(scaled, h, 9.0)
    outs = [torch.full((h.size,), 9.0, dtype=torch.float64, device="cuda") for _ in range(5)]
    with pytest.raises(ChsError) as ei:
        ptab.update(_dev(h), *outs)
    assert ei.value.code == CHS_E_RANGE
    assert "3 cells" in str(ei.value)   # below, above, at the first knot
    assert bad == 3
    for a, b in zip(outs, ref):
        got = a.cpu().numpy()
        assert np.array_equal(got[[0, 6]], b[[0, 6]])
        assert np.all(got[[1, 2, 3, 5]] == 9.0)
        assert np.isnan(got[4]) and np.isnan(b[4])   # a NaN enthalpy propagates, as in the reference


def test_rows_of_a_ghost_aware_array(scaled, ptab):
    """ENTHALPY(NCL1, 0:N2DO+1, NCL3): interior rows only, ghost planes untouched."""
    L, M, N = 32, 12, 8
    rng = np.random.default_rng(4)
    lo, hi = ptab.h_range
    H3 = np.asfortranarray(rng.uniform(lo, hi, (L, M + 2, N)))
    flat = H3.ravel(order="F")
    interior = H3[:, 1:M + 1, :].ravel(order="F")
    bad, ref = _oracle(scaled, interior, 0.0)
    assert bad == 0
    dh = _dev(flat)
    outs = [torch.full((flat.size,), 5.0, dtype=torch.float64, device="cuda") for _ in range(5)]
    # views starting at (1, 1, 1): pass device addresses offset by L values
    off = [t[L:] for t in [dh] + outs]
    ptab.update(off[0], *off[1:], rows=(N, L * M, L * (M + 2)))
    for a, b in zip(outs, ref):
        got = a.cpu().numpy().reshape((L, M + 2, N), order="F")
        assert np.array_equal(got[:, 1:M + 1, :].ravel(order="F"), b)
        assert np.all(got[:, 0, :] == 5.0) and np.all(got[:, M + 1, :] == 5.0)


@pytest.mark.parametrize("n,peak", [(4096, 4.0), (2048, 40.0), (512, 2000.0)])
def test_uneven_knots_and_bucket_edges(n, peak):
    """Tables whose H knots are spaced unevenly (c_p peak/base = 5 at 4096
    knots, the largest count accepted; 41; 2001: long walks from a bucket's
    first candidate): random values, every knot and its neighbours one ulp
    away, and the edges of power-of-two bucket counts from n to 64 n and their
    neighbours, bit for bit against the oracle's bisection."""
    print(f"props uneven table n={n} peak={peak}", flush=True)
    sc = props.scale_table(props.synthetic_table(n, peak=peak), t0=1.0)
    t = props.PropertyTable(sc)
    try:
        H = sc["H"]
        lo, hi = H[0], H[-1]
        rng = np.random.default_rng(n)
        vals = [rng.uniform(lo, hi, 200000), H[1:-1], np.nextafter(H[1:-1], np.inf),
                np.nextafter(H[1:-1], -np.inf)]
        for m in range(7):
            nb = n << m
            e = lo + np.arange(1, nb) * (hi - lo) / nb
            vals += [e, np.nextafter(e, np.inf), np.nextafter(e, -np.inf)]
        h = np.concatenate(vals)
        h = h[(h > lo) & (h < hi)]
        bad, ref = _oracle(sc, h, 0.0)
        outs = [torch.zeros(h.size, dtype=torch.float64, device="cuda") for _ in range(5)]
        from chapsim1ww_amd.native import ChsError
        try:
            nf = t.update(_dev(h), *outs)
        except ChsError as e:
            nf = int(str(e).split(" cells")[0].split()[-1])
        torch.cuda.synchronize()
        assert nf == bad
        for name, a, b in zip(("T", "rho", "mu", "kap", "cp"), outs, ref):
            got = a.cpu().numpy()
            assert np.array_equal(got, b), (name, np.abs(got - b).max())
    finally:
        t.close()
