"""chs_tdma_batched_noalloc (the entry point that allocates nothing) against
chs_tdma_batched through the C ABI: bit-identical results on two calls in a
row, the second smaller, for lines that take the static scratch workspace
(cyclic lines of odd length, non-cyclic lines of 600 rows) and for lines the
segmented kernel solves without one; oversized calls are refused on the host."""
import ctypes as C

import numpy as np
import pytest

from chapsim1ww_amd import native as nat

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.from_numpy(np.asfortranarray(a).ravel(order="F").copy()).to("cuda")


def _system(rng, n, nl, march):
    shape, es, ls = ((nl, n), nl, 1) if march == "IJJ" else ((n, nl), 1, n)
    a = rng.uniform(-0.9, -0.1, shape)
    c = rng.uniform(-0.9, -0.1, shape)
    b = 2.0 + np.abs(rng.standard_normal(shape))
    d = rng.standard_normal(shape)
    bc = rng.standard_normal((nl, 2))
    return [_dev(v) for v in (a, b, c, d, bc)], es, ls


def _solve(fn, arrs, n, nl, es, ls, cyclic, r):
    a, b, c, _, bc = arrs
    rc = fn(a.data_ptr(), b.data_ptr(), c.data_ptr(), r.data_ptr(), None if cyclic else bc.data_ptr(), n, nl, es,
            ls, int(cyclic), None)
    nat.check(rc)


# (first call, second smaller call): (n, nlines, cyclic, march)
CASES = [
    ((7, 20000, True, "IJJ"), (5, 300, True, "IJJ")),        # scratch path, cyclic odd n
    ((600, 2000, False, "IJI"), (513, 7, False, "IJJ")),     # scratch path, long lines
    ((192, 4096, True, "IJJ"), (64, 129, False, "IJI")),     # segmented path, no workspace
]


@pytest.mark.parametrize("first,second", CASES)
def test_noalloc_matches_allocating_entry(first, second):
    lib = nat.load()
    rng = np.random.default_rng(7)
    for (n, nl, cyclic, march) in (first, second):
        print(f"noalloc n={n} nlines={nl} cyclic={cyclic} {march}", flush=True)
        arrs, es, ls = _system(rng, n, nl, march)
        r_ref, r_new = arrs[3].clone(), arrs[3].clone()
        _solve(lib.chs_tdma_batched, arrs, n, nl, es, ls, cyclic, r_ref)
        _solve(lib.chs_tdma_batched_noalloc, arrs, n, nl, es, ls, cyclic, r_new)
        torch.cuda.synchronize()
        x_ref, x_new = r_ref.cpu().numpy(), r_new.cpu().numpy()
        assert np.isfinite(x_ref).all()
        assert np.array_equal(x_ref, x_new)


def test_noalloc_refuses_calls_above_its_scratch():
    lib = nat.load()
    # 3 * 7 * 200000 doubles > 4194304: refused before any device work
    rc = lib.chs_tdma_batched_noalloc(1, 1, 1, 1, None, 7, 200000, 200000, 1, 1, None)
    assert rc == nat.CHS_E_ARG
    assert b"static scratch" in lib.chs_last_error()
