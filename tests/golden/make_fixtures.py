"""Writes the golden fixtures of this directory from the reference's own
routines, compiled from the reference sources into oracle/_ref by
`make -C oracle ref` (this container only; the reference tree does not travel):

  tdma_ref.npz   TDMAIJJ_nonCYC, TDMAIJJ_CYC, TDMAIJI_nonCYC, TDMAIJI_CYC
                 (src/TDMAIJ{I,J}_{CYC,nonCYC}.f90:67) on random 4x7, 3x12 and
                 6x5 systems (b = 2 + |N(0,1)|, a, c ~ U(-0.9, -0.1), r ~ N(0,1),
                 boundary values BC ~ N(0,1) for the non-cyclic solvers)
  fft99_ref.npz  FFT99 forward (ISIGN = -1) and inverse (ISIGN = +1) and CFFT99
                 both ways (libs/fft99/FFT99.f90:256, CFFT99.f90:314) on LOT = 3
                 vectors of N = 8, 12, 16, 30

The files hold inputs and the reference's outputs only.

    make -C oracle ref && python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

TDMA_SHAPES = [(4, 7), (3, 12), (6, 5)]
FFT_NS = [8, 12, 16, 30]


def tdma_fixtures(rng):
    out = {}
    for march in ("IJJ", "IJI"):
        for cyclic in (0, 1):
            for isz, jsz in TDMA_SHAPES:
                key = f"{march}_{'cyc' if cyclic else 'noncyc'}_{isz}x{jsz}"
                sh = (isz, jsz)
                a = rng.uniform(-0.9, -0.1, sh)
                c = rng.uniform(-0.9, -0.1, sh)
                b = 2.0 + np.abs(rng.standard_normal(sh))
                r = rng.standard_normal(sh)
                nl = isz if march == "IJJ" else jsz
                bc = rng.standard_normal((nl, 2)) if not cyclic else np.zeros((nl, 2))
                x = O.ref_tdma_solve(a, b, c, r, None if cyclic else bc, march, cyclic, lo=(2, 0))
                for n, v in (("a", a), ("b", b), ("c", c), ("r", r), ("bc", bc), ("x", x)):
                    out[f"{key}_{n}"] = np.asfortranarray(v)
    return out


def fft_fixtures(rng):
    out = {}
    lot = 3
    for n in FFT_NS:
        x = rng.standard_normal((n, lot))
        data = np.zeros((n + 2, lot))
        data[1:n + 1] = x
        data[0] = x[n - 1]
        data[n + 1] = x[0]
        out[f"fft99_fwd_{n}_x"] = x
        out[f"fft99_fwd_{n}_ab"] = O.ref_fft99_transform(data, n, -1)
        ab = rng.standard_normal((n + 2, lot))
        out[f"fft99_inv_{n}_ab"] = ab
        out[f"fft99_inv_{n}_x"] = O.ref_fft99_transform(ab, n, +1)[1:n + 1]
        z = rng.standard_normal((n, lot)) + 1j * rng.standard_normal((n, lot))
        out[f"cfft99_{n}_z"] = z
        for s in (-1, 1):
            out[f"cfft99_{n}_{'m' if s < 0 else 'p'}"] = O.ref_cfft99_transform(z, s)
    return out


def main():
    if O.ref_tdma() is None or O.ref_fft99() is None:
        raise SystemExit("oracle/_ref is not built: run `make -C oracle ref` where the reference tree exists")
    rng = np.random.default_rng(19731203)
    np.savez(os.path.join(HERE, "tdma_ref.npz"), **tdma_fixtures(rng))
    np.savez(os.path.join(HERE, "fft99_ref.npz"), **fft_fixtures(rng))
    print("wrote", os.path.join(HERE, "tdma_ref.npz"), os.path.join(HERE, "fft99_ref.npz"))


if __name__ == "__main__":
    main()
