"""The Fortran route: fortran/test_shim and fortran/test_dropin (built by
__graft_entry__.build() with amdflang) call the library through the
iso_c_binding shim, the latter through the drop-ins under the reference's own
names on module arrays: the four TDMA solvers against the reference's golden
outputs, FFT99_POIS3D_periodicxz, SOLVERRK3_MOM_tg for NS = 0..3,
THERM_PROP_UPDATE, SOLVERRK3_ENG_IO(1) and SOLVERRK3_MOM_IO(1) against the
oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from chapsim1ww_amd import fields, props
from chapsim1ww_amd.grid import ChannelGrid

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "fortran")
GOLD = os.path.join(ROOT, "tests", "golden", "tdma_ref.npz")
HEAT = ("rhoh", "h", "T", "rho", "rho0", "rho1", "mu", "mu0", "kap", "rhs0", "g", "g0", "wallflux", "q", "pr",
        "drhodtp", "divu", "rho_stg", "mu_stg", "explt0", "drhoi_stg")


def _exe(name):
    p = os.path.join(FDIR, name)
    if not os.path.exists(p):
        pytest.skip(f"{p} not built (no amdflang where build() ran)")
    return p


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _f(a):
    return np.asfortranarray(a, dtype=np.float64).tobytes(order="F")


def test_fortran_shim_tdma():
    print("fortran test_shim", flush=True)
    r = subprocess.run([_exe("test_shim")], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "max residual" in r.stdout


def test_fortran_dropins_against_oracle(tmp_path):
    print("fortran test_dropin", flush=True)
    exe = _exe("test_dropin")
    L, M, N = 32, 24, 16
    g = ChannelGrid(L, M, N)
    og = O.OracleGrid(L, M, N)
    scaled = props.scale_table(props.synthetic_table(256), t0=1.0)
    st, tab, n = O.io_case(og, scaled, seed=12)
    q, pr, c0 = fields.init_tg_fields(g, seed=4)
    f = fields.poisson_rhs(g, seed=6)
    dt_tg = 0.5 / O.OracleState(og, q, pr, c0).cfl()
    dt_io, cthecd = 1e-3, 1.0 / 2800.0
    ni = np.array([L, M, N, len(scaled["H"]), 1, 1, 1, 0, 0], "<i4")
    nr = np.array([g.dxi, g.dzi, g.dxqi, g.dzqi, g.cvisc, g.volm, dt_tg, dt_io, cthecd, 0.0, 0.0], "<f8")
    gold = np.load(GOLD)
    tdma = [("IJJ_noncyc_4x7", "bc"), ("IJI_noncyc_4x7", "bc"), ("IJJ_cyc_4x7", None), ("IJI_cyc_4x7", None)]
    with open(tmp_path / "in.bin", "wb") as fh:
        fh.write(ni.tobytes() + nr.tobytes())
        for k in ("dyfi", "dyci", "wff", "wfb", "amph", "acph", "apph", "amvr", "acvr", "apvr", "ak1", "ak3"):
            fh.write(np.ascontiguousarray(getattr(g, k), np.float64).tobytes())
        fh.write(np.array([4, 7], "<i4").tobytes())
        for key, bc in tdma:
            for v in ("a", "b", "c", "r"):
                fh.write(_f(gold[f"{key}_{v}"]))
            if bc:
                fh.write(_f(gold[f"{key}_bc"]))
        fh.write(_f(f))
        fh.write(_f(q) + _f(pr) + _f(c0))
        for k in ("H", "T", "D", "M", "K", "CP"):
            fh.write(np.ascontiguousarray(scaled[k], np.float64).tobytes())
        for k in HEAT:
            fh.write(_f(st[k]))
    r = subprocess.run([exe, str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "done" in r.stdout
    raw = np.fromfile(tmp_path / "out.bin", dtype="<f8")
    pos = [0]

    def take(shape):
        k = int(np.prod(shape))
        a = raw[pos[0]:pos[0] + k].reshape(shape, order="F")
        pos[0] += k
        return a

    for key, _ in tdma:
        x = take((4, 7))
        assert _rel(x, gold[f"{key}_x"]) < 1e-12, key
    # Poisson: the oracle, ghost planes zero
    dph = take((L, M + 2, N))
    assert _rel(dph, O.poisson(og, f)) < 1e-12
    # SOLVERRK3_MOM_tg NS = 0..3
    ost = O.OracleState(og, q, pr, c0)
    for ns in (0, 1, 2, 3):
        ost.substep(ns, dt_tg)
    for name, shp, ref in (("Q", g.shape_q, ost.q), ("PR", g.shape_ghost, ost.pr),
                           ("CONVH0", g.shape_convh0, ost.convh0)):
        e = _rel(take(shp), ref)
        print(name, e)
        assert e < 1e-12, name
    # THERM_PROP_UPDATE: rows 1..M from ENTHALPY, ghost rows untouched (-1)
    outs = [take((L, M + 2, N)) for _ in range(4)]
    h_in = np.ascontiguousarray(st["h"][:, 1:M + 1, :]).ravel(order="F")
    ref = [np.zeros(h_in.size) for _ in range(4)]
    assert O.therm_prop_update(tab, n, np.ascontiguousarray(h_in), *ref) == 0
    for o, rr in zip(outs, ref):
        assert np.array_equal(o[:, 1:M + 1, :].ravel(order="F"), rr)
        assert np.all(o[:, 0] == -1.0) and np.all(o[:, M + 1] == -1.0)
    # SOLVERRK3_ENG_IO(1), staging, SOLVERRK3_MOM_IO(1)
    io = O.OracleIO(og, st)
    assert O.energy_substep(og, st, tab, n, 1, dt_io, cthecd, O.rhoh_max(scaled)) == 0
    io.prepare(1)
    io.substep(1, dt_io)
    shapes = {"rhs0": (L, M, N), "explt0": (L, M, N, 3), "wallflux": (L, 2)}
    for k in ("g", "g0", "q", "rho_stg", "mu_stg", "drhoi_stg"):
        shapes[k] = (L, M + 2, N, 3)
    for k in HEAT:
        got = take(shapes.get(k, (L, M + 2, N)))
        e = _rel(got, st[k]) if np.abs(st[k]).max() > 0 else float(np.abs(got).max())
        print(k, e)
        assert e < 1e-12, k
    assert pos[0] == raw.size
