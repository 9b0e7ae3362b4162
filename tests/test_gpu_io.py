"""The heated channel's io momentum substep on the device against the oracle:
the property staging bit for bit, and full substeps (energy, staging, io
momentum in SOLVE's order) within 1e-12 of the oracle, field by field."""
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
       "g": "g", "g0": "g0", "wallflux": "wallflux", "q": "q", "pr": "pr", "drhodtp": "drhodtp", "divu": "divu",
       "rho_stg": "rho_stg", "mu_stg": "mu_stg", "explt0": "explt0", "drhoi_stg": "drhoi_stg"}
FP64_TOL = 1e-12


@pytest.fixture(scope="module")
def scaled():
    return props.scale_table(props.synthetic_table(), t0=1.0)


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _pair(dims, scaled, seed, **kw):
    from chapsim1ww_amd.heated import HeatedChannel
    from chapsim1ww_amd.solver import ChannelSolver
    g = ChannelGrid(*dims)
    og = O.OracleGrid(*dims)
    st, tab, n = O.io_case(og, scaled, seed=seed)
    s = ChannelSolver(g)
    pt = props.PropertyTable(scaled)
    hc = HeatedChannel(s, pt, **kw)
    hc.upload({MAP[k]: v for k, v in st.items()})
    io = O.OracleIO(og, st, **kw)
    return og, st, tab, n, s, pt, hc, io


@pytest.mark.parametrize("dims,mode", [((16, 12, 8), "auto"), ((64, 48, 32), "auto"), ((64, 48, 32), "marching")])
def test_io_prepare_bitwise(scaled, dims, mode):
    """The property staging bit for bit: the row-tile form (auto on these small
    grids) and the y-marching one (mode 'marching')."""
    print(f"io prepare {dims} {mode}", flush=True)
    og, st, tab, n, s, pt, hc, io = _pair(dims, scaled, seed=sum(dims), visimplicit=1)
    s.set_cn_solver(mode)
    for ns in (1, 2):
        io.prepare(ns)
        hc.io_prepare(ns)
        got = hc.download()
        for k in ("drhodtp", "rho_stg", "mu_stg", "q", "drhoi_stg"):
            assert np.array_equal(got[k], st[k]), (ns, k, np.abs(got[k] - st[k]).max())
    pt.close()
    s.close()


@pytest.mark.parametrize("dims,kw", [((16, 12, 8), dict(thermal=1, gravdir=1, f_a=-0.05)),
                                     ((64, 48, 32), dict(thermal=1, gravdir=2, f_a=0.02)),
                                     ((32, 24, 16), dict(thermal=1, flowdrvtp=2, cfgv=0.01, gravdir=3, f_a=0.03)),
                                     ((32, 24, 16), dict(thermal=1, gravdir=1, f_a=-0.05, visimplicit=1)),
                                     ((64, 48, 32), dict(thermal=1, nflow=3, visimplicit=1)),
                                     ((32, 24, 16), dict(thermal=1, nflow=3, gravdir=2, f_a=0.02)),
                                     ((32, 24, 16), dict(thermal=1, nflow=2))])
def test_heated_substeps(scaled, dims, kw):
    """Two RK3 steps (NS = 1, 2, 3 twice) of energy + staging + io momentum;
    every field within 1e-12 of the oracle after every substep and the
    projected divergence below 1e-10 on both sides.  The explicit cases with
    NFLOW = 1 or 3 take the folded path (G + RHS and G0 written by the terms
    kernel, DPGRNS taken off in the update); NFLOW = 2 and the RK-CN cases
    keep the separate copy and mass-flux pass."""
    print(f"heated substeps {dims} {kw}", flush=True)
    og, st, tab, n, s, pt, hc, io = _pair(dims, scaled, seed=3 + sum(dims), **kw)
    rmax = O.rhoh_max(scaled)
    dt = 1e-3
    # The random start is far from divergence-free, so the first projections
    # put pressures of O(100-1000) into PR that later substeps cancel: PR is
    # compared relative to the largest |PR| of the trajectory so far (the
    # operands of that cancellation), the other fields to their own maximum.
    pr_scale = np.abs(st["pr"]).max()
    for step, ns in enumerate((1, 2, 3, 1, 2, 3)):
        assert O.energy_substep(og, st, tab, n, ns, dt, CTHECD, rmax) == 0
        io.prepare(ns)
        io.substep(ns, dt)
        hc.substep(ns, dt, CTHECD)
        got = hc.download()
        errs = {k: _rel(got[MAP[k]], st[k]) for k in ("g", "g0", "explt0", "divu", "q", "rho_stg", "mu_stg",
                                                      "drhodtp", "rho", "T", "h", "drhoi_stg")}
        pr_scale = max(pr_scale, np.abs(st["pr"]).max())
        errs["pr"] = float(np.abs(got["pr"] - st["pr"]).max() / pr_scale)
        dg, _ = hc.divgck_io()
        do, _ = io.divgck()
        print(step, ns, {k: f"{v:.1e}" for k, v in errs.items()}, f"div {dg:.1e} {do:.1e}", flush=True)
        for k, e in errs.items():
            assert e < FP64_TOL, (step, ns, k, e)
        assert dg < 1e-10 and do < 1e-10
    pt.close()
    s.close()


def test_isothermal_io_and_divgck_nan(scaled):
    """thermal = 0 (RHO_STG = MU_STG = 1, DrhoDtP = 0 as allocated): the
    momentum substep alone against the oracle, then NaN propagation of
    DIVGCK_io."""
    print("isothermal io", flush=True)
    dims = (32, 24, 16)
    og, st, tab, n, s, pt, hc, io = _pair(dims, scaled, seed=5, thermal=0)
    L, M, N = dims
    for k in ("rho", "rho0", "mu", "mu0"):
        st[k][...] = 1.0
    hc.upload({MAP[k]: st[k] for k in ("rho", "rho0", "mu", "mu0")})
    pr_scale = np.abs(st["pr"]).max()
    for ns in (1, 2, 3):
        io.prepare(ns)
        io.substep(ns, 2e-3)
        hc.io_prepare(ns)
        hc.solverrk3_mom_io(ns, 2e-3)
        pr_scale = max(pr_scale, np.abs(st["pr"]).max())
    got = hc.download()
    for k in ("g", "explt0", "q"):
        assert _rel(got[k], st[k]) < FP64_TOL, k
    assert np.abs(got["pr"] - st["pr"]).max() < FP64_TOL * pr_scale
    g = got["g"].copy(order="F")
    g[3, 5, 7, 0] = np.nan
    hc.upload({"g": g})
    dg, _ = hc.divgck_io()
    assert np.isnan(dg)
    pt.close()
    s.close()


@pytest.mark.parametrize("dims,kw", [((64, 48, 32), dict(thermal=1, gravdir=2, f_a=0.02)),
                                     ((32, 24, 16), dict(thermal=1, nflow=2))])
def test_io_terms_marching_matches_rowtile(scaled, dims, kw):
    """The y-marching terms kernel against the row-tile one (chs_set_cn_solver
    mode 2) on the same start: they evaluate the same expressions, only the
    mass-flux partial sums group differently, so after three substeps every
    field agrees to a few ulps."""
    print(f"io terms marching vs rowtile {dims} {kw}", flush=True)
    outs = []
    for mode in ("marching", "rowtile"):
        og, st, tab, n, s, pt, hc, io = _pair(dims, scaled, seed=41, **kw)
        s.set_cn_solver(mode)
        for ns in (1, 2, 3):
            hc.substep(ns, 1e-3, CTHECD)
        outs.append(hc.download())
        pt.close()
        s.close()
    for k in ("g", "g0", "explt0", "q", "rho", "T"):
        e = _rel(outs[0][MAP[k]], outs[1][MAP[k]])
        print(k, e, flush=True)
        assert e < 1e-13, (k, e)
    pr_scale = np.abs(outs[1]["pr"]).max()
    assert np.abs(outs[0]["pr"] - outs[1]["pr"]).max() < 1e-13 * pr_scale


def test_mean_xz_io_and_time_means(scaled):
    """PP_MEAN_ZX_FLOW_Xperiodic_IO on the device against the oracle on the
    device's own fields after each of three heated substeps (only the
    summation order differs: 1e-13 of each column's largest value), with the
    time means of PP_MEAN_T (asymptotic, then summing) against the recurrence
    on the oracle's moments; the tg time means too."""
    print("mean xz io + time means", flush=True)
    dims = (32, 24, 16)
    og, st, tab, n, s, pt, hc, io = _pair(dims, scaled, seed=77)
    rmax = O.rhoh_max(scaled)
    acc, ns_o = np.zeros((dims[1], 37)), 0
    for ns in (1, 2, 3):
        assert O.energy_substep(og, st, tab, n, ns, 1e-3, CTHECD, rmax) == 0
        io.prepare(ns)
        io.substep(ns, 1e-3)
        hc.substep(ns, 1e-3, CTHECD)
        dev = hc.download()
        ref = O.mean_xz_io(og, {"q": dev["q"], "g": dev["g"], "pr": dev["pr"], "T": dev["temperature"],
                                "rho": dev["density"], "h": dev["enthalpy"], "mu": dev["viscosity"]})
        got = hc.mean_xz_io()
        sc = np.maximum(np.abs(ref).max(axis=0), 1e-300)
        err = float((np.abs(got - ref) / sc).max())
        # and the trajectory itself stays with the oracle's
        traj = float((np.abs(got - O.mean_xz_io(og, st)) / sc).max())
        print(ns, "mean_xz_io rel", err, "vs oracle trajectory", traj, flush=True)
        assert err < 1e-13 and traj < 1e-10
        acc, ns_o = O.mean_t_update(acc, ref, ns_o, tp=1)
        tm, nst = s.mean_t("io", tp=1)
        assert nst == ns_o
        assert float((np.abs(tm - acc) / sc).max()) < 1e-13
    s.mean_t_reset("io")
    hc.mean_xz_io()
    tm, nst = s.mean_t("io", tp=2, pp_instn_sz=4)
    assert nst == 1 and float((np.abs(tm - ref / 4) / sc).max()) < 1e-13
    # tg domain: chs_mean_xz feeds PP_MEAN_T_TG
    s.mean_t_reset("tg")
    m1 = s.mean_xz()
    tm, nst = s.mean_t("tg", tp=1)
    assert nst == 1 and np.array_equal(tm, m1)
    pt.close()
    s.close()


@pytest.mark.parametrize("dims", [(256, 24, 16), (64, 24, 16)])
def test_heated_on_mixed_handle(scaled, dims):
    """The heated channel on a mixed-precision handle (FP64 io stencils, the
    Poisson solve FP32 refined against an FP64 residual; NCL1 = 256: the
    fused pair of the tg substep, checked) against the oracle: after three
    substeps max|div G + DrhoDtP| <= 1e-10 and every field within what the
    refined solve leaves (the pressure carries the Poisson tolerance)."""
    from chapsim1ww_amd.heated import HeatedChannel
    from chapsim1ww_amd.solver import ChannelSolver
    print(f"heated on mixed handle {dims}", flush=True)
    g = ChannelGrid(*dims)
    og = O.OracleGrid(*dims)
    st, tab, n = O.io_case(og, scaled, seed=77)
    s = ChannelSolver(g, "mixed")
    pt = props.PropertyTable(scaled)
    hc = HeatedChannel(s, pt)
    hc.upload({MAP[k]: v for k, v in st.items()})
    io = O.OracleIO(og, st)
    rmax = O.rhoh_max(scaled)
    dt = 1e-3
    pr_scale = np.abs(st["pr"]).max()
    for ns in (1, 2, 3):
        assert O.energy_substep(og, st, tab, n, ns, dt, CTHECD, rmax) == 0
        io.prepare(ns)
        io.substep(ns, dt)
        hc.substep(ns, dt, CTHECD)
        pr_scale = max(pr_scale, np.abs(st["pr"]).max())
        dg, _ = hc.divgck_io()
        assert dg <= 1e-10, (ns, dg)
    pt.check()
    got = hc.download()
    for k in ("g", "T", "rho", "h"):
        e = _rel(got[MAP[k]], st[k])
        print(k, e, flush=True)
        assert e < 1e-11, (k, e)
    # PR relative to the largest |PR| of the trajectory (the start is far
    # from divergence-free, as in the FP64 test)
    e = float(np.abs(got["pr"] - st["pr"]).max() / pr_scale)
    print("pr", e, flush=True)
    assert e < 1e-8, e
    pt.close()
    s.close()
