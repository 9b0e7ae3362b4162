"""Parity of the HIP library (through its C ABI) with the CPU oracle.

Tolerances (FP64 plain form): max-norm relative error 1e-12 against the
oracle on the same inputs, as the reference's FP64 path is required to match.
The mixed form is checked on what it guarantees: the FP64 residual of the
refined Poisson solve and the projected divergence max|div u| <= 1e-10.
"""
import numpy as np
import pytest

import oracle as O
from chapsim1ww_amd import fields
from chapsim1ww_amd import native as nat
from chapsim1ww_amd.grid import ChannelGrid

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def _dev(a):
    return torch.from_numpy(np.asfortranarray(a).ravel(order="F").copy()).to("cuda")


def _host(t, shape):
    return t.detach().cpu().numpy().reshape(shape, order="F")


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def solver_mod():
    from chapsim1ww_amd import solver
    return solver


# --------------------------------------------------------------------------
# batched tridiagonal solves (the four TDMAIJ? solvers)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("n,nl", [(192, 65536), (7, 5), (12, 3), (256, 1000), (64, 129), (200, 77), (3, 4)])
@pytest.mark.parametrize("cyclic", [False, True])
@pytest.mark.parametrize("march", ["IJJ", "IJI"])
def test_tdma_batched(solver_mod, n, nl, cyclic, march):
    print(f"tdma n={n} nlines={nl} cyclic={cyclic} {march}", flush=True)
    rng = np.random.default_rng(n * 1000 + nl)
    if march == "IJJ":
        shape, es, ls = (nl, n), nl, 1
    else:
        shape, es, ls = (n, nl), 1, n
    a = np.asfortranarray(rng.uniform(-0.9, -0.1, shape))
    c = np.asfortranarray(rng.uniform(-0.9, -0.1, shape))
    b = np.asfortranarray(2.0 + np.abs(rng.standard_normal(shape)))
    d = np.asfortranarray(rng.standard_normal(shape))
    bc = None if cyclic else np.asfortranarray(rng.standard_normal((nl, 2)))
    ref = d.copy(order="F")
    if cyclic:
        O.tdma_cyc(a, b, c, ref, n, nl, es, ls)
    else:
        O.tdma_noncyc(a, b, c, ref, bc, n, nl, es, ls)
    da, db, dc, dd = (_dev(v) for v in (a, b, c, d))
    dbc = None if bc is None else _dev(bc)
    solver_mod.tdma_batched(da, db, dc, dd, dbc, n, nl, es, ls, cyclic)
    torch.cuda.synchronize()
    out = _host(dd, shape)
    assert _rel(out, ref) < FP64_TOL


def test_tdma_golden_fixtures(solver_mod):
    """The reference's own solvers' outputs (tests/golden/tdma_ref.npz, written
    from src/TDMAIJ{I,J}_{CYC,nonCYC}.f90 compiled from source): all four
    solvers on 4x7, 3x12 and 6x5 systems through chs_tdma_batched."""
    import os
    gold = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tdma_ref.npz"))
    keys = sorted({k.rsplit("_", 1)[0] for k in gold.files})
    assert len(keys) == 12
    for key in keys:
        print(f"tdma golden {key}", flush=True)
        d = {n: np.array(gold[f"{key}_{n}"], order="F") for n in ("a", "b", "c", "r", "bc", "x")}
        isz, jsz = d["a"].shape
        n, nl, es, ls = (jsz, isz, isz, 1) if key.startswith("IJJ") else (isz, jsz, 1, isz)
        cyclic = "_cyc_" in key
        da, db, dc, dd = (_dev(d[v]) for v in ("a", "b", "c", "r"))
        dbc = None if cyclic else _dev(d["bc"])
        solver_mod.tdma_batched(da, db, dc, dd, dbc, n, nl, es, ls, cyclic)
        torch.cuda.synchronize()
        assert _rel(_host(dd, (isz, jsz)), d["x"]) < FP64_TOL, key


def test_tdma_reference_wrappers(solver_mod):
    """The reference-named wrappers map their Fortran argument lists correctly."""
    rng = np.random.default_rng(5)
    isz, jsz = 33, 24
    a = np.asfortranarray(rng.uniform(-0.9, -0.1, (isz, jsz)))
    c = np.asfortranarray(rng.uniform(-0.9, -0.1, (isz, jsz)))
    b = np.asfortranarray(3.0 + rng.random((isz, jsz)))
    d = np.asfortranarray(rng.standard_normal((isz, jsz)))
    for name, oracle_args in (("tdmaijj_noncyc", (jsz, isz, isz, 1)), ("tdmaiji_noncyc", (isz, jsz, 1, isz)),
                              ("tdmaijj_cyc", (jsz, isz, isz, 1)), ("tdmaiji_cyc", (isz, jsz, 1, isz))):
        ref = d.copy(order="F")
        n, nl, es, ls = oracle_args
        bcn = np.asfortranarray(rng.standard_normal((nl, 2)))
        if "noncyc" in name:
            O.tdma_noncyc(a, b, c, ref, bcn, n, nl, es, ls)
        else:
            O.tdma_cyc(a, b, c, ref, n, nl, es, ls)
        dd = _dev(d)
        fn = getattr(solver_mod, name)
        if name == "tdmaijj_noncyc":
            fn(_dev(a), _dev(b), _dev(c), dd, _dev(bcn), 1, jsz, 1, isz)
        elif name == "tdmaiji_noncyc":
            fn(_dev(a), _dev(b), _dev(c), dd, _dev(bcn), 1, isz, 1, jsz)
        elif name == "tdmaijj_cyc":
            fn(_dev(a), _dev(b), _dev(c), dd, 1, jsz, 1, isz)
        else:
            fn(_dev(a), _dev(b), _dev(c), dd, 1, isz, 1, jsz)
        torch.cuda.synchronize()
        assert _rel(_host(dd, (isz, jsz)), ref) < FP64_TOL, name


# --------------------------------------------------------------------------
# Poisson solve
# --------------------------------------------------------------------------
@pytest.mark.parametrize("dims", [(8, 6, 8), (16, 12, 16), (12, 10, 20), (64, 64, 64), (128, 128, 128), (256, 192, 256)])
def test_poisson_fp64(solver_mod, dims):
    print(f"poisson fp64 {dims}", flush=True)
    g = ChannelGrid(*dims)
    og = O.OracleGrid(*dims)
    f = fields.poisson_rhs(g, seed=17)
    ref = O.poisson(og, f)
    s = solver_mod.ChannelSolver(g, "fp64")
    df = _dev(f)
    dph = torch.full((g.L * (g.M + 2) * g.N,), 7.0, dtype=torch.float64, device="cuda")
    s.fft99_pois3d(df, dph)
    torch.cuda.synchronize()
    out = _host(dph, g.shape_ghost)
    assert _rel(out, ref) < FP64_TOL
    assert np.all(out[:, 0, :] == 0.0) and np.all(out[:, -1, :] == 0.0)
    mx, _ = O.poisson_residual(og, out, f)
    assert mx < 1e-11 * np.abs(f).max()
    s.close()


@pytest.mark.parametrize("dims", [(16, 12, 16), (12, 10, 20), (64, 64, 64), (256, 192, 256)])
def test_poisson_mixed_refined(solver_mod, dims):
    print(f"poisson mixed {dims}", flush=True)
    g = ChannelGrid(*dims)
    og = O.OracleGrid(*dims)
    f = fields.poisson_rhs(g, seed=23)
    ref = O.poisson(og, f)
    s = solver_mod.ChannelSolver(g, "mixed")
    s.set_refinement(1e-11, 10)
    dph = torch.zeros(g.L * (g.M + 2) * g.N, dtype=torch.float64, device="cuda")
    s.fft99_pois3d(_dev(f), dph)
    torch.cuda.synchronize()
    out = _host(dph, g.shape_ghost)
    it, rel = s.poisson_stats()
    mx, _ = O.poisson_residual(og, out, f)
    assert rel <= 1e-11 and mx <= 1.01e-11 * np.abs(f).max(), (it, rel)
    assert it >= 1
    # the FP64 oracle solution to the accuracy the residual buys
    assert _rel(out, ref) < 1e-8
    s.close()


@pytest.mark.parametrize("dims", [(16, 12, 16), (256, 16, 16)])
def test_poisson_mixed_nan_reported(solver_mod, dims):
    """A NaN in f reaches the residual maxima (they propagate NaN), so the
    mixed solve reports it instead of passing its convergence check (NCL1 =
    256: through the pair with the residual formed in the x transform)."""
    print(f"poisson mixed nan {dims}", flush=True)
    g = ChannelGrid(*dims)
    f = fields.poisson_rhs(g, seed=3)
    f[2, 3, 4] = np.nan
    s = solver_mod.ChannelSolver(g, "mixed")
    dph = torch.zeros(g.L * (g.M + 2) * g.N, dtype=torch.float64, device="cuda")
    with pytest.raises(nat.ChsError, match="NaN"):
        s.fft99_pois3d(_dev(f), dph)
    torch.cuda.synchronize()
    s.close()


# --------------------------------------------------------------------------
# full RK3 substeps of the tg channel
# --------------------------------------------------------------------------
def _run_pair(dims, nsteps, seed=1973, precision="fp64", solver_mod=None, cn="auto", expect_path=None):
    g = ChannelGrid(*dims)
    og = O.OracleGrid(*dims)
    q, pr, c0 = fields.init_tg_fields(g, seed=seed)
    st = O.OracleState(og, q, pr, c0)
    dt = 0.5 / st.cfl()
    s = solver_mod.ChannelSolver(g, precision)
    s.set_cn_solver(cn)
    if expect_path is not None:
        for ns in (0, 1, 2, 3):
            info = s.cn_info(ns, dt)
            assert info["path"] == expect_path, info
    s.upload_state(q, pr, c0)
    out = []
    ns_seq = [0] + [1, 2, 3] * nsteps
    for ns in ns_seq:
        st.substep(ns, dt)
        s.solverrk3_mom_tg(ns, dt)
        qg, pg, cg = s.download_state()
        out.append((ns, (qg, pg, cg), (st.q.copy(order="F"), st.pr.copy(order="F"), st.convh0.copy(order="F")),
                    s.divgck(), st.divgck(), s.last_dpgrns(), st.dpgrns))
    s.close()
    return g, out


@pytest.mark.parametrize("dims,cn,path", [((32, 16, 16), "auto", "thomas"), ((64, 64, 64), "auto", "stencil"),
                                          ((64, 64, 64), "thomas", "thomas"), ((128, 96, 64), "marching", "stencil"),
                                          ((64, 64, 64), "marching", "stencil"), ((32, 40, 64), "marching", "stencil"),
                                          ((32, 48, 128), "marching", "stencil"), ((32, 43, 64), "marching", "stencil"),
                                          ((96, 48, 64), "marching", "stencil"), ((96, 48, 64), "rowtile", "stencil"),
                                          ((192, 40, 64), "marching", "stencil"), ((40, 24, 16), "auto", "thomas"),
                                          ((120, 24, 16), "thomas", "thomas")])
def test_substep_fp64(solver_mod, dims, cn, path):
    """7 substeps (NS = 0, then 1,2,3 twice) against the oracle; every state
    field within 1e-12 of the oracle after every substep.  NCL1 = 96, 192,
    40, 120 (factors 3 and 5: no power of two) run the same kernels with
    their thread indices split by division."""
    print(f"substep fp64 {dims} cn={cn}", flush=True)
    g, out = _run_pair(dims, 2, solver_mod=solver_mod, cn=cn, expect_path=path)
    for i, (ns, gpu, ref, dg, dr, pg, pr_) in enumerate(out):
        errs = {name: _rel(a, b) for name, a, b in zip(("Q", "PR", "CONVH0"), gpu, ref)}
        print(i, ns, errs, dg, flush=True)
        for name, e in errs.items():
            assert e < FP64_TOL, (i, ns, name, e)
        assert dg < 1e-10 and dr < 1e-10
        assert abs(pg - pr_) <= 1e-12 * max(abs(pr_), 1e-300) + 1e-18


@pytest.mark.parametrize("dims", [(256, 192, 256), (64, 48, 64)])
def test_rhs_marching_matches_rowtile(solver_mod, dims):
    """The y-marching RHS kernel against the per-row one on the same start
    state, through NS = 0, 1, 2, 3: the two evaluate the same expressions
    (only the block sums of the mass-flux integral group differently), so
    the states agree to a few ulps."""
    print(f"rhs marching vs rowtile {dims}", flush=True)
    g = ChannelGrid(*dims)
    q, pr, c0 = fields.init_tg_fields(g, seed=29)
    outs = []
    for mode in ("marching", "rowtile"):
        s = solver_mod.ChannelSolver(g)
        s.set_cn_solver(mode)
        s.upload_state(q, pr, c0)
        dt = 0.5 / s.cfl()
        assert s.cn_info(1, dt)["path"] == "stencil"
        for ns in (0, 1, 2, 3):
            s.solverrk3_mom_tg(ns, dt)
        outs.append(s.download_state())
        s.close()
    for name, a, b in zip(("Q", "PR", "CONVH0"), *outs):
        e = _rel(a, b)
        print(name, e, flush=True)
        assert e < 1e-13, (name, e)


@pytest.mark.parametrize("dims", [(64, 64, 64), (192, 40, 64)])
def test_substep_mixed_divergence(solver_mod, dims):
    """The mixed form (FP32 stencils and x/z solves, FP64 y elimination, FP32
    Poisson refined once with its residual checked in the projection): the
    projected divergence stays below 1e-10, the state close to the FP64
    oracle, and the check's verdict is read back (no error, residual known).
    NCL1 = 192: the Stockham x transforms and the stored FP32 residual."""
    g, out = _run_pair(dims, 1, precision="mixed", solver_mod=solver_mod)
    for ns, gpu, ref, dg, dr, *_ in out:
        print(ns, "div", dg, "Q", _rel(gpu[0], ref[0]), "PR", _rel(gpu[1], ref[1]), flush=True)
        assert dg <= 1e-10, (ns, dg)
        assert _rel(gpu[0], ref[0]) < 1e-6


def test_mixed_convh0_store_round_trip(solver_mod):
    """The mixed form keeps CONVH0 in FP32 between substeps.  Downloading the
    state widens it (exactly), uploading it back and going on narrows it
    again (exactly): the trajectory equals one without the round trip, bit
    for bit; the downloaded CONVH0 is FP32-representable and within FP32
    rounding of the FP64 form's after the same substeps."""
    g = ChannelGrid(64, 48, 64)
    q, pr, c0 = fields.init_tg_fields(g, seed=21)
    runs = []
    for trip in (False, True):
        s = solver_mod.ChannelSolver(g, "mixed")
        s.upload_state(q, pr, c0)
        dt = 0.5 / s.cfl()
        for ns in (0, 1, 2):
            s.solverrk3_mom_tg(ns, dt)
        if trip:
            st = s.download_state()
            h0 = st[2]
            assert np.array_equal(h0.astype(np.float32).astype(np.float64), h0)
            s.upload_state(*st)
        for ns in (3, 1, 2):
            s.solverrk3_mom_tg(ns, dt)
        runs.append(s.download_state())
        s.close()
    for a, b in zip(*runs):
        assert np.array_equal(a, b)
    s = solver_mod.ChannelSolver(g)
    s.upload_state(q, pr, c0)
    dt = 0.5 / s.cfl()
    for ns in (0, 1, 2, 3, 1, 2):
        s.solverrk3_mom_tg(ns, dt)
    ref = s.download_state()
    s.close()
    e = _rel(runs[0][2], ref[2])
    print("CONVH0 mixed vs fp64", e, flush=True)
    assert e < 1e-5


def test_mixed_check_reports_nan(solver_mod):
    """A NaN planted in the state reaches the mixed substep's fused residual
    check; the verdict, read at the next synchronising call, is an error."""
    g = ChannelGrid(64, 48, 64)
    q, pr, c0 = fields.init_tg_fields(g, seed=8)
    s = solver_mod.ChannelSolver(g, "mixed")
    s.upload_state(q, pr, c0)
    dt = 0.5 / s.cfl()
    s.solverrk3_mom_tg(0, dt)
    it, rel = s.poisson_stats()
    assert it == 1 and 0.0 <= rel < 1e-9, (it, rel)
    q2 = q.copy(order="F")
    q2[5, 7, 9, 0] = np.nan
    s.upload_state(q2, pr, c0)
    s.solverrk3_mom_tg(1, dt)
    # the host runs ahead of the device: a substep reads a verdict only if it
    # is ready, so the NaN surfaces at one of these calls (which one depends
    # on timing), kept on the device until a verdict is read
    with pytest.raises(nat.ChsError, match="NaN"):
        for ns in (2, 3, 1):
            s.solverrk3_mom_tg(ns, dt)
        s.poisson_stats()
    # reported once: a clean state checks clean again
    s.upload_state(q, pr, c0)
    s.solverrk3_mom_tg(0, dt)
    s.solverrk3_mom_tg(1, dt)
    it, rel = s.poisson_stats()
    assert 0.0 <= rel < 1e-9, rel
    s.close()


def test_divgck_cfl_and_nan(solver_mod):
    g = ChannelGrid(32, 24, 16)
    og = O.OracleGrid(32, 24, 16)
    q, pr, c0 = fields.init_tg_fields(g, seed=3)
    st = O.OracleState(og, q, pr, c0)
    s = solver_mod.ChannelSolver(g)
    s.upload_state(q, pr, c0)
    assert abs(s.divgck() - st.divgck()) <= 1e-14 * st.divgck()
    assert abs(s.cfl() - st.cfl()) <= 1e-14 * st.cfl()
    q2 = q.copy(order="F")
    q2[3, 5, 7, 1] = np.nan
    s.upload_state(q2, None, None)
    assert np.isnan(s.divgck()) and np.isnan(s.cfl())
    s.close()


MEAN_GROUPS = ([0, 1, 2], [3], list(range(4, 10)), [10, 11, 12])


def _mean_err(a, b):
    """Largest difference of two (NCL2, 13) moment arrays, each column scaled
    by the largest value of its group (<u_i>, <p>, <u_i u_j>, <u_i p>): the
    xz-mean of v vanishes analytically, so that column holds only round-off
    and is measured against the velocity scale.  An all-zero group (p = 0 on
    the start state) is compared absolutely."""
    err = 0.0
    for grp in MEAN_GROUPS:
        sc = np.abs(b[:, grp]).max()
        err = max(err, float(np.abs(a[:, grp] - b[:, grp]).max() / (sc if sc > 0 else 1.0)))
    return err


@pytest.mark.parametrize("dims", [(32, 24, 16), (64, 48, 32)])
def test_mean_xz(solver_mod, dims):
    """xz-plane moments (PP_MEAN_ZX_FLOW_TG) of the device state against the
    oracle: on the uploaded state (same fields, only the summation order
    differs) and after NS = 0, 1, 2, 3 (fields within 1e-12 of each other)."""
    print(f"mean_xz {dims}", flush=True)
    g = ChannelGrid(*dims)
    og = O.OracleGrid(*dims)
    q, pr, c0 = fields.init_tg_fields(g, seed=11)
    st = O.OracleState(og, q, pr, c0)
    s = solver_mod.ChannelSolver(g)
    s.upload_state(q, pr, c0)
    dt = 0.5 / st.cfl()

    def close(tol):
        a, b = s.mean_xz(), st.mean_xz()
        assert a.shape == (dims[1], 13)
        err = _mean_err(a, b)
        print("rel", err, flush=True)
        assert err < tol, err

    close(1e-14)
    for ns in (0, 1, 2, 3):
        st.substep(ns, dt)
        s.solverrk3_mom_tg(ns, dt)
    close(1e-12)
    s.close()


def test_trajectory_64cube_re2800(solver_mod):
    """The CPU-runnable channel case: 64^3, Re_b = 2800, seeded laminar +
    noise start, CFL 0.5, 100 RK3 steps (NS = 0 then 1,2,3 x 100), FP64 device
    path against the oracle.  Round-off differences grow along the
    trajectory; after 301 substeps the state and the xz-mean moments stay
    within 1e-10 of the oracle's."""
    print("trajectory 64^3 Re_b=2800, 100 steps", flush=True)
    dims = (64, 64, 64)
    g = ChannelGrid(*dims, ren=2800.0)
    og = O.OracleGrid(*dims, ren=2800.0)
    q, pr, c0 = fields.init_tg_fields(g, seed=1973)
    st = O.OracleState(og, q, pr, c0)
    dt = 0.5 / st.cfl()
    s = solver_mod.ChannelSolver(g)
    s.upload_state(q, pr, c0)
    for i, ns in enumerate([0] + [1, 2, 3] * 100):
        st.substep(ns, dt)
        s.solverrk3_mom_tg(ns, dt)
        if i % 60 == 0 or i == 300:
            (qg, pg, _), dg = s.download_state(), s.divgck()
            eq, ep = _rel(qg, st.q), _rel(pg, st.pr)
            a, b = s.mean_xz(), st.mean_xz()
            em = _mean_err(a, b)
            print(i, ns, f"Q {eq:.2e} PR {ep:.2e} means {em:.2e} div {dg:.2e}", flush=True)
            assert eq < 1e-10 and ep < 1e-10 and em < 1e-10, (i, eq, ep, em)
            assert dg < 1e-10
    s.close()


def test_instant_files_restart_device_run(solver_mod, tmp_path):
    """WRT_INSTANT_VARS_TG of a device run after a whole RK3 step, then
    RESTART_INSTANT_VARS_TG into a fresh handle: both continue with
    NS = 1, 2, 3 to the same bits (NS = 1 does not read CONVH0, so the files
    hold the whole state a restart needs)."""
    print("instant files restart", flush=True)
    g = ChannelGrid(64, 48, 32)
    q, pr, c0 = fields.init_tg_fields(g, seed=31)
    a = solver_mod.ChannelSolver(g)
    a.upload_state(q, pr, c0)
    dt = 0.5 / a.cfl()
    for ns in (0, 1, 2, 3):
        a.solverrk3_mom_tg(ns, dt)
    d = str(tmp_path) + "/"
    paths = a.write_instant_vars_tg(4, 4 * dt, dt, d)
    assert len(paths) == 4
    b = solver_mod.ChannelSolver(g)
    it, t = b.restart_instant_vars_tg(4 * dt, d)
    assert it == 4 and t == 4 * dt
    for ns in (1, 2, 3):
        a.solverrk3_mom_tg(ns, dt)
        b.solverrk3_mom_tg(ns, dt)
    for x, y in zip(a.download_state(), b.download_state()):
        assert np.array_equal(x, y)
    a.close()
    b.close()


@pytest.mark.parametrize("flowdrvtp,nflow", [(2, 1), (2, 3), (1, 3)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_flow_drive_substeps(solver_mod, flowdrvtp, nflow, precision):
    """FLOWDRVTP = 2 (constant pressure gradient -CFGV/2) and NFLOW = 3
    (spanwise streamwise direction) on the tg substep: NS = 0, 1, 2, 3 against
    the oracle (FP64 within 1e-12; the mixed form within its FP32 stencils'
    1e-6 of the velocities and max|div u| <= 1e-10), and DPGRNS equal to the
    oracle's after every substep."""
    print(f"flow drive flowdrvtp={flowdrvtp} nflow={nflow} {precision}", flush=True)
    dims = (64, 48, 32)
    g = ChannelGrid(*dims)
    og = O.OracleGrid(*dims)
    q, pr, c0 = fields.init_tg_fields(g, seed=41 + nflow)
    if nflow == 3:   # the laminar profile along z
        q[..., 0], q[..., 2] = q[..., 2].copy(), q[..., 0].copy()
    st = O.OracleState(og, q, pr, c0)
    cfgv = 0.0163
    s = solver_mod.ChannelSolver(g, precision, flowdrvtp=flowdrvtp, nflow=nflow, cfgv=cfgv)
    s.upload_state(q, pr, c0)
    dt = 0.5 / st.cfl()
    for ns in (0, 1, 2, 3):
        st.substep(ns, dt, flowdrvtp=flowdrvtp, nflow=nflow, cfgv=cfgv)
        s.solverrk3_mom_tg(ns, dt)
        gq, gp, gc = s.download_state()
        errs = {n: _rel(a, b) for n, a, b in zip(("Q", "PR", "CONVH0"), (gq, gp, gc), (st.q, st.pr, st.convh0))}
        print(ns, errs, s.last_dpgrns(), st.dpgrns, flush=True)
        if precision == "fp64":
            for n, e in errs.items():
                assert e < FP64_TOL, (ns, n, e)
            assert abs(s.last_dpgrns() - st.dpgrns) <= 1e-12 * abs(st.dpgrns) + 1e-18
        else:
            assert errs["Q"] < 1e-6, errs
            assert abs(s.last_dpgrns() - st.dpgrns) <= 1e-6 * abs(st.dpgrns) + 1e-12
        if flowdrvtp == 2:
            from chapsim1ww_amd.grid import TALP
            assert st.dpgrns == -0.5 * cfgv * (TALP[ns] * dt)
        assert s.divgck() <= 1e-10
    s.close()
