"""Throughput of the per-RK3-substep solver core at 256x192x256 on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one RK3 substep (SOLVERRK3_MOM_tg) of the tg channel at
256x192x256 (Re_b = 6880, seeded laminar + noise start) on the
device-resident FP64 state; the timed region holds exactly K substeps
(NS cycling 1, 2, 3), bracketed by device synchronisation.  The JSON line
also carries Poisson solves/s and batched-TDMA rates, the mixed-precision
form (both forms continued to NS = 0 + 500 RK3 steps, then their fields and
xz-mean profiles compared), the per-kernel HBM roofline of the dominant
kernel (HIP events on the library's stream), and the CPU oracle timed on this
host: a bounded sample at 256x192x256 and the whole 64^3, Re_b = 2800,
100-step case beside the device path over the same trajectory.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
CFG = dict(L=256, M=192, N=256, ren=6880.0)

# algorithmic bytes per cell and launch (FP64 fields x 8 B; half-complex rows
# count L reals), the figures behind roofline.achieved
def alg_bytes_per_cell(L, precision="fp64"):
    sp = 16.0 * (L // 2 + 1) / L        # one spectral sweep (complex, LH per L)
    if precision == "mixed":
        sp32 = sp / 2.0
        return {"resid_fft_x": 12 + sp32,  # read phi1 (FP32), f; write S
                "fft_x_fwd": 8 + sp32, "fft_z": 2 * sp32, "poisson_y": 2 * sp32,
                "fft_x_inv": sp32 + 8,     # read S; write phi1 (FP32) / read phi1, write phi: 4 and 12, averaged
                "poisson_residual": 24.0,
                "rhs_x": 68.0,             # read Q x3, PR, CONVH0 x3 (FP32); write RHS x3, CONVH0 x3 (FP32)
                "z_solve": 24.0, "y_solve": 60.0, "divg": 32.0, "update": 80.0}
    return {"rhs_x": 104.0,          # read Q x3, PR, CONVH0 x3; write RHS x3, CONVH0 x3
            "z_solve": 48.0,         # read/write RHS x3
            "y_solve": 72.0,         # read RHS x3, Q x3; write Q x3
            "divg": 32.0,            # read Q x3; write f
            "fft_x_fwd": 8 + sp,     # read f; write S
            "fft_z": 2 * sp,         # read/write S
            "poisson_y": 2 * sp,     # read/write S
            "fft_x_inv": sp + 8,     # read S; write phi
            "update": 80.0}          # read phi, f, Q x3, PR; write Q x3, PR


def _note(msg):
    """Progress on stderr (stdout carries only the JSON line)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def _sync():
    import torch
    torch.cuda.synchronize()


def run_substeps(solver, dt, nsteps, ns0=0):
    seq = [1, 2, 3]
    for i in range(nsteps):
        solver.solverrk3_mom_tg(seq[(ns0 + i) % 3], dt)


def gpu_bench(steps, warmup, cmp_steps):
    import torch
    from chapsim1ww_amd import fields
    from chapsim1ww_amd.grid import ChannelGrid
    from chapsim1ww_amd.solver import ChannelSolver, tdma_batched

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    g = ChannelGrid(CFG["L"], CFG["M"], CFG["N"], ren=CFG["ren"])
    q, pr, c0 = fields.init_tg_fields(g, seed=1973)
    res = {}
    ncell = g.ncell

    for prec in ("fp64", "mixed"):
        s = ChannelSolver(g, prec, stream=stream)
        s.upload_state(q, pr, c0)
        rate = s.cfl()
        dt = 0.5 / rate
        s.solverrk3_mom_tg(0, dt)      # the reference's initial NS=0 step
        run_substeps(s, dt, warmup)
        _sync()
        t0 = time.perf_counter()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        run_substeps(s, dt, steps, warmup)
        e1.record(stream)
        _sync()
        wall = time.perf_counter() - t0
        ms = e0.elapsed_time(e1) / steps
        div = s.divgck()
        # per-kernel durations with HIP events over a second region of K substeps
        s.profile(True)
        s.profile_reset()
        run_substeps(s, dt, steps, warmup + steps)
        _sync()
        prof = s.profile_read()
        s.profile(False)
        iters = s.poisson_stats()
        # continue both forms to the same substep count (NS = 0, then cmp_steps
        # RK3 steps) for the mixed-vs-FP64 comparison
        done = warmup + 2 * steps
        if 3 * cmp_steps > done:
            run_substeps(s, dt, 3 * cmp_steps - done, done)
        nsub = 1 + max(3 * cmp_steps, done)
        state = s.download_state()
        means = s.mean_xz()
        _note(f"{prec}: {ms:.4f} ms/substep, compared after {nsub} substeps")
        end = dict(substeps=nsub, maxdiv=s.divgck(), cfl=s.cfl() * dt)
        # Poisson solves/s on the handle's workspace
        f = torch.from_numpy(np.asfortranarray(fields.poisson_rhs(g, seed=9)).ravel(order="F").copy()).to(dev)
        dph = torch.zeros(g.L * (g.M + 2) * g.N, dtype=torch.float64, device=dev)
        if prec == "mixed":
            s.set_refinement(1e-10, 10)
        for _ in range(3):
            s.fft99_pois3d(f, dph)
        _sync()
        npois = max(20, steps)
        p0, p1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        p0.record(stream)
        for _ in range(npois):
            s.fft99_pois3d(f, dph)
        p1.record(stream)
        _sync()
        pois_ms = p0.elapsed_time(p1) / npois
        res[prec] = dict(ms_per_substep=ms, wall_s=wall, maxdiv=div, dt=dt, prof=prof,
                         poisson_ms=pois_ms, poisson_refine=s.poisson_stats(), substep_refine=iters,
                         cn={ns: s.cn_info(ns, dt) for ns in (1, 2, 3)}, state=state,
                         means=means, end=end)
        s.close()
        del f, dph

    # batched TDMA, 65536 lines x 192 (lines interleaved: IJJ form)
    n, nl = 192, 65536
    rng = np.random.default_rng(2)
    a = torch.from_numpy(rng.uniform(-0.9, -0.1, n * nl)).to(dev)
    c = torch.from_numpy(rng.uniform(-0.9, -0.1, n * nl)).to(dev)
    b = torch.from_numpy(2.0 + np.abs(rng.standard_normal(n * nl))).to(dev)
    d0 = torch.from_numpy(rng.standard_normal(n * nl)).to(dev)
    r = d0.clone()
    tdma = {}
    for cyc in (False, True):
        for _ in range(3):
            r.copy_(d0)
            tdma_batched(a, b, c, r, None, n, nl, nl, 1, cyc, stream)
        _sync()
        reps = 50
        t_tot = 0.0
        for _ in range(reps):
            r.copy_(d0)
            q0, q1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            q0.record(stream)
            tdma_batched(a, b, c, r, None, n, nl, nl, 1, cyc, stream)
            q1.record(stream)
            q1.synchronize()
            t_tot += q0.elapsed_time(q1)
        ms_t = t_tot / reps
        tdma["cyclic" if cyc else "noncyclic"] = dict(ms=ms_t, batches_per_s=1e3 / ms_t,
                                                      gbs=40.0 * n * nl / (ms_t * 1e-3) / 1e9)
    res["tdma"] = tdma
    res["ncell"] = ncell
    _note("tdma done")
    res["props"] = props_bench(g, res["fp64"]["state"][0], stream, steps)
    _note("properties done")
    return res


def _heated_start(g, sc, q_fp64, stream):
    """A heated-channel state on g: T = 1.2 - 0.3 y^2 + a small x-z wave
    mapped to h through the table, mass fluxes from the FP64 tg state,
    q_w = 0.1 at both walls; returns (solver, table, channel) per precision."""
    import torch
    from chapsim1ww_amd import props
    from chapsim1ww_amd.heated import HeatedChannel
    from chapsim1ww_amd.solver import ChannelSolver
    dev = torch.device("cuda:0")
    L, M, N = g.L, g.M, g.N
    yc = np.concatenate([[-1.0], np.asarray(g.ycc), [1.0]])
    xx = np.arange(L) / L * 2 * np.pi
    zz = np.arange(N) / N * 2 * np.pi
    T3 = 1.2 - 0.3 * yc[None, :, None] ** 2 + 0.02 * np.sin(xx)[:, None, None] * np.cos(zz)[None, None, :]
    hg = torch.from_numpy(np.interp(T3.ravel(order="F"), sc["T"], sc["H"])).to(dev)
    qd = torch.from_numpy(np.asfortranarray(q_fp64).ravel(order="F")).to(dev)
    out = {}
    for prec in ("fp64", "mixed"):
        s = ChannelSolver(g, prec, stream=stream)
        pt = props.PropertyTable(sc, stream=stream)
        ts = HeatedChannel(s, pt, thermal=1, nflow=1, flowdrvtp=1, gravdir=0)
        f = ts.dev
        f["enthalpy"].copy_(hg)
        pt.update(hg, f["temperature"], f["density"], f["viscosity"], f["thermcondt"])
        f["rhoh"].copy_(hg * f["density"])
        for k in ("density0", "density1"):
            f[k].copy_(f["density"])
        f["viscosity0"].copy_(f["viscosity"])
        f["g"].copy_(qd)
        f["g0"].copy_(qd)
        f["wallflux"].fill_(0.1)
        out[prec] = (s, pt, ts)
    return out


def props_bench(g, q_fp64, stream, steps):
    """The heated channel at 256x192x256 on the synthetic table: the whole
    heated substep (FP64 and the mixed-precision solver side by side, their
    difference stated), THERM_PROP_UPDATE by table and by the surrogate
    trained on the heated solver's own ENTHALPY, and the energy substep with
    either property update."""
    import torch
    from chapsim1ww_amd import props
    dev = torch.device("cuda:0")
    sc = props.scale_table(props.synthetic_table())
    n = g.ncell
    L, M, N = g.L, g.M, g.N
    twins = _heated_start(g, sc, q_fp64, stream)
    s, pt, ts = twins["fp64"]
    out = {}

    def timed(fn, reps):
        fn()
        _sync()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        _sync()
        return e0.elapsed_time(e1) / reps

    reps = max(10, steps)
    seq = [1, 2, 3]
    cthecd, dt = 1.0 / 2800.0, 1e-3
    # three substeps of both forms from the same start: the solver snapshot
    for i in range(3):
        for (_, _, t_) in twins.values():
            t_.substep(seq[i % 3], dt, cthecd)
    # the snapshot's interior ENTHALPY: the property workload and the training samples
    hs = ts.dev["enthalpy"].view(N, M + 2, L)[:, 1:M + 1, :].reshape(-1).contiguous()
    outs = [torch.empty_like(hs) for _ in range(5)]
    ms_t = timed(lambda: pt.update(hs, *outs), reps)
    out["table"] = dict(ms=round(ms_t, 4), cells_per_s=round(n / (ms_t * 1e-3), 1),
                        gbs=round(48.0 * n / (ms_t * 1e-3) / 1e9, 1), field="heated-solver ENTHALPY after 3 substeps")
    ref = [o.clone() for o in outs]
    ms_t4 = timed(lambda: pt.update(hs, *outs[:4]), reps)
    gen = torch.Generator(device=dev).manual_seed(3)
    samples = hs[torch.randint(0, n, (1 << 20,), device=dev, generator=gen)].contiguous()
    t0 = time.perf_counter()
    pt.train(samples, 0, seed=11)
    loss = pt.train(samples, 40000, 16384, 2e-3, seed=11)
    _sync()
    t_train = time.perf_counter() - t0
    c3 = pt.certify(1e-3, 2048)
    pt.certify(1e-4, 2048)     # the 1e-4 certificate stays in force
    c5, c4 = pt.certified()
    nf = pt.update(hs, *outs, mode="surrogate")
    err = max(float(((o - r).abs() / r.abs()).max()) for o, r in zip(outs, ref))
    ms_s = timed(lambda: pt.update(hs, *outs, mode="surrogate"), reps)
    nf4 = pt.update(hs, *outs[:4], mode="surrogate")
    err4 = max(float(((o - r).abs() / r.abs()).max()) for o, r in zip(outs[:4], ref[:4]))
    ms_s4 = timed(lambda: pt.update(hs, *outs[:4], mode="surrogate"), reps)
    # the network alone on every cell (what a fully certified field would cost)
    net_out = torch.empty(5 * n, dtype=torch.float64, device=dev)
    ms_n = timed(lambda: pt.eval(hs, net_out), max(3, reps // 3))
    del net_out
    out["surrogate"] = {"ms": round(ms_s, 4), "gain_vs_table": round(ms_t / ms_s, 3), "train_s": round(t_train, 2),
                        "train_steps": 40000, "train_samples": "2^20 of the heated solver's ENTHALPY",
                        "final_loss": loss, "certified_fraction_tol1e-4": c5, "certified_fraction_tol1e-3": c3,
                        "certified_fraction_tol1e-4_without_cp": c4, "fallback_fraction": round(nf / n, 4),
                        "max_rel_err_vs_table": err,
                        "without_cp": {"table_ms": round(ms_t4, 4), "surrogate_ms": round(ms_s4, 4),
                                       "gain_vs_table": round(ms_t4 / ms_s4, 3), "fallback_fraction": round(nf4 / n, 4),
                                       "max_rel_err_vs_table": err4},
                        "network_only_ms": round(ms_n, 4),
                        # model FLOPs (8960 per cell) and the matrix-core FLOPs executed (three BF16
                        # products per weight block, layer 3 padded to 16 rows: 30720 per cell)
                        "network_model_tflops": round(8960.0 * n / (ms_n * 1e-3) / 1e12, 2),
                        "network_mfma_tflops": round(30720.0 * n / (ms_n * 1e-3) / 1e12, 2),
                        "mfma_peak_bf16_tflops": 2500.0}
    cnt = [0]

    def eng():
        ts.solverrk3_eng_io(seq[cnt[0] % 3], dt, cthecd)
        cnt[0] += 1
    ms_e = timed(eng, reps)
    sur_e = None
    if c4 > 0.0:
        pt.set_mode("surrogate")
        ms_es = timed(eng, reps)
        sur_e = dict(ms=round(ms_es, 4), fallback_cells=pt.last_fallback())
        pt.set_mode("table")
    out["energy_substep"] = dict(ms=round(ms_e, 4), substeps_per_s=round(1e3 / ms_e, 2), surrogate_mode=sur_e,
                                 note="SOLVERRK3_ENG_IO: RHS + RK3 + property cascade + iso-flux walls; "
                                      "its range check is read by the next call")
    # the whole heated-channel substep in SOLVE's order: energy, property
    # staging, io momentum (explicit viscous terms, FFT99 projection)

    def heated():
        ts.substep(seq[cnt[0] % 3], dt, cthecd)
        cnt[0] += 1

    def mom():
        ns = seq[cnt[0] % 3]
        ts.io_prepare(ns)
        ts.solverrk3_mom_io(ns, dt)
        cnt[0] += 1
    cnt[0] = 0
    ms_h = timed(heated, reps)
    ms_m = timed(mom, reps)
    s.profile(True)
    s.profile_reset()
    for _ in range(reps):
        mom()
    _sync()
    prof = s.profile_read()
    s.profile(False)
    dg, _ = ts.divgck_io()
    # the RK-CN branch (visthemflg = 1) of the momentum part on the same state
    ts.params.visimplicit = 1
    ms_cn = timed(mom, reps)
    s.profile(True)
    s.profile_reset()
    for _ in range(reps):
        mom()
    _sync()
    prof_cn = s.profile_read()
    s.profile(False)
    dg_cn, _ = ts.divgck_io()
    ts.params.visimplicit = 0
    # algorithmic bytes per cell of the explicit branch with NFLOW = 1: the
    # terms kernel also writes G0 and G + RHS (no separate copy or mass-flux pass)
    ab = {"io_prepare": 160.0, "io_divu": 32.0, "io_terms": 200.0, "io_massflux": 72.0, "io_divg": 40.0,
          "io_update": 72.0}
    kern = {}
    for name, (ms_tot, calls) in prof.items():
        if calls:
            kern[name] = dict(avg_ms=round(ms_tot / calls, 5), calls=calls)
            if name in ab:
                kern[name]["alg_gbs"] = round(ab[name] * n / (ms_tot / calls * 1e-3) / 1e9, 1)
    out["heated_substep"] = dict(ms=round(ms_h, 4), substeps_per_s=round(1e3 / ms_h, 2),
                                 momentum_ms=round(ms_m, 4), max_div_g=dg, kernels=kern,
                                 momentum_cn_ms=round(ms_cn, 4), max_div_g_cn=dg_cn,
                                 kernels_cn={k: round(v[0] / v[1], 5) for k, v in prof_cn.items() if v[1]},
                                 note="SOLVE order: SOLVERRK3_ENG_IO, DENSITY_TIME_DERIVATION + DENSITY/MU_Staggered"
                                      " + VELOCITY_CALC_io, SOLVERRK3_MOM_IO (explicit viscous, FFT99 projection)")
    _note("heated substep done")
    # the mixed-precision heated form beside the FP64 one: both continued
    # from the common snapshot (3 substeps) by the same 30 substeps
    sm, ptm, tsm = twins["mixed"]
    s64b = _heated_start(g, sc, q_fp64, stream)["fp64"]
    s64b_s, s64b_p, s64b_t = s64b
    for i in range(3):
        s64b_t.substep(seq[i % 3], dt, cthecd)
    cnt[0] = 0

    def heated_m():
        tsm.substep(seq[cnt[0] % 3], dt, cthecd)
        cnt[0] += 1
    ms_hm = timed(heated_m, reps)
    nmix = 1 + reps
    for i in range(nmix):
        s64b_t.substep(seq[i % 3], dt, cthecd)
    dgm, _ = tsm.divgck_io()
    a, b = s64b_t.download(), tsm.download()
    rel = lambda x, y: float(np.abs(x - y).max() / max(np.abs(x).max(), 1e-300))
    errs = {k: float(f"{rel(a[k], b[k]):.3e}") for k in ("temperature", "density", "enthalpy", "g", "pr")}
    mx = {k: float(f"{v:.3e}") for k, v in zip(
        ("T_mean", "rho_mean", "h_mean"),
        [rel(ma[:, c], mb[:, c]) for ma, mb in [(s64b_t.mean_xz_io(), tsm.mean_xz_io())] for c in (22, 23, 24)])}
    out["heated_mixed"] = dict(ms=round(ms_hm, 4), speedup_vs_fp64=round(ms_h / ms_hm, 3), max_div_g=dgm,
                               substeps_compared=3 + nmix, err_vs_fp64=errs, mean_profile_rel_vs_fp64=mx,
                               note="heated substep on the mixed-precision handle (FP32 Poisson refined, "
                                    "FP64 io stencils) vs the FP64 handle from the same snapshot")
    _note("heated mixed done")
    for (s_, p_, _) in list(twins.values()) + [s64b]:
        s_.close()
        p_.close()
    # the variable-property heated channel's own grid: 128^3, Re = 2800,
    # the same table and q_w = 0.1, mass fluxes from the seeded tg start
    from chapsim1ww_amd import fields
    from chapsim1ww_amd.grid import ChannelGrid
    g3 = ChannelGrid(128, 128, 128, ren=2800.0)
    q3, _, _ = fields.init_tg_fields(g3, seed=1973)
    t3 = _heated_start(g3, sc, q3, stream)
    s3, p3, ts3 = t3["fp64"]
    cnt3 = [0]

    def heated3():
        ts3.substep(seq[cnt3[0] % 3], 1e-3, 1.0 / 2800.0)
        cnt3[0] += 1
    ms3 = timed(heated3, reps)
    dg3, _ = ts3.divgck_io()
    out["heated_substep_128cube"] = dict(ms=round(ms3, 4), substeps_per_s=round(1e3 / ms3, 1),
                                         grid_point_substeps_per_s=round(g3.ncell * 1e3 / ms3, 1), max_div_g=dg3,
                                         note="FP64 energy + staging + io momentum at 128^3, Re = 2800")
    for (s_, p_, _) in t3.values():
        s_.close()
        p_.close()
    _note("heated 128^3 done")
    return out


def cpu_baseline(max_seconds=30.0):
    """The CPU oracle (C, OpenMP) on this host at the same workload: a bounded
    sample of substeps and Poisson solves."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from chapsim1ww_amd import fields
    from chapsim1ww_amd.grid import ChannelGrid

    g = ChannelGrid(CFG["L"], CFG["M"], CFG["N"], ren=CFG["ren"])
    og = O.OracleGrid(CFG["L"], CFG["M"], CFG["N"], ren=CFG["ren"])
    q, pr, c0 = fields.init_tg_fields(g, seed=1973)
    st = O.OracleState(og, q, pr, c0)
    dt = 0.5 / st.cfl()
    st.substep(0, dt)
    n, t0 = 0, time.perf_counter()
    while True:
        st.substep(1 + n % 3, dt)
        n += 1
        el = time.perf_counter() - t0
        if el > max_seconds / 2 or n >= 3:
            break
    sub_rate = n / el
    f = fields.poisson_rhs(g, seed=9)
    m, t1 = 0, time.perf_counter()
    while True:
        O.poisson(og, f)
        m += 1
        el2 = time.perf_counter() - t1
        if el2 > max_seconds / 4 or m >= 5:
            break
    nthr = O.num_threads()
    # the same substep on one thread (one substep: about 10-25 s of CPU work)
    O.set_num_threads(1)
    t2 = time.perf_counter()
    st.substep(1 + n % 3, dt)
    one = 1.0 / (time.perf_counter() - t2)
    O.set_num_threads(nthr)
    return dict(substeps_per_s=sub_rate, nsub=n, poisson_per_s=m / el2, npois=m, cores=nthr,
                substeps_per_s_1thread=one, host=host_cpu())


def host_cpu():
    """CPU model, logical CPUs of the host and CPUs this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return dict(model=model, logical_cpus=os.cpu_count(), usable_cpus=len(os.sched_getaffinity(0)),
                omp_num_threads=os.environ.get("OMP_NUM_THREADS"))


MEAN_COLS = {"u": 0, "v": 1, "w": 2, "p": 3, "uu": 4, "uv": 5, "vv": 7, "ww": 9}


def mean_err(a, b):
    """Largest difference of two (NCL2, 13) xz-moment arrays, each column
    scaled by the largest value of its group (<u_i>, <p>, <u_i u_j>,
    <u_i p>; <v> vanishes analytically and holds only round-off)."""
    err = 0.0
    for grp in ([0, 1, 2], [3], list(range(4, 10)), [10, 11, 12]):
        sc = np.abs(b[:, grp]).max()
        err = max(err, float(np.abs(a[:, grp] - b[:, grp]).max() / (sc if sc > 0 else 1.0)))
    return err


def small_case(steps=100):
    """The CPU-runnable channel case (64^3, Re_b = 2800, CFL 0.5, NS = 0 then
    `steps` RK3 steps): the oracle and the FP64 device path over the same
    trajectory, each timed, and their difference at the end."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import torch
    from chapsim1ww_amd import fields
    from chapsim1ww_amd.grid import ChannelGrid
    from chapsim1ww_amd.solver import ChannelSolver

    g = ChannelGrid(64, 64, 64, ren=2800.0)
    og = O.OracleGrid(64, 64, 64, ren=2800.0)
    q, pr, c0 = fields.init_tg_fields(g, seed=1973)
    st = O.OracleState(og, q, pr, c0)
    dt = 0.5 / st.cfl()
    seq = [0] + [1, 2, 3] * steps
    t0 = time.perf_counter()
    for ns in seq:
        st.substep(ns, dt)
    cpu_s = time.perf_counter() - t0
    s = ChannelSolver(g, "fp64", stream=torch.cuda.current_stream())
    s.upload_state(q, pr, c0)
    s.solverrk3_mom_tg(0, dt)
    _sync()
    t1 = time.perf_counter()
    for ns in seq[1:]:
        s.solverrk3_mom_tg(ns, dt)
    _sync()
    gpu_s = time.perf_counter() - t1
    qg, pg, _ = s.download_state()
    a, b = s.mean_xz(), st.mean_xz()
    rel = lambda x, y: float(np.abs(x - y).max() / np.abs(y).max())
    _note(f"64^3 case: oracle {cpu_s:.1f} s, device {gpu_s:.2f} s")
    out = dict(grid="64x64x64", ren=2800.0, substeps=len(seq),
               cpu_substeps_per_s=round(len(seq) / cpu_s, 3),
               gpu_substeps_per_s=round((len(seq) - 1) / gpu_s, 2),
               q_rel_err=float(f"{rel(qg, st.q):.3e}"), pr_rel_err=float(f"{rel(pg, st.pr):.3e}"),
               mean_rel_err=float(f"{mean_err(a, b):.3e}"),
               gpu_max_div=s.divgck(), cpu_max_div=st.divgck())
    s.close()
    return out


def mixed_errors(a, b, ma, mb):
    """Max-norm relative difference of the mixed form's state (b) from the FP64
    form's (a) after the same substeps, per field; and the largest difference
    of the xz-mean profiles (device PP_MEAN_ZX moments, velocities in units of
    the bulk velocity), absolute and relative to the profile's maximum."""
    (qa, pa, _), (qb, pb, _) = a, b
    rel = lambda x, y: float(np.abs(x - y).max() / max(np.abs(x).max(), 1e-300))
    out = {n: rel(qa[:, 1:-1, :, c], qb[:, 1:-1, :, c]) for c, n in enumerate("uvw")}
    out["p"] = rel(pa[:, 1:-1, :], pb[:, 1:-1, :])
    out = {k: float(f"{v:.3e}") for k, v in out.items()}
    d = np.abs(ma - mb)
    out["mean_profile_abs"] = {k: float(f"{d[:, c].max():.3e}") for k, c in MEAN_COLS.items()}
    out["mean_profile_rel"] = {k: float(f"{d[:, c].max() / max(np.abs(ma[:, c]).max(), 1e-300):.3e}")
                               for k, c in MEAN_COLS.items() if k in ("u", "p", "uu", "vv", "ww")}
    out["u_mean_profile"] = out["mean_profile_abs"]["u"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--compare-steps", type=int, default=500,
                    help="RK3 steps of both forms before the mixed-vs-FP64 comparison")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                                   "(sharding across GPUs is not built)"}))
        sys.exit(2)
    res = gpu_bench(args.steps, args.warmup, args.compare_steps)
    ncell = res["ncell"]
    f64 = res["fp64"]
    mixed = res["mixed"]
    # roofline of the dominant kernel of the FP64 substep
    ab = alg_bytes_per_cell(CFG["L"])
    kern = {}
    for name, (ms_tot, calls) in f64["prof"].items():
        if name in ab and calls:
            avg = ms_tot / calls
            kern[name] = dict(avg_ms=avg, calls=calls, total_ms=ms_tot,
                              gbs=ab[name] * ncell / (avg * 1e-3) / 1e9)
    dom = max(kern, key=lambda k: kern[k]["total_ms"])
    abm = alg_bytes_per_cell(CFG["L"], "mixed")
    kern_mx = {}
    for name, (ms_tot, calls) in mixed["prof"].items():
        if calls:
            avg = ms_tot / calls
            kern_mx[name] = dict(avg_ms=avg, calls=calls, total_ms=ms_tot)
            if name in abm:
                kern_mx[name]["gbs"] = abm[name] * ncell / (avg * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get(dom)
    roof = dict(bound="hbm", kernel=dom, achieved=round(kern[dom]["gbs"], 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(kern[dom]["gbs"] / HBM_PEAK_GBS, 4), traffic=traffic,
                alg_bytes_per_launch=ab[dom] * ncell, avg_launch_ms=round(kern[dom]["avg_ms"], 5))
    cpu = None
    if not args.no_cpu:
        cb = cpu_baseline()
        _note("cpu baseline done")
        cpu = dict(value=round(cb["substeps_per_s"], 4), unit="substeps/s", cores=cb["cores"], kind="port",
                   sample=f"{cb['nsub']} FP64 RK3 substeps of the C oracle at 256x192x256 "
                          f"(+{cb['npois']} Poisson solves: {cb['poisson_per_s']:.3f}/s), OpenMP, "
                          f"gcc -O2 -fopenmp; one more substep on 1 thread",
                   value_1thread=round(cb["substeps_per_s_1thread"], 4), host=cb["host"],
                   poisson_per_s=round(cb["poisson_per_s"], 4), case_64cube=small_case())
    sub_s = 1e3 / f64["ms_per_substep"]
    out = {
        "metric": "RK3 substeps/s & Poisson solves/s at 256x192x256, div≤1e-10, mixed vs FP64",
        "value": round(sub_s, 3),
        "unit": "substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(f64["ms_per_substep"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded laminar Poiseuille + U(-0.1,0.1) noise)",
        "config": {"workload": "tg channel RK3 substep, 256x192x256, Re_b=6880, FP64, device-resident state",
                   "grid": "256x192x256", "parallelism": "single GPU"},
        "grid_point_substeps_per_s": round(sub_s * ncell, 1),
        "poisson_solves_per_s": round(1e3 / f64["poisson_ms"], 2),
        "max_div_after_timed_steps": f64["maxdiv"],
        "dt": f64["dt"],
        "cn_solver": f64["cn"],
        "mixed": {"err_vs_fp64": mixed_errors(f64["state"], mixed["state"], f64["means"], mixed["means"]),
                  "compared_after": {"substeps": mixed["end"]["substeps"], "fp64_max_div": f64["end"]["maxdiv"],
                                     "mixed_max_div": mixed["end"]["maxdiv"], "cfl": round(f64["end"]["cfl"], 4)},
                  "substeps_per_s": round(1e3 / mixed["ms_per_substep"], 3),
                  "ms_per_substep": round(mixed["ms_per_substep"], 4),
                  "speedup_vs_fp64": round(f64["ms_per_substep"] / mixed["ms_per_substep"], 3),
                  "poisson_solves_per_s": round(1e3 / mixed["poisson_ms"], 2),
                  "poisson_refine_iters_rel": mixed["poisson_refine"],
                  "max_div_after_timed_steps": mixed["maxdiv"]},
        "properties_256x192x256": res["props"],
        "tdma_65536x192": {k: {kk: round(vv, 4) for kk, vv in v.items()} for k, v in res["tdma"].items()},
        "kernels_fp64": {k: {kk: round(vv, 5) for kk, vv in v.items()} for k, v in kern.items()},
        "kernels_mixed": {k: {kk: round(vv, 5) for kk, vv in v.items()} for k, v in kern_mx.items()},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
