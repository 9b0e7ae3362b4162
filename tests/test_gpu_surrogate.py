"""The learned THERM_PROP_UPDATE surrogate on the device: forward and
gradient against the CPU oracle, training, certification, the certified
update with table fallback, and what the network gains over the table."""
import time

import numpy as np
import pytest

import mlp_oracle as MO
import oracle as O
from chapsim1ww_amd import fields, props
from chapsim1ww_amd.grid import ChannelGrid

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scaled():
    return props.scale_table(props.synthetic_table(), t0=1.0)


@pytest.fixture(scope="module")
def snapshot_h(scaled):
    """Enthalpy samples from a channel-solver snapshot (64^3, a few substeps)."""
    from chapsim1ww_amd.solver import ChannelSolver
    g = ChannelGrid(64, 64, 64)
    q, pr, c0 = fields.init_tg_fields(g, seed=5)
    s = ChannelSolver(g)
    s.upload_state(q, pr, c0)
    dt = 0.5 / s.cfl()
    for ns in (0, 1, 2, 3, 1, 2, 3):
        s.solverrk3_mom_tg(ns, dt)
    qg, _, _ = s.download_state()
    s.close()
    return props.snapshot_enthalpy(qg[:, 1:-1, :, 0], scaled, 1 << 18, seed=3)


def _labels(scaled, h):
    tab, n = O.build_prop_tab(scaled)
    out = [np.zeros_like(h) for _ in range(5)]
    assert O.therm_prop_update(tab, n, h, *out) == 0
    return np.stack(out, 1)


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _oracle_norm(norm):
    return (norm[0], norm[1], norm[4:9], norm[9:14])


def test_forward_matches_oracle(scaled, snapshot_h):
    """Initial network (Glorot, seeded): device forward on the matrix cores vs
    the oracle's bf16-emulated forward (checks the lane layouts of the MFMA
    chain as a whole)."""
    pt = props.PropertyTable(scaled)
    pt.train(_dev(snapshot_h), steps=0, seed=7)
    par, norm = pt.params()
    h = np.concatenate([snapshot_h[:3000], np.linspace(norm[2], norm[3], 1001)])
    out = torch.zeros(5 * h.size, dtype=torch.float64, device="cuda")
    pt.eval(_dev(h), out)
    got = out.cpu().numpy().reshape(5, h.size).T
    ref, _ = MO.forward(par, _oracle_norm(norm), h)
    err = np.abs(got - ref) / np.abs(ref).clip(1e-3)
    print("forward max rel err", err.max(), flush=True)
    assert err.max() < 2e-5
    pt.close()


def test_gradient_matches_oracle(scaled, snapshot_h):
    pt = props.PropertyTable(scaled)
    pt.train(_dev(snapshot_h), steps=0, seed=7)
    par, norm = pt.params()
    h = snapshot_h[:2000]
    loss, g = pt.grad(_dev(h))
    t = _labels(scaled, h).astype(np.float32)
    L, gr = MO.loss_grad(par, _oracle_norm(norm), h, t)
    print("loss", loss, L, flush=True)
    assert abs(loss - L) <= 1e-4 * L
    for name, lo, hi in (("w1", MO.O_W1, MO.O_B1), ("b1", MO.O_B1, MO.O_W2), ("W2", MO.O_W2, MO.O_B2),
                         ("b2", MO.O_B2, MO.O_W3), ("W3", MO.O_W3, MO.O_B3), ("b3", MO.O_B3, MO.NPAR)):
        e = np.linalg.norm(g[lo:hi] - gr[lo:hi]) / np.linalg.norm(gr[lo:hi])
        print(name, e, flush=True)
        assert e < 1e-4, name
    pt.close()


def test_train_certify_and_update(scaled, snapshot_h):
    """Train on the snapshot samples, certify at 1e-4, then THERM_PROP_UPDATE
    in surrogate mode against table mode on a 2^22-cell field: certified cells
    within 1e-4 of the table, all others equal to the table bit for bit; the
    timing of both modes is what the network gains (or costs)."""
    pt = props.PropertyTable(scaled)
    dh = _dev(snapshot_h)
    l0 = pt.train(dh, steps=0, seed=11)
    l1 = pt.train(dh, steps=20000, batch=8192, lr=2e-3, seed=11)
    print(f"loss init {l0:.3e} trained {l1:.3e}", flush=True)
    assert l1 < 1e-4 * l0
    loose = pt.certify(1e-3, 2048)
    frac = pt.certify(1e-4, 2048)
    print(f"certified fraction of the trained range: {frac:.4f} at 1e-4 ({loose:.4f} at 1e-3)", flush=True)
    assert loose > 0.0
    par, norm = pt.params()
    # a field: the snapshot distribution plus cells outside the trained range
    rng = np.random.default_rng(9)
    n = 1 << 22
    h = snapshot_h[rng.integers(0, snapshot_h.size, n)]
    lo, hi = pt.h_range
    h[:1000] = rng.uniform(lo + 1e-6, norm[2], 1000)    # below the trained range: table
    h[1000:2000] = 0.0                                  # |h| < 1e-10: untouched
    dh2 = _dev(h)
    tab = [torch.full((n,), -1.0, dtype=torch.float64, device="cuda") for _ in range(5)]
    sur = [torch.full((n,), -1.0, dtype=torch.float64, device="cuda") for _ in range(5)]
    pt.update(dh2, *tab)
    nf = pt.update(dh2, *sur, mode="surrogate")
    torch.cuda.synchronize()
    A = np.stack([t.cpu().numpy() for t in tab], 1)
    B = np.stack([t.cpu().numpy() for t in sur], 1)
    same = np.all(A == B, 1)
    rel = (np.abs(B - A) / np.abs(A).clip(1e-300)).max(1)
    used_net = ~same
    print(f"fallback cells {nf} of {n}; network cells {used_net.sum()}; max rel err {rel.max():.3e}", flush=True)
    assert rel.max() <= 1e-4
    assert np.all(A[1000:2000] == -1.0) and np.all(B[1000:2000] == -1.0)
    assert np.all(same[:1000])
    assert nf >= 1000
    # timing: table vs surrogate on the same field
    def timed(mode):
        for _ in range(2):
            pt.update(dh2, *sur, mode=mode)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(10):
            pt.update(dh2, *sur, mode=mode)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / 10
    tt, ts = timed("table"), timed("surrogate")
    print(f"THERM_PROP_UPDATE 2^22 cells: table {tt*1e3:.3f} ms, surrogate {ts*1e3:.3f} ms, "
          f"gain {tt/ts:.3f}x, certified {frac:.3f}", flush=True)
    pt.close()


HEAT_MAP = {"rhoh": "rhoh", "h": "enthalpy", "T": "temperature", "rho": "density", "rho0": "density0",
            "rho1": "density1", "mu": "viscosity", "mu0": "viscosity0", "kap": "thermcondt", "rhs0": "rhs_energy0",
            "g": "g", "g0": "g0", "wallflux": "wallflux", "q": "q", "pr": "pr", "drhodtp": "drhodtp",
            "divu": "divu", "rho_stg": "rho_stg", "mu_stg": "mu_stg", "explt0": "explt0", "drhoi_stg": "drhoi_stg"}


def test_surrogate_on_heated_snapshots_256x192x256():
    """The surrogate workload at its own size: the heated channel at
    256x192x256 (2048-point table, q_w = 0.1, Re = 6880, Pr = 1) on the
    mixed-precision solver; 2^20 enthalpy samples from the solver's own
    ENTHALPY after three substeps; training, certification at 1e-4;
    THERM_PROP_UPDATE on the whole snapshot in both modes (network cells within
    1e-4 of the table, fallback cells equal to it bit for bit, the time of both
    modes recorded); then three more substeps with the energy substep's
    property update in surrogate mode beside a table-mode twin (divergence
    below 1e-10, T, rho, mu, kappa within 1e-4 of the twin)."""
    from chapsim1ww_amd.heated import HeatedChannel
    from chapsim1ww_amd.solver import ChannelSolver
    dims, ren = (256, 192, 256), 6880.0
    L, M, N = dims
    print(f"surrogate on heated snapshots {dims}", flush=True)
    scaled = props.scale_table(props.synthetic_table(2048), t0=1.0)
    g = ChannelGrid(*dims, ren=ren)
    og = O.OracleGrid(*dims, ren=ren)
    st, _, _ = O.io_case(og, scaled, seed=2026, qw=0.1)
    cthecd, dt = 1.0 / ren, 5e-4
    twins = []
    for _ in range(2):
        s = ChannelSolver(g, "mixed")
        pt = props.PropertyTable(scaled)
        hc = HeatedChannel(s, pt, thermal=1)
        hc.upload({HEAT_MAP[k]: v for k, v in st.items()})
        twins.append((s, pt, hc))
    del st
    s, pt, hc = twins[0]
    for ns in (1, 2, 3):
        for (_, _, h_) in twins:
            h_.substep(ns, dt, cthecd)
        assert hc.divgck_io()[0] <= 1e-10
    # 2^20 samples of the solver's ENTHALPY (interior rows), drawn on the device
    H = hc.dev["enthalpy"].view(N, M + 2, L)[:, 1:M + 1, :].reshape(-1)
    gen = torch.Generator(device="cuda").manual_seed(7)
    samples = H[torch.randint(0, H.numel(), (1 << 20,), device="cuda", generator=gen)].contiguous()
    t0 = time.perf_counter()
    loss = pt.train(samples, steps=40000, batch=4096, lr=3e-3, seed=5)
    t_train = time.perf_counter() - t0
    frac5 = pt.certify(1e-4, 4096)
    frac5, frac4 = pt.certified()
    print(f"trained 40000 Adam steps on 2^20 snapshot samples in {t_train:.2f} s, loss {loss:.3e}; certified "
          f"at 1e-4: {frac5:.4f} (all outputs), {frac4:.4f} (T, rho, mu, kappa)", flush=True)
    # THERM_PROP_UPDATE on the snapshot's interior in both modes (no c_p)
    Hc = H.contiguous()
    n = Hc.numel()
    tab = [torch.full((n,), -1.0, dtype=torch.float64, device="cuda") for _ in range(4)]
    sur = [torch.full((n,), -1.0, dtype=torch.float64, device="cuda") for _ in range(4)]
    pt.update(Hc, *tab)
    nf = pt.update(Hc, *sur, mode="surrogate")
    torch.cuda.synchronize()
    A = torch.stack(tab, 1)
    B = torch.stack(sur, 1)
    rel = ((B - A).abs() / A.abs().clamp_min(1e-300)).max(1).values
    net = (B != A).any(1)
    worst = float(rel.max())
    print(f"snapshot update: {n} cells, network {int(net.sum())}, fallback {nf}, max rel err {worst:.3e}",
          flush=True)
    assert worst <= 1e-4
    assert int(net.sum()) <= n - nf
    if frac4 == 0.0:
        assert nf == n

    def timed(mode):
        for _ in range(2):
            pt.update(Hc, *sur, mode=mode)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(10):
            pt.update(Hc, *sur, mode=mode)
        torch.cuda.synchronize()
        return (time.perf_counter() - t1) / 10
    tt, ts = timed("table"), timed("surrogate")
    print(f"THERM_PROP_UPDATE {n} cells: table {tt * 1e3:.3f} ms, surrogate {ts * 1e3:.3f} ms, "
          f"gain {tt / ts:.3f}x", flush=True)
    # the energy substep with the surrogate in its property update, beside the table twin
    if frac4 > 0.0:
        pt.set_mode("surrogate")
    for ns in (1, 2, 3):
        for (_, _, h_) in twins:
            h_.substep(ns, dt, cthecd)
        assert hc.divgck_io()[0] <= 1e-10 and twins[1][2].divgck_io()[0] <= 1e-10
    fb = pt.last_fallback()
    a, b = hc.download(), twins[1][2].download()
    errs = {k: float(np.abs(a[k] - b[k]).max() / np.abs(b[k]).max())
            for k in ("temperature", "density", "viscosity", "thermcondt", "enthalpy")}
    print(f"3 substeps, energy update in {'surrogate' if frac4 > 0 else 'table'} mode vs table twin: "
          f"{ {k: f'{v:.2e}' for k, v in errs.items()} }, last fallback {fb}", flush=True)
    for k in ("temperature", "density", "viscosity", "thermcondt"):
        assert errs[k] <= 1e-4, k
    for (s_, p_, _) in twins:
        p_.close()
        s_.close()
