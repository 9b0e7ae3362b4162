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
