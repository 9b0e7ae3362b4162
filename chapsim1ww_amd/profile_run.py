"""Short, fixed workload for hardware-counter profiles (rocprofv3 --pmc / --kernel-trace).

    rocprofv3 --pmc FETCH_SIZE -- python -m chapsim1ww_amd.profile_run

Runs, at 256x192x256 (or --grid L,M,N, --ren): NSUB FP64 RK3 substeps, NPOIS FP64 Poisson solves,
NTDMA batched tridiagonal solves (65536 lines x 192, IJJ form, non-cyclic
and cyclic), NTDMA table-lookup THERM_PROP_UPDATEs over the grid (and, with
--surrogate K, K evaluations of the learned surrogate over it), NHEAT
heated-channel substeps (energy, property staging, io momentum), and a plain
device copy of 1 GiB (read 1 GiB + write 1 GiB)
whose counter readings calibrate the byte counters of the others.
"""
from __future__ import annotations

import argparse

import numpy as np


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--nsub", type=int, default=4)
    ap.add_argument("--npois", type=int, default=4)
    ap.add_argument("--ntdma", type=int, default=4)
    ap.add_argument("--nheat", type=int, default=3)
    ap.add_argument("--precision", default="fp64")
    ap.add_argument("--surrogate", type=int, default=0,
                    help="network evaluations over the grid after a short training (k_mlp_eval counters)")
    ap.add_argument("--grid", default="256,192,256", help="NCL1,NCL2,NCL3")
    ap.add_argument("--ren", type=float, default=6880.0)
    ap.add_argument("--cn", default="auto", help="chs_set_cn_solver mode: auto, thomas, rowtile or marching")
    a = ap.parse_args()
    import torch

    from chapsim1ww_amd import fields
    from chapsim1ww_amd.grid import ChannelGrid
    from chapsim1ww_amd.solver import ChannelSolver, tdma_batched

    dev = torch.device("cuda:0")
    st = torch.cuda.current_stream(dev)
    L, M, N = (int(v) for v in a.grid.split(","))
    g = ChannelGrid(L, M, N, ren=a.ren)
    q, pr, c0 = fields.init_tg_fields(g, seed=1973)
    s = ChannelSolver(g, a.precision, stream=st)
    s.set_cn_solver(a.cn)
    s.upload_state(q, pr, c0)
    dt = 0.5 / s.cfl()
    s.solverrk3_mom_tg(0, dt)
    for i in range(a.nsub):
        s.solverrk3_mom_tg(1 + i % 3, dt)
    f = torch.from_numpy(np.asfortranarray(fields.poisson_rhs(g, seed=9)).ravel(order="F").copy()).to(dev)
    dph = torch.zeros(g.L * (g.M + 2) * g.N, dtype=torch.float64, device=dev)
    for _ in range(a.npois):
        s.fft99_pois3d(f, dph)
    n, nl = 192, 65536
    rng = np.random.default_rng(2)
    A = torch.from_numpy(rng.uniform(-0.9, -0.1, n * nl)).to(dev)
    C = torch.from_numpy(rng.uniform(-0.9, -0.1, n * nl)).to(dev)
    B = torch.from_numpy(2.0 + np.abs(rng.standard_normal(n * nl))).to(dev)
    R = torch.from_numpy(rng.standard_normal(n * nl)).to(dev)
    for i in range(a.ntdma):
        tdma_batched(A, B, C, R, None, n, nl, nl, 1, i % 2 == 1, st)
    # THERM_PROP_UPDATE by table lookup over the whole grid
    from chapsim1ww_amd import props
    sc = props.scale_table(props.synthetic_table())
    pt = props.PropertyTable(sc, stream=st)
    rng2 = np.random.default_rng(4)
    hh = torch.from_numpy(rng2.uniform(sc["H"][1], sc["H"][-2], g.ncell)).to(dev)
    outs = [torch.empty_like(hh) for _ in range(5)]
    for _ in range(a.ntdma):
        pt.update(hh, *outs)
    if a.surrogate > 0:
        # the learned surrogate on all cells: 2000 Adam steps on 2^18 of them,
        # then a.surrogate evaluations of the network over the grid
        pt.train(hh[: 1 << 18].contiguous(), 2000, 4096, 3e-3, seed=1)
        net = torch.empty(5 * g.ncell, dtype=torch.float64, device=dev)
        for _ in range(a.surrogate):
            pt.eval(hh, net)
        del net
    # heated channel: energy + staging + io momentum from the tg velocity as mass flux
    from chapsim1ww_amd.heated import HeatedChannel
    hc = HeatedChannel(s, pt)
    L, M, N = g.L, g.M, g.N
    yc = np.concatenate([[-1.0], np.asarray(g.ycc), [1.0]])
    T3 = 1.2 - 0.3 * yc[None, :, None] ** 2 * np.ones((L, 1, N))
    hg = torch.from_numpy(np.interp(T3.ravel(order="F"), sc["T"], sc["H"])).to(dev)
    fd = hc.dev
    fd["enthalpy"].copy_(hg)
    pt.update(hg, fd["temperature"], fd["density"], fd["viscosity"], fd["thermcondt"])
    fd["rhoh"].copy_(hg * fd["density"])
    for k in ("density0", "density1"):
        fd[k].copy_(fd["density"])
    fd["viscosity0"].copy_(fd["viscosity"])
    qs, _, _ = s.download_state()
    qd = torch.from_numpy(np.asfortranarray(qs).ravel(order="F")).to(dev)
    fd["g"].copy_(qd)
    fd["g0"].copy_(qd)
    fd["wallflux"].fill_(0.1)
    for i in range(a.nheat):
        hc.substep(1 + i % 3, 1e-3, 1.0 / 2800.0)
    pt.close()
    # calibration: 1 GiB device-to-device copy
    x = torch.empty(1 << 27, dtype=torch.float64, device=dev)
    y = torch.empty_like(x)
    x.fill_(1.0)
    for _ in range(2):
        y.copy_(x)
    torch.cuda.synchronize()
    s.close()
    print("profile_run done", flush=True)


if __name__ == "__main__":
    main()
