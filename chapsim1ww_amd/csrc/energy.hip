// Energy substep of the heated channel on one GPU: SOLVERRK3_ENG_IO(NS)
// (src/SOLVERRK3_ENG_IO.f90:1-77) with periodic x and z and iso-flux walls.
//
//   k_energy_rhs    RHS_ENERGY_EXPLICIT (src/RHS_ENERGY_EXPLICIT.f90:1-150) and
//                   the RK3 update RHOH += TGAM DT RHS + TROH DT RHS0, clamped
//                   to RHOHmax + REALMIN; RHS0 = RHS
//   k_energy_props  VISCOUSITY0 = VISCOUSITY, DENSITY1 = DENSITY0,
//                   DENSITY0 = DENSITY; then per interior cell ENTHALPY_UPDATE,
//                   DENSITY_UPDATE, ENTHALPY_UPDATE, THERM_PROP_UPDATE
//                   (src/THERM_PROP_UPDATE.f90:1-33, 70-205)
//   k_energy_wall   BC_WALL_THERMAL -> BC_WALL_ISOFLUX at both walls
//                   (src/BC_WALL.f90:522-650)
// INTFC_MFD_THERMAL_io exchanges h, rho, T, mu, kappa and rho*h across ranks;
// on one rank the wall conditions then overwrite all six on the ghost rows,
// so it has nothing to do here.
//
// Arrays are the reference's (NCL1, 0:NCL2+1, NCL3), x fastest, except
// RHS_ENERGY0 (NCL1, NCL2, NCL3), G/G0 (NCL1, 0:NCL2+1, NCL3, 3) and
// WALLFLUX (NCL1, 2).  The stencil and the property arithmetic are evaluated
// as written, without contraction into fused multiply-adds, so the results
// equal an FP64 evaluation of the reference's expressions bit for bit.
#include <algorithm>
#include <cmath>

#include "props.hpp"

namespace chs {

struct EGeo {
    int L, M, N;
    double dxi, dzi, dxqi, dzqi;
    const double *dyfi, *dyci, *wff, *wfb;
};

struct EFields {
    double *rhoh, *h, *T, *rho, *rho0, *rho1, *mu, *mu0, *kap, *rhs0;
    const double *g, *g0, *wallflux;
};

// one thread per interior cell; block (64 i) x (4 k), grid (i tiles, k tiles, j)
__global__ void __launch_bounds__(256) k_energy_rhs(EGeo g, EFields f, double rk1, double rk0, double cthecd,
                                                    double rhoh_cap) {
#pragma clang fp contract(off)
    const int L = g.L, M = g.M, N = g.N;
    int it, jt, kt, tile;
    if (!row_tile(L, N, M, &it, &jt, &kt, &tile)) return;
    const int IC = it * 64 + threadIdx.x;            // 0-based i
    const int KC = kt * 4 + threadIdx.y;
    const int JC = jt + 1;                           // Fortran J (ghost-aware row)
    if (IC >= L || KC >= N) return;
    const int IP = IC == L - 1 ? 0 : IC + 1, IM = IC == 0 ? L - 1 : IC - 1;
    const int KP = KC == N - 1 ? 0 : KC + 1, KM = KC == 0 ? N - 1 : KC - 1;
    const int JJ = JC, JJP = JJ + 1;
    const int LP = L * (M + 2), CS = LP * N;
    auto at = [&](int i, int j, int k) { return i + L * j + LP * k; };
    const double COE1 = g.dxi * 0.5 * 0.5;
    const double COE11 = g.dxqi * 0.5 * cthecd;
    const double COE2 = ldc(g.dyfi, JJ - 1) * 1.0 * 0.5;
    const double COE3 = g.dzi * 0.5 * 1.0 * 0.5;
    const double COE33 = g.dzqi * 0.5 * 1.0 * cthecd;
    const double COE22 = ldc(g.dyfi, JJ - 1) * 1.0 * cthecd;
    double COE221, COE222;
    if (JJ == 1) {
        COE221 = ldc(g.dyci, JJP - 1) / 1.0;
        COE222 = 1.0 / 1.0;
    } else if (JJ == M) {
        COE222 = ldc(g.dyci, JJ - 1) / 1.0;
        COE221 = 1.0 / 1.0;
    } else {
        COE221 = ldc(g.dyci, JJP - 1) / 1.0;
        COE222 = ldc(g.dyci, JJ - 1) / 1.0;
    }
    const double wffp = ldc(g.wff, JJP - 1), wfbp = ldc(g.wfb, JJP - 1), wff = ldc(g.wff, JJ - 1), wfb = ldc(g.wfb, JJ - 1);
    const double* H = f.h;
    const double* TT = f.T;
    const double* KK = f.kap;
    const double* G = f.g;
    const double* G0 = f.g0;
    const int c = at(IC, JC, KC);
    double DX_RHOHU = (H[at(IP, JC, KC)] + H[c]) * (G0[at(IP, JC, KC)] + G[at(IP, JC, KC)]) -
                      (H[c] + H[at(IM, JC, KC)]) * (G0[c] + G[c]);
    DX_RHOHU = DX_RHOHU * COE1;
    double DY_RHOHV = (wffp * H[at(IC, JC + 1, KC)] + wfbp * H[c]) *
                          (G0[CS + at(IC, JC + 1, KC)] + G[CS + at(IC, JC + 1, KC)]) -
                      (wff * H[c] + wfb * H[at(IC, JC - 1, KC)]) * (G0[CS + c] + G[CS + c]);
    DY_RHOHV = DY_RHOHV * COE2;
    double DZ_RHOHW = (H[at(IC, JC, KP)] + H[c]) * (G0[2 * CS + at(IC, JC, KP)] + G[2 * CS + at(IC, JC, KP)]) -
                      (H[c] + H[at(IC, JC, KM)]) * (G0[2 * CS + c] + G[2 * CS + c]);
    DZ_RHOHW = DZ_RHOHW * COE3;
    double DX_KDT = (KK[at(IP, JC, KC)] + KK[c]) * (TT[at(IP, JC, KC)] - TT[c]) -
                    (KK[c] + KK[at(IM, JC, KC)]) * (TT[c] - TT[at(IM, JC, KC)]);
    DX_KDT = DX_KDT * COE11;
    double KDTy1 = (wffp * KK[at(IC, JC + 1, KC)] + wfbp * KK[c]) * (TT[at(IC, JC + 1, KC)] - TT[c]) * COE221;
    double KDTy2 = (wff * KK[c] + wfb * KK[at(IC, JC - 1, KC)]) * (TT[c] - TT[at(IC, JC - 1, KC)]) * COE222;
    if (JJ == 1)
        KDTy2 = -f.wallflux[IC] * COE222;
    else if (JJ == M)
        KDTy1 = -f.wallflux[IC + L] * COE221;
    const double DY_KDT = (KDTy1 - KDTy2) * COE22;
    double DZ_KDT = (KK[at(IC, JC, KP)] + KK[c]) * (TT[at(IC, JC, KP)] - TT[c]) -
                    (KK[at(IC, JC, KM)] + KK[c]) * (TT[c] - TT[at(IC, JC, KM)]);
    DZ_KDT = DZ_KDT * COE33;
    const double RHS = -DX_RHOHU - DY_RHOHV - DZ_RHOHW + DX_KDT + DY_KDT + DZ_KDT;
    const int r0 = IC + L * (JC - 1) + L * M * KC;
    double r = f.rhoh[c] + rk1 * RHS + rk0 * f.rhs0[r0];
    r = fmin(rhoh_cap, r);
    f.rhoh[c] = r;
    f.rhs0[r0] = RHS;
}

constexpr int EPT = 512;   // threads per property block

// rotations of the property history on every cell, the property cascade on
// interior rows; persistent blocks over k planes
__global__ void __launch_bounds__(EPT) k_energy_props(PropTab t, EGeo g, EFields f,
                                                      unsigned long long* __restrict__ cnt) {
#pragma clang fp contract(off)
    extern __shared__ unsigned char smem[];
    const double *xh, *xt;
    const int *sh, *st;
    stage_tables(t, smem, xh, xt, sh, st);
    // 32-bit cell offsets (chs_create keeps 3*L*(M+2)*N below 2^31): the
    // row of a cell costs one 32-bit division instead of two 64-bit ones
    const int L = g.L, M = g.M;
    const int ntot = L * (M + 2) * g.N;
    unsigned long long bad = 0;
    for (int q = blockIdx.x * EPT + threadIdx.x; q < ntot; q += gridDim.x * EPT) {
        const int j = (q / L) % (M + 2);
        const double mu_old = f.mu[q], rho_old = f.rho[q], rho0_old = f.rho0[q];
        f.mu0[q] = mu_old;
        f.rho1[q] = rho0_old;
        f.rho0[q] = rho_old;
        if (j == 0 || j == M + 1) continue;
        const double rh = f.rhoh[q];
        double hv = rh / rho_old;
        double Tt, D, Mu, K, CP;
        if (!spline_ht(t, xh, sh, hv, Tt, true) || !spline_t(t, xt, st, Tt, D, Mu, K, CP)) {
            ++bad;
            continue;
        }
        f.rho[q] = D;
        hv = rh / D;
        f.h[q] = hv;
        if (fabs(hv) < 1.0e-10) continue;   // THERM_PROP_UPDATE leaves such cells
        if (!spline_ht(t, xh, sh, hv, Tt, false) || !spline_t(t, xt, st, Tt, D, Mu, K, CP)) {
            ++bad;
            continue;
        }
        f.T[q] = Tt;
        f.mu[q] = Mu;
        f.kap[q] = K;
        f.rho[q] = D;
    }
    for (int o = 32; o > 0; o >>= 1) bad += __shfl_xor(bad, o);
    if ((threadIdx.x & 63) == 0 && bad) atomicAdd(cnt, bad);
}

// BC_WALL_ISOFLUX: ghost rows J = 0 and M+1 from the two nearest interior rows
__global__ void __launch_bounds__(256) k_energy_wall(PropTab t, EGeo g, EFields f,
                                                     unsigned long long* __restrict__ cnt) {
#pragma clang fp contract(off)
    extern __shared__ unsigned char smem[];
    const double *xh, *xt;
    const int *sh, *st;
    stage_tables(t, smem, xh, xt, sh, st);
    const int L = g.L, M = g.M;
    const long long LP = (long long)L * (M + 2);
    const long long nw = 2LL * L * g.N;
    unsigned long long bad = 0;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < nw; e += (long long)gridDim.x * blockDim.x) {
        const int top = e < (long long)L * g.N ? 1 : 0;
        const long long ik = top ? e : e - (long long)L * g.N;
        const int i = (int)(ik % L);
        const long long k = ik / L;
        const long long base = i + LP * k;
        double hv;
        long long q;
        if (top) {
            const int J = M + 1;
            q = base + (long long)L * J;
            hv = 2.0 * f.h[q - L] - ldc(g.wff, J - 2) * f.h[q - L] - ldc(g.wfb, J - 2) * f.h[q - 2 * L];
        } else {
            q = base;
            hv = 2.0 * f.h[q + L] - ldc(g.wff, 1) * f.h[q + 2 * L] - ldc(g.wfb, 1) * f.h[q + L];
        }
        f.h[q] = hv;
        double Tt, D, Mu, K, CP;
        if (!spline_ht(t, xh, sh, hv, Tt, true) || !spline_t(t, xt, st, Tt, D, Mu, K, CP)) {
            ++bad;
            continue;
        }
        f.rho[q] = D;
        f.T[q] = Tt;
        f.mu[q] = Mu;
        f.kap[q] = K;
        f.rhoh[q] = hv * D;
    }
    for (int o = 32; o > 0; o >>= 1) bad += __shfl_xor(bad, o);
    if ((threadIdx.x & 63) == 0 && bad) atomicAdd(cnt, bad);
}

__global__ void k_zero_u64_e(unsigned long long* c, int n) {
    if ((int)threadIdx.x < n) c[threadIdx.x] = 0ull;
}

static const double TGAM_E[4] = {1.0, 8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
static const double TROH_E[4] = {0.0, 0.0, -17.0 / 60.0, -5.0 / 12.0};

int energy_substep(chs_handle* h, chs_props* p, const chs_thermal_fields* tf, int ns, double dt, double cthecd) {
    CHS_CHECK(h && p && tf, CHS_E_ARG, "chs_energy_substep: NULL argument");
    CHS_CHECK(ns >= 0 && ns <= 3, CHS_E_ARG, "chs_energy_substep: NS must be 0..3");
    CHS_CHECK(h->M >= 2, CHS_E_ARG, "chs_energy_substep: NCL2 >= 2 needed");
    const double* req[] = {tf->rhoh, tf->enthalpy, tf->temperature, tf->density, tf->density0, tf->density1,
                           tf->viscosity, tf->viscosity0, tf->thermcondt, tf->rhs_energy0, tf->g, tf->g0,
                           tf->wallflux};
    for (const double* r : req) CHS_CHECK(r != nullptr, CHS_E_ARG, "chs_energy_substep: NULL field");
    EGeo g{h->L, h->M, h->N, h->dxi, h->dzi, h->dxqi, h->dzqi, h->dyfi, h->dyci, h->wff, h->wfb};
    EFields f{tf->rhoh, tf->enthalpy, tf->temperature, tf->density, tf->density0, tf->density1, tf->viscosity,
              tf->viscosity0, tf->thermcondt, tf->rhs_energy0, tf->g, tf->g0, tf->wallflux};
    const hipStream_t s = h->stream;
    const double cap = p->rhoh_max + 1.0e-14;   // RHOHmax + REALMIN
    hipEvent_t ev;
    CHS_LAUNCH(k_zero_u64_e, dim3(1), dim3(64), 0, s, p->counters, 4);
    h->timer.begin("energy_rhs", s, &ev);
    CHS_LAUNCH(k_energy_rhs, dim3(row_tile_blocks(h->L, h->N, h->M)), dim3(64, 4), 0, s, g, f, TGAM_E[ns] * dt,
               TROH_E[ns] * dt, cthecd, cap);
    h->timer.end(ev, "energy_rhs", s);
    const PropTab t = tab_of(p);
    h->timer.begin("energy_props", s, &ev);
    CHS_LAUNCH(k_energy_props, dim3(512), dim3(EPT), props_lds_bytes(p), s, t, g, f, p->counters);
    h->timer.end(ev, "energy_props", s);
    CHS_LAUNCH(k_energy_wall, dim3(64), dim3(256), props_lds_bytes(p), s, t, g, f, p->counters);
    unsigned long long bad = 0;
    CHS_HIP(hipMemcpyAsync(&bad, p->counters, sizeof(bad), hipMemcpyDeviceToHost, s));
    CHS_HIP(hipStreamSynchronize(s));
    if (bad) {
        set_error("chs_energy_substep: " + std::to_string(bad) + " property lookups outside the table");
        return CHS_E_RANGE;
    }
    return CHS_OK;
}

}  // namespace chs

extern "C" int chs_energy_substep(chs_handle* h, chs_props* p, const chs_thermal_fields* f, int ns, double dt,
                                  double cthecd) {
    return chs::energy_substep(h, p, f, ns, dt, cthecd);
}
