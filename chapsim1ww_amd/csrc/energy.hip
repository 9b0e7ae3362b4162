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

int mlp_update(chs_props* p, const double* h, Rows rw, double* T, double* rho, double* mu, double* kap, double* cp,
               unsigned long long* cnt, hipStream_t stream);

struct EGeo {
    int L, M, N;
    double dxi, dzi, dxqi, dzqi;
    const double *dyfi, *dyci, *wff, *wfb;
};

struct EFields {
    double *rhoh, *h, *T, *rho, *rho0, *rho1, *mu, *mu0, *kap, *rhs0;
    const double *g, *g0, *wallflux;
};

// G0 and G of one cell: u at i+1 and at the cell, v at j+1 and at the cell,
// w at k+1 and at the cell
struct EGv {
    double u0p, up, u0c, uc, v0p, vp, v0c, vc, w0p, wp, w0c, wc;
};

// RHS_ENERGY_EXPLICIT at the interior cell (IC, JC, KC): H, TT, KK access
// the enthalpy, temperature and conductivity around it, q the mass fluxes.
template <class A>
__device__ __forceinline__ double energy_rhs_cell(const EGeo& g, const A& H, const A& TT, const A& KK,
                                                  const EGv& q, int IC, int JC, double cthecd,
                                                  const double* __restrict__ wallflux) {
#pragma clang fp contract(off)
    const int M = g.M, L = g.L;
    const int JJ = JC, JJP = JJ + 1;
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
    const double wffp = ldc(g.wff, JJP - 1), wfbp = ldc(g.wfb, JJP - 1), wff = ldc(g.wff, JJ - 1),
                 wfb = ldc(g.wfb, JJ - 1);
    const double hc = H(0, 0, 0);
    double DX_RHOHU = (H(1, 0, 0) + hc) * (q.u0p + q.up) - (hc + H(-1, 0, 0)) * (q.u0c + q.uc);
    DX_RHOHU = DX_RHOHU * COE1;
    double DY_RHOHV = (wffp * H(0, 1, 0) + wfbp * hc) * (q.v0p + q.vp) - (wff * hc + wfb * H(0, -1, 0)) * (q.v0c + q.vc);
    DY_RHOHV = DY_RHOHV * COE2;
    double DZ_RHOHW = (H(0, 0, 1) + hc) * (q.w0p + q.wp) - (hc + H(0, 0, -1)) * (q.w0c + q.wc);
    DZ_RHOHW = DZ_RHOHW * COE3;
    const double kc = KK(0, 0, 0), tc = TT(0, 0, 0);
    double DX_KDT = (KK(1, 0, 0) + kc) * (TT(1, 0, 0) - tc) - (kc + KK(-1, 0, 0)) * (tc - TT(-1, 0, 0));
    DX_KDT = DX_KDT * COE11;
    double KDTy1 = (wffp * KK(0, 1, 0) + wfbp * kc) * (TT(0, 1, 0) - tc) * COE221;
    double KDTy2 = (wff * kc + wfb * KK(0, -1, 0)) * (tc - TT(0, -1, 0)) * COE222;
    if (JJ == 1)
        KDTy2 = -wallflux[IC] * COE222;
    else if (JJ == M)
        KDTy1 = -wallflux[IC + L] * COE221;
    const double DY_KDT = (KDTy1 - KDTy2) * COE22;
    double DZ_KDT = (KK(0, 0, 1) + kc) * (TT(0, 0, 1) - tc) - (KK(0, 0, -1) + kc) * (tc - TT(0, 0, -1));
    DZ_KDT = DZ_KDT * COE33;
    return -DX_RHOHU - DY_RHOHV - DZ_RHOHW + DX_KDT + DY_KDT + DZ_KDT;
}

// The RK3 update of one cell: RHOH += rk1 RHS + rk0 RHS0, clamped to the cap
__device__ __forceinline__ double energy_update(double rhoh, double rhs, double rhs0, double rk1, double rk0,
                                                double cap) {
#pragma clang fp contract(off)
    const double r = rhoh + rk1 * rhs + rk0 * rhs0;
    return fmin(cap, r);
}

// one thread per interior cell in row tiles (64 i x 4 k)
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
    const int LP = L * (M + 2), CS = LP * N;
    const int ip = IC == L - 1 ? 1 - L : 1, im = IC == 0 ? L - 1 : -1;
    const int kp = KC == N - 1 ? (1 - N) * LP : LP, km = KC == 0 ? (N - 1) * LP : -LP;
    const int c = IC + L * JC + LP * KC;
    const double *G = f.g, *G0 = f.g0;
    EGv q;
    q.u0p = G0[c + ip], q.up = G[c + ip], q.u0c = G0[c], q.uc = G[c];
    q.v0p = G0[CS + c + L], q.vp = G[CS + c + L], q.v0c = G0[CS + c], q.vc = G[CS + c];
    q.w0p = G0[2 * CS + c + kp], q.wp = G[2 * CS + c + kp], q.w0c = G0[2 * CS + c], q.wc = G[2 * CS + c];
    const double RHS = energy_rhs_cell(g, cell_acc(f.h, c, ip, im, kp, km, L), cell_acc(f.T, c, ip, im, kp, km, L),
                                       cell_acc(f.kap, c, ip, im, kp, km, L), q, IC, JC, cthecd, f.wallflux);
    const int r0 = IC + L * (JC - 1) + L * M * KC;
    f.rhoh[c] = energy_update(f.rhoh[c], RHS, f.rhs0[r0], rk1, rk0, rhoh_cap);
    f.rhs0[r0] = RHS;
}

// ---------------------------------------------------------------------------
// k_energy_rhs_m: k_energy_rhs marching along y, as k_rhs_xm: a block owns
// EKR = 4 consecutive k rows of every x line and a chunk of rows, one cell
// of each row per thread (4L threads).  LDS holds a ring of the rows JC-1,
// JC, JC+1 of H, T and kappa (EKR + 2 k rows each); the mass fluxes, RHOH
// and RHS0 of a row are loaded one step ahead in registers, G0 and G of v at
// the cell carried from the step before.
// ---------------------------------------------------------------------------
constexpr int EKR = 4, EKP = EKR + 2;

inline size_t energy_rhs_m_smem(int L) { return sizeof(double) * (size_t)3 * 3 * EKP * L; }

__global__ void __launch_bounds__(1024) k_energy_rhs_m(EGeo g, EFields f, double rk1, double rk0, double cthecd,
                                                       double rhoh_cap, int nyc, int jc, int lgL) {
#pragma clang fp contract(off)
    extern __shared__ double sm[];
    const int L = g.L, M = g.M, N = g.N;
    const int PL = EKP * L, SL = 3 * PL;  // one array / one ring row (H, T, kappa)
    double* ring = sm;
    const int nt = (N / EKR) * nyc, per = (nt + 7) / 8;
    const int t = (blockIdx.x % 8) * per + blockIdx.x / 8;  // XCD x: tiles x*per .. (x+1)*per-1
    if (t >= nt) return;
    const int kg = t / nyc;
    const int j0 = (t - kg * nyc) * jc, j1 = min(M, j0 + jc);
    const int k0 = kg * EKR;
    const int tid = threadIdx.x, r = tid >> lgL, IC = tid & (L - 1);
    const int KC = k0 + r;
    const int LP = L * (M + 2), CS = LP * N;
    // one halo value per thread: the k0-1 row (r == 0), the k0+EKR row
    // (r == EKR-1), or its own cell once more (same value, same place)
    const bool lo = r == 0, hi = r == EKR - 1;
    const int kh = lo ? (k0 == 0 ? N - 1 : k0 - 1) : hi ? (k0 + EKR == N ? 0 : k0 + EKR) : KC;
    const int gc = IC + LP * KC, gh = IC + LP * kh;
    const int lc = (r + 1) * L + IC, lh = lo ? IC : hi ? (EKP - 1) * L + IC : lc;
    const int ip = IC == L - 1 ? 1 - L : 1, im = IC == 0 ? L - 1 : -1;
    const int kp = KC == N - 1 ? (1 - N) * LP : LP;
    const double* src[3] = {f.h, f.T, f.kap};
    const double *G = f.g, *G0 = f.g0;
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        const int row = L * (j0 + s);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const double a = src[q][gc + row], ah = src[q][gh + row];
            ring[s * SL + q * PL + lc] = a;
            ring[s * SL + q * PL + lh] = ah;
        }
    }
    // a row's values from global memory besides the ring (row JC; v at JC+1)
    struct Row {
        double u0p, up, u0c, uc, v0p, vp, w0p, wp, w0c, wc, rhoh, rhs0;
    };
    auto load_row = [&](int JC, Row& w) {
        const int JL = min(JC, M), c = gc + L * JL;
        w.u0p = G0[c + ip], w.up = G[c + ip], w.u0c = G0[c], w.uc = G[c];
        w.v0p = G0[CS + c + L], w.vp = G[CS + c + L];
        w.w0p = G0[2 * CS + c + kp], w.wp = G[2 * CS + c + kp], w.w0c = G0[2 * CS + c], w.wc = G[2 * CS + c];
        w.rhoh = f.rhoh[c];
        w.rhs0 = f.rhs0[IC + L * (JL - 1) + L * M * KC];
    };
    Row cur;
    load_row(j0 + 1, cur);
    double v0c = G0[CS + gc + L * (j0 + 1)], vc = G[CS + gc + L * (j0 + 1)];
    __builtin_amdgcn_s_waitcnt(0xF70);  // nothing above in flight when the loop starts: s_waitcnt vmcnt(0)
    __syncthreads();
    int sM = 0, s0 = 1, sP = 2;
    for (int j = j0; j < j1; ++j) {
        const int JC = j + 1;
        // the next step's row values, then the ring's next row JC+2 (both
        // clamped on the chunk's last row: loaded, never used)
        Row nxt;
        load_row(JC + 1, nxt);
        const int rq = L * min(JC + 2, M + 1);
        double pn[3], ph[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            pn[q] = src[q][gc + rq];
            ph[q] = src[q][gh + rq];
        }
        auto R = [&](int q) {
            return CellRing{ring + sM * SL + q * PL + lc, ring + s0 * SL + q * PL + lc,
                            ring + sP * SL + q * PL + lc, ip, im, L};
        };
        const EGv q{cur.u0p, cur.up, cur.u0c, cur.uc, cur.v0p, cur.vp, v0c, vc, cur.w0p, cur.wp, cur.w0c, cur.wc};
        const double RHS = energy_rhs_cell(g, R(0), R(1), R(2), q, IC, JC, cthecd, f.wallflux);
        const int c = gc + L * JC;
        f.rhoh[c] = energy_update(cur.rhoh, RHS, cur.rhs0, rk1, rk0, rhoh_cap);
        f.rhs0[IC + L * (JC - 1) + L * M * KC] = RHS;
        __syncthreads();
        // row JC+2 replaces row JC-1 in the ring; v at JC+1 carries over
#pragma unroll
        for (int q2 = 0; q2 < 3; ++q2) {
            ring[sM * SL + q2 * PL + lc] = pn[q2];
            ring[sM * SL + q2 * PL + lh] = ph[q2];
        }
        v0c = cur.v0p, vc = cur.vp;
        cur = nxt;
        __syncthreads();
        const int s_ = sM;
        sM = s0;
        s0 = sP;
        sP = s_;
    }
}

constexpr int EPT = 512;   // threads per property block

// cells per thread and pass of k_energy_props (heated substep at 256x192x256:
// 1 cell 2330 us, 2 cells 2324 us, 4 cells 2310 us, 8 cells 2345 us)
constexpr int ECELL = 4;

// The lookups of k_energy_props without branches, so that the chains of a
// thread's cells interleave: the query is clamped into the table (its result
// is then discarded), ok says whether the value was inside; a NaN passes and
// gives NaN (as the reference's evaluators do).
// S_HT (hack: ENTHALPY_UPDATE's |h| < 1e-10 gives T = 1)
template <int KW>
__device__ __forceinline__ double ht_nb(const PropTab& t, const double* xh, const int* sh, double hv, bool hack,
                                        bool& ok) {
    const bool nan = hv != hv, in = hv > t.h_lo && hv < t.h_hi;
    ok = nan || in;
    const double hq = in ? hv : t.h_lo;
    const int i = find_iv<KW>(xh, sh, t.n, t.nbh, t.h_lo, t.hbi, hq);
    const double4 r = reinterpret_cast<const double4*>(t.rh)[i];
    double T = cubic_nc(r.x, r.y, r.z, r.w, hq - xh[i]);
    if (hack && fabs(hv) < 1.0e-10) T = 1.0;
    return nan ? hv : T;
}

// the interval of T (clamped) and its record's first double4 (S_TD) or all
// four (S_TD, S_TM, S_TK, S_TCp)
template <int KW>
__device__ __forceinline__ int t_iv_nb(const PropTab& t, const double* xt, const int* st, double tv, bool& ok,
                                       double& tq) {
    const bool nan = tv != tv, in = tv > t.t_lo && tv < t.t_hi;
    ok = nan || in;
    tq = in ? tv : t.t_lo;
    return find_iv<KW>(xt, st, t.n, t.nbt, t.t_lo, t.tbi, tq);
}

// rotations of the property history on every cell, the property cascade on
// interior rows (ENTHALPY_UPDATE -> DENSITY_UPDATE -> ENTHALPY_UPDATE ->
// THERM_PROP_UPDATE, src/THERM_PROP_UPDATE.f90); persistent blocks, ECELL
// cells per thread and pass with every load issued first and the lookup
// chains of the cells side by side (the same arithmetic per cell as the
// one-cell form, so the outputs are unchanged bit for bit).
// second = 0: the final THERM_PROP_UPDATE lookup is left to the learned
// surrogate (chs_props_set_mode), which runs after this kernel on h
template <int KW>
__global__ void __launch_bounds__(EPT) k_energy_props(PropTab t, EGeo g, EFields f,
                                                      unsigned long long* __restrict__ cnt, int second) {
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
    for (int q0 = blockIdx.x * (EPT * ECELL) + threadIdx.x; q0 < ntot; q0 += gridDim.x * (EPT * ECELL)) {
        int q[ECELL];
        bool live[ECELL], inner[ECELL];
        double mu_old[ECELL], rho_old[ECELL], rho0_old[ECELL], rh[ECELL];
#pragma unroll
        for (int u = 0; u < ECELL; ++u) {
            const int qq = q0 + u * EPT;
            live[u] = qq < ntot;
            q[u] = live[u] ? qq : q0;
            const int j = (q[u] / L) % (M + 2);
            inner[u] = live[u] && j != 0 && j != M + 1;
            mu_old[u] = f.mu[q[u]];
            rho_old[u] = f.rho[q[u]];
            rho0_old[u] = f.rho0[q[u]];
            rh[u] = f.rhoh[q[u]];
        }
#pragma unroll
        for (int u = 0; u < ECELL; ++u) {
            if (!live[u]) continue;
            f.mu0[q[u]] = mu_old[u];
            f.rho1[q[u]] = rho0_old[u];
            f.rho0[q[u]] = rho_old[u];
        }
        // ENTHALPY_UPDATE, DENSITY_UPDATE: only the density of the first lookup is used
        double D1[ECELL];
        bool ok1[ECELL];
        {
            double T1[ECELL];
            bool okh[ECELL];
#pragma unroll
            for (int u = 0; u < ECELL; ++u) T1[u] = ht_nb<KW>(t, xh, sh, rh[u] / rho_old[u], true, okh[u]);
            int k1[ECELL];
            double tq[ECELL];
            double4 cd[ECELL];
#pragma unroll
            for (int u = 0; u < ECELL; ++u) {
                bool okt;
                k1[u] = t_iv_nb<KW>(t, xt, st, T1[u], okt, tq[u]);
                ok1[u] = okh[u] && okt;
                cd[u] = reinterpret_cast<const double4*>(t.rt)[4 * (size_t)k1[u]];
            }
#pragma unroll
            for (int u = 0; u < ECELL; ++u) {
                double D = cubic_nc(cd[u].x, cd[u].y, cd[u].z, cd[u].w, tq[u] - xt[k1[u]]);
                if (fabs(T1[u] - 1.0) < 1.0e-10) D = 1.0;
                D1[u] = T1[u] != T1[u] ? T1[u] : D;
            }
        }
        // ENTHALPY_UPDATE with the new density, then THERM_PROP_UPDATE
        double h2[ECELL], T2[ECELL];
        bool ok2[ECELL];
        int k2[ECELL];
        double tq2[ECELL];
#pragma unroll
        for (int u = 0; u < ECELL; ++u) {
            h2[u] = rh[u] / D1[u];
            bool okh, okt;
            T2[u] = ht_nb<KW>(t, xh, sh, h2[u], false, okh);
            k2[u] = t_iv_nb<KW>(t, xt, st, T2[u], okt, tq2[u]);
            ok2[u] = okh && okt;
        }
#pragma unroll
        for (int u = 0; u < ECELL; ++u) {
            if (!inner[u]) continue;
            if (!ok1[u]) {
                ++bad;
                continue;
            }
            f.h[q[u]] = h2[u];
            // THERM_PROP_UPDATE leaves |h| < 1e-10 as it is (and the
            // surrogate form does this lookup itself)
            const bool look = second && !(fabs(h2[u]) < 1.0e-10);
            if (look && !ok2[u]) {
                ++bad;
                f.rho[q[u]] = D1[u];
                continue;
            }
            if (!look) {
                f.rho[q[u]] = D1[u];
                continue;
            }
            const double4* r4 = reinterpret_cast<const double4*>(t.rt) + 4 * (size_t)k2[u];
            const double4 cd = r4[0], cm = r4[1], ck = r4[2];
            const double dt = tq2[u] - xt[k2[u]];
            double D = cubic_nc(cd.x, cd.y, cd.z, cd.w, dt);
            double Mu = cubic_nc(cm.x, cm.y, cm.z, cm.w, dt);
            double K = cubic_nc(ck.x, ck.y, ck.z, ck.w, dt);
            if (fabs(T2[u] - 1.0) < 1.0e-10) D = Mu = K = 1.0;
            if (T2[u] != T2[u]) D = Mu = K = T2[u];
            f.T[q[u]] = T2[u];
            f.mu[q[u]] = Mu;
            f.kap[q[u]] = K;
            f.rho[q[u]] = D;
        }
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
    // the range check of the previous call on p, read now (its copy has
    // long finished when the momentum work was queued in between)
    CHS_TRY(energy_range_check(p, "chs_energy_substep"));
    const hipStream_t s = h->stream;
    const double cap = p->rhoh_max + 1.0e-14;   // RHOHmax + REALMIN
    hipEvent_t ev;
    CHS_LAUNCH(k_zero_u64_e, dim3(1), dim3(64), 0, s, p->ecnt, 4);
    // the marching form needs L a power of two with 4L threads in whole
    // waves (at most 1024) and its LDS ring on this device; chs_set_cn_solver
    // mode 2 selects the row-tile form
    const int L = h->L;
    const bool march = h->cn_mode != 2 && (L & (L - 1)) == 0 && L >= 16 && EKR * L <= 1024 && h->N % EKR == 0 &&
                       energy_rhs_m_smem(L) <= (size_t)h->lds_max &&
                       (h->cn_mode == 3 ||
                        march_chunk_rows(h->M, h->N / EKR, energy_rhs_m_smem(L), h->lds_max, false) >= MARCH_MIN);
    h->timer.begin("energy_rhs", s, &ev);
    if (march) {
        const int nkg = h->N / EKR;
        const size_t sm = energy_rhs_m_smem(L);
        // one block per CU: filling the LDS (two per CU at 128^3) measured no faster
        const int jc = march_chunk_rows(h->M, nkg, sm, h->lds_max, false);
        const int nyc = (h->M + jc - 1) / jc;
        (void)hipFuncSetAttribute((const void*)k_energy_rhs_m, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
        CHS_LAUNCH(k_energy_rhs_m, dim3(8 * ((nkg * nyc + 7) / 8)), dim3(EKR * L), sm, s, g, f, TGAM_E[ns] * dt,
                   TROH_E[ns] * dt, cthecd, cap, nyc, jc, __builtin_ctz((unsigned)L));
    } else {
        CHS_LAUNCH(k_energy_rhs, dim3(row_tile_blocks(h->L, h->N, h->M)), dim3(64, 4), 0, s, g, f, TGAM_E[ns] * dt,
                   TROH_E[ns] * dt, cthecd, cap);
    }
    h->timer.end(ev, "energy_rhs", s);
    const PropTab t = tab_of(p);
    h->timer.begin("energy_props", s, &ev);
    CHS_TRY(props_lds_attr((const void*)k_energy_wall, p));
    const bool sur = p->mode == CHS_PROPS_SURROGATE;
#define CHS_EPROPS(KK)                                                                                         \
    do {                                                                                                       \
        CHS_TRY(props_lds_attr((const void*)k_energy_props<KK>, p));                                           \
        CHS_LAUNCH(k_energy_props<KK>, dim3(props_grid_occ(p, (const void*)k_energy_props<KK>, EPT, 1 << 20)),  \
                   dim3(EPT), props_lds_bytes(p), s, t, g, f, p->ecnt, sur ? 0 : 1);                           \
    } while (0)
    switch (p->kwalk) {
        case 2: CHS_EPROPS(2); break;
        case 3: CHS_EPROPS(3); break;
        case 4: CHS_EPROPS(4); break;
        default: CHS_EPROPS(0); break;
    }
#undef CHS_EPROPS
    if (sur) {
        // THERM_PROP_UPDATE of the interior rows by the surrogate where it is
        // certified (the table elsewhere), on this handle's stream
        const Rows rw{(long long)h->N, (long long)h->L * h->M, (long long)h->L * (h->M + 2)};
        const size_t o = (size_t)h->L;   // row J = 1
        CHS_TRY(mlp_update(p, tf->enthalpy + o, rw, tf->temperature + o, tf->density + o, tf->viscosity + o,
                           tf->thermcondt + o, nullptr, p->ecnt, s));
    }
    h->timer.end(ev, "energy_props", s);
    CHS_LAUNCH(k_energy_wall, dim3(64), dim3(256), props_lds_bytes(p), s, t, g, f, p->ecnt);
    // the counts of lookups outside the table and of surrogate fallbacks to
    // pinned host memory, read by the next call on p (or chs_props_check,
    // chs_props_last_fallback) without a stream sync here
    CHS_HIP(hipMemcpyAsync(p->ecnt_host, p->ecnt, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    CHS_HIP(hipEventRecord(p->ecnt_ev, s));
    p->ecnt_pending = true;
    p->ecnt_sur = sur;
    return CHS_OK;
}

}  // namespace chs

extern "C" int chs_energy_substep(chs_handle* h, chs_props* p, const chs_thermal_fields* f, int ns, double dt,
                                  double cthecd) {
    return chs::energy_substep(h, p, f, ns, dt, cthecd);
}
