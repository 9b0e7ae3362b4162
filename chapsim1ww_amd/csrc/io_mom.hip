// The io (main-domain) momentum substep of the heated periodic channel on one
// GPU: variable properties, mass fluxes G = rho u, explicit viscous terms
// (visthemflg = 0), no weighted pressure.
//
//   k_io_prepare   DENSITY_TIME_DERIVATION(NS) (src/THERM_PROP_UPDATE.f90:
//                  225-291), DENSITY_Staggered, MU_Staggered and
//                  VELOCITY_CALC_io (src/MASSFLUX_CALC_IO.f90:82-311)
//   k_io_divu      DIVG_U_io (src/DIVG.f90:153-259)
//   k_io_terms     CONVECTION_X/Y/Z_io (src/CONVECTION_X.f90:77-176,
//                  src/CONVECTION_Y.f90:128-287, src/CONVECTION_Z.f90:115-314),
//                  VISCOUS_ALL_EXPLT_X/Y/Z_io (src/VISCOUS_X_EXPLICIT_io.f90:
//                  1-113, src/VISCOUS_Y_EXPLICIT_io.f90:1-99,
//                  src/VISCOUS_Z_EXPLICIT_io.f90:1-100) and the explicit
//                  branch of RHS_MOM_EXPLICIT_io (src/RHS_MOM_EXPLICIT_io.f90:
//                  1-297) up to DPGRNS, whose volume integral it sums per block
//   k_io_dpgrns    the constant-mass-flux DPGRNS from the block sums
//   k_io_massflux  RHS - DPGRNS, MASSFLUX_CALC_IO (src/MASSFLUX_CALC_IO.f90:
//                  1-79) and BC_WALL_G_io (src/BC_WALL.f90:412-458)
//   k_io_divg      DIVG_io(NS) (src/DIVG.f90:60-150)
//   FFT99 Poisson  the reference's "Method One" for the io domain
//                  (src/SOLVERRK3_MOM_IO.f90:123), Neumann ghosts
//   k_io_update    PRCALC_io (explicit branch, src/PRESSURE_CORRECTION.f90:
//                  88-156) and MASSFLUX_UPDATE_IO (src/VELO_CORRECTION.f90:
//                  109-185) with the wall conditions
//   k_io_divgck    DIVGCK_io (src/DIVGCK.f90:61-124)
//
// INTFC_VARS on one GPU is the periodic y wrap the reference's exchange
// produces across its ranks (row 0 <- row M, row M+1 <- row 1), after which
// the wall routines overwrite what they set.  Channel metrics: RCCI1 = RNDI1 =
// RCCI2 = 1, XND2CL = YND2CL = ZND2CL = 1/2.  Every expression is evaluated
// in the reference's order without contraction into fused multiply-adds, so
// the element-wise results are an FP64 evaluation of the reference's
// expressions; only the DPGRNS volume sum is summed in a different order.
//
// Arrays (x fastest): Q, G, G0, RHO_STG, MU_STG (NCL1, 0:NCL2+1, NCL3, 3);
// PR, DrhoDtP, DivU, DENSITY*, VISCOUSITY* (NCL1, 0:NCL2+1, NCL3);
// EXPLT0 (NCL1, NCL2, NCL3, 3).
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "chs_common.hpp"
#include "tdma_dev.hpp"

namespace chs {

int poisson_solve(chs_handle* h, const double* f, double* dph, int ghost, double abs_tol);
bool tdma_seg_config(int n, bool cyclic, int* MS, int* P, int* LPB, int maxthr, bool lpb_small);

struct IoGeo {
    int L, M, N;
    double dxi, dzi, cvisc, volm;
    const double *dyfi, *dyci, *wff, *wfb;
};

struct IoArrays {
    double *q, *g, *g0, *pr, *drhodtp, *divu, *rs, *ms, *explt0;
    const double *rho, *rho0, *mu, *mu0;
};

struct IoParams {
    int thermal, nflow, flowdrvtp, gravdir;
    double cfgv, f_a;
};

static const double TGAM_IO[4] = {1.0, 8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
static const double TROH_IO[4] = {0.0, 0.0, -17.0 / 60.0, -5.0 / 12.0};

#define IO_CELL(NJ, J0)                                                          \
    const int L = g.L, M = g.M, N = g.N;                                         \
    int it_, jt_, kt_, tile_;                                                    \
    if (!row_tile(L, N, (NJ), &it_, &jt_, &kt_, &tile_)) return;                  \
    const int IC = it_ * 64 + threadIdx.x;                                       \
    const int KC = kt_ * 4 + threadIdx.y;                                        \
    const int JT = jt_ + (J0);                                                   \
    if (IC >= L || KC >= N) return;                                              \
    const int IP = IC == L - 1 ? 0 : IC + 1, IM = IC == 0 ? L - 1 : IC - 1;      \
    const int KP = KC == N - 1 ? 0 : KC + 1, KM = KC == 0 ? N - 1 : KC - 1;      \
    const int LP = L * (M + 2), CS = LP * N;                                     \
    auto at = [&](int i, int j, int k) { return i + L * j + LP * k; };           \
    (void)IP; (void)IM; (void)KP; (void)KM; (void)CS; (void)JT

// DENSITY_TIME_DERIVATION, DENSITY_Staggered, MU_Staggered (thermal = 1) and
// VELOCITY_CALC_io with BC_WALL_Q_io; one thread per cell of rows 0..M+1.
// Every array is distinct (restrict): all loads can be issued before the
// first store.
__global__ void __launch_bounds__(256) k_io_prepare(IoGeo g, const double* __restrict__ D,
                                                    const double* __restrict__ D0, const double* __restrict__ V,
                                                    const double* __restrict__ V0, const double* __restrict__ G,
                                                    const double* __restrict__ G0, double* __restrict__ drhodtp,
                                                    double* __restrict__ rs, double* __restrict__ ms,
                                                    double* __restrict__ q, double* __restrict__ drhoi, int thermal,
                                                    double ca, double cb, double cd, int nj, int j0) {
#pragma clang fp contract(off)
    IO_CELL(nj, j0);   // rows j0 .. j0+nj-1 of 0..M+1
    const int J = JT;   // ghost-aware row 0..M+1
    const int c = at(IC, J, KC);
    double r1, r2, r3;
    if (thermal) {
        // Every branch below only selects values: the ghost rows evaluate
        // the interior expressions at clamped rows and discard them, so the
        // loads form one straight line.
        const bool interior = J >= 1 && J <= M;
        double dr;
        {
            const int Jd = J < 1 ? 1 : (J > M ? M : J), cd0 = at(IC, Jd, KC);
            const double COE2 = ldc(g.dyfi, Jd - 1) * 1.0, COE3 = g.dzi * 1.0;
            const double DIVX = (G[at(IP, Jd, KC)] - G[cd0]) * g.dxi;
            const double DIVX0 = (G0[at(IP, Jd, KC)] - G0[cd0]) * g.dxi;
            const double DIVY = (G[CS + at(IC, Jd + 1, KC)] - G[CS + cd0]) * COE2;
            const double DIVY0 = (G0[CS + at(IC, Jd + 1, KC)] - G0[CS + cd0]) * COE2;
            const double DIVZ = (G[2 * CS + at(IC, Jd, KP)] - G[2 * CS + cd0]) * COE3;
            const double DIVZ0 = (G0[2 * CS + at(IC, Jd, KP)] - G0[2 * CS + cd0]) * COE3;
            dr = (ca * (DIVX + DIVY + DIVZ) + cb * (DIVX0 + DIVY0 + DIVZ0)) / cd;
        }
        // rows whose values the y wrap copies: 0 <- M, M+1 <- 1
        const int Jw = J == 0 ? M : (J == M + 1 ? 1 : J);
        // RHO_STG x and z faces: rows 1..M, wrapped ghosts
        r1 = (D[at(IC, Jw, KC)] + D[at(IM, Jw, KC)] + D0[at(IC, Jw, KC)] + D0[at(IM, Jw, KC)]) * 0.5 * 0.5;
        r3 = (D[at(IC, Jw, KC)] + D[at(IC, Jw, KM)] + D0[at(IC, Jw, KC)] + D0[at(IC, Jw, KM)]) * 0.5 * 0.5;
        // MU_STG x-z edge: rows 1..M, wrapped ghosts
        const double m2 = (V[at(IC, Jw, KC)] + V[at(IM, Jw, KC)] + V0[at(IC, Jw, KC)] + V0[at(IM, Jw, KC)] +
                           V[at(IC, Jw, KM)] + V[at(IM, Jw, KM)] + V0[at(IC, Jw, KM)] + V0[at(IM, Jw, KM)]) *
                          0.5 * 0.5 * 0.5;
        // RHO_STG y face and MU_STG x-y, y-z edges: rows 2..M; row 0 <- M;
        // rows 1 and M+1 are the wall values
        const bool wall = J == 1 || J == M + 1;
        const int Jg = J == M + 1 ? M + 1 : 0;
        const double r2w = 0.5 * (D0[at(IC, Jg, KC)] + D[at(IC, Jg, KC)]);
        const double m1w = 0.5 * (V0[at(IC, Jg, KC)] + V[at(IC, Jg, KC)]);
        const int Jr = J == 0 ? M : J, Jm = Jr - 1;
        const double wff = ldc(g.wff, Jr - 1), wfb = ldc(g.wfb, Jr - 1);
        const double r2i =
            (wff * (D[at(IC, Jr, KC)] + D0[at(IC, Jr, KC)]) + wfb * (D[at(IC, Jm, KC)] + D0[at(IC, Jm, KC)])) * 0.5;
        const double m1i = ((V[at(IC, Jr, KC)] + V[at(IM, Jr, KC)] + V0[at(IC, Jr, KC)] + V0[at(IM, Jr, KC)]) * wff +
                            (V[at(IC, Jm, KC)] + V[at(IM, Jm, KC)] + V0[at(IC, Jm, KC)] + V0[at(IM, Jm, KC)]) * wfb) *
                           0.5 * 0.5;
        const double m3i = ((V[at(IC, Jr, KM)] + V[at(IC, Jr, KC)] + V0[at(IC, Jr, KM)] + V0[at(IC, Jr, KC)]) * wff +
                            (V[at(IC, Jm, KM)] + V[at(IC, Jm, KC)] + V0[at(IC, Jm, KM)] + V0[at(IC, Jm, KC)]) * wfb) *
                           0.5 * 0.5;
        r2 = wall ? r2w : r2i;
        const double m1 = wall ? m1w : m1i, m3 = wall ? m1w : m3i;
        if (drhoi) {
            // DENSITY_implicit_add (src/MASSFLUX_CALC_IO.f90:316-400): x and z
            // faces on rows 1..M with the wrapped ghosts, y faces on rows
            // 2..M, row 0 <- M, zero on the wall faces
            const int Jr2 = J == 0 ? M : J;
            const double d1 = 1.0 / r1 - 1.0 / ((D0[at(IC, Jw, KC)] + D0[at(IM, Jw, KC)]) * 0.5);
            const double d3 = 1.0 / r3 - 1.0 / ((D0[at(IC, Jw, KC)] + D0[at(IC, Jw, KM)]) * 0.5);
            const double d2 = 1.0 / r2i -
                              1.0 / (ldc(g.wff, Jr2 - 1) * D0[at(IC, Jr2, KC)] + ldc(g.wfb, Jr2 - 1) * D0[at(IC, Jr2 - 1, KC)]);
            drhoi[c] = d1;
            drhoi[CS + c] = wall ? 0.0 : d2;
            drhoi[2 * CS + c] = d3;
        }
        if (interior) drhodtp[c] = dr;
        rs[c] = r1;
        rs[CS + c] = r2;
        rs[2 * CS + c] = r3;
        ms[c] = m1;
        ms[CS + c] = m2;
        ms[2 * CS + c] = m3;
    } else {
        r1 = rs[c];
        r2 = rs[CS + c];
        r3 = rs[2 * CS + c];
    }
    // VELOCITY_CALC_io + BC_WALL_Q_io (the wrapped ghost rows are then zeroed)
    if (J == 0 || J == M + 1) {
        q[c] = 0.0;
        q[CS + c] = 0.0;
        q[2 * CS + c] = 0.0;
    } else {
        q[c] = G[c] / r1;
        q[2 * CS + c] = G[2 * CS + c] / r3;
        q[CS + c] = J >= 2 ? G[CS + c] / r2 : 0.0;
    }
}

// k_io_prepare (thermal = 1) on the rows 2..M, marching along y: a thread
// owns the column (IC, KC) of a chunk of rows.  There, Jw = Jr = J and Jm =
// J-1, and the row J-1 parts of RHO_STG y and MU_STG x-y, y-z are the row
// parts of the step before (the same expressions), so they carry over in
// registers with v of G and G0 at J; the values of row J+1 are loaded one
// step ahead.  Rows 0, 1 and M+1 (wrap and wall rows) are k_io_prepare's.
struct PrepRow {
    double d, dim, dkm, d0, d0im, d0km;              // RHO at the cell, i-1, k-1; RHO0 likewise
    double v, vim, vkm, vimkm, v0, v0im, v0km, v0imkm;  // MU, MU0 at the cell, i-1, k-1, i-1 k-1
    double gu, gup, g0u, g0up, gvp, g0vp, gw, gwp, g0w, g0wp;  // G, G0: u at i, i+1; v at j+1; w at k, k+1
};

__global__ void __launch_bounds__(256) k_io_prepare_m(IoGeo g, const double* __restrict__ D,
                                                      const double* __restrict__ D0, const double* __restrict__ V,
                                                      const double* __restrict__ V0, const double* __restrict__ G,
                                                      const double* __restrict__ G0, double* __restrict__ drhodtp,
                                                      double* __restrict__ rs, double* __restrict__ ms,
                                                      double* __restrict__ q, double* __restrict__ drhoi, double ca,
                                                      double cb, double cd, int nyc, int jc) {
#pragma clang fp contract(off)
    const int L = g.L, M = g.M, N = g.N;
    const int ni = (L + 63) / 64, nk = (N + 3) / 4;
    const int nt = ni * nk * nyc, per = (nt + 7) / 8;
    const int t = (blockIdx.x % 8) * per + blockIdx.x / 8;  // XCD x: tiles x*per .. (x+1)*per-1
    if (t >= nt) return;
    const int it = t % ni, kt = (t / ni) % nk, yc = t / (ni * nk);
    const int IC = it * 64 + threadIdx.x, KC = kt * 4 + threadIdx.y;
    if (IC >= L || KC >= N) return;
    const int J0 = 2 + yc * jc, J1 = min(M + 1, J0 + jc);   // rows J0 .. J1-1 of 2..M
    if (J0 >= J1) return;
    const int LP = L * (M + 2), CS = LP * N;
    const int ip = IC == L - 1 ? 1 - L : 1, im = IC == 0 ? L - 1 : -1;
    const int kp = KC == N - 1 ? (1 - N) * LP : LP, km = KC == 0 ? (N - 1) * LP : -LP;
    const int col = IC + LP * KC;
    auto load_row = [&](int J, PrepRow& r) {
        const int c = col + L * min(J, M);
        r.d = D[c], r.dim = D[c + im], r.dkm = D[c + km];
        r.d0 = D0[c], r.d0im = D0[c + im], r.d0km = D0[c + km];
        r.v = V[c], r.vim = V[c + im], r.vkm = V[c + km], r.vimkm = V[c + im + km];
        r.v0 = V0[c], r.v0im = V0[c + im], r.v0km = V0[c + km], r.v0imkm = V0[c + im + km];
        r.gu = G[c], r.gup = G[c + ip], r.g0u = G0[c], r.g0up = G0[c + ip];
        r.gvp = G[CS + c + L], r.g0vp = G0[CS + c + L];
        r.gw = G[2 * CS + c], r.gwp = G[2 * CS + c + kp], r.g0w = G0[2 * CS + c], r.g0wp = G0[2 * CS + c + kp];
    };
    // the row J0-1 parts, and v at J0
    double sdm, s1m, s3m, d0m, gv, g0v;
    {
        const int c = col + L * (J0 - 1);
        sdm = D[c] + D0[c];
        s1m = V[c] + V[c + im] + V0[c] + V0[c + im];
        s3m = V[c + km] + V[c] + V0[c + km] + V0[c];
        d0m = D0[c];
        gv = G[CS + c + L];
        g0v = G0[CS + c + L];
    }
    PrepRow cur;
    load_row(J0, cur);
    for (int J = J0; J < J1; ++J) {
        PrepRow nxt;
        load_row(J + 1, nxt);   // clamped on the chunk's last row: loaded, never used
        const int c = col + L * J;
        // DENSITY_TIME_DERIVATION
        const double COE2 = ldc(g.dyfi, J - 1) * 1.0, COE3 = g.dzi * 1.0;
        const double DIVX = (cur.gup - cur.gu) * g.dxi;
        const double DIVX0 = (cur.g0up - cur.g0u) * g.dxi;
        const double DIVY = (cur.gvp - gv) * COE2;
        const double DIVY0 = (cur.g0vp - g0v) * COE2;
        const double DIVZ = (cur.gwp - cur.gw) * COE3;
        const double DIVZ0 = (cur.g0wp - cur.g0w) * COE3;
        const double dr = (ca * (DIVX + DIVY + DIVZ) + cb * (DIVX0 + DIVY0 + DIVZ0)) / cd;
        // RHO_STG and MU_STG (rows 2..M: no wrap, no wall)
        const double r1 = (cur.d + cur.dim + cur.d0 + cur.d0im) * 0.5 * 0.5;
        const double r3 = (cur.d + cur.dkm + cur.d0 + cur.d0km) * 0.5 * 0.5;
        const double m2 = (cur.v + cur.vim + cur.v0 + cur.v0im + cur.vkm + cur.vimkm + cur.v0km + cur.v0imkm) *
                          0.5 * 0.5 * 0.5;
        const double wff = ldc(g.wff, J - 1), wfb = ldc(g.wfb, J - 1);
        const double sd = cur.d + cur.d0;
        const double s1 = cur.v + cur.vim + cur.v0 + cur.v0im;
        const double s3 = cur.vkm + cur.v + cur.v0km + cur.v0;
        const double r2 = (wff * sd + wfb * sdm) * 0.5;
        const double m1 = (s1 * wff + s1m * wfb) * 0.5 * 0.5;
        const double m3 = (s3 * wff + s3m * wfb) * 0.5 * 0.5;
        if (drhoi) {
            const double d1 = 1.0 / r1 - 1.0 / ((cur.d0 + cur.d0im) * 0.5);
            const double d3 = 1.0 / r3 - 1.0 / ((cur.d0 + cur.d0km) * 0.5);
            const double d2 = 1.0 / r2 - 1.0 / (wff * cur.d0 + wfb * d0m);
            drhoi[c] = d1;
            drhoi[CS + c] = d2;
            drhoi[2 * CS + c] = d3;
        }
        drhodtp[c] = dr;
        rs[c] = r1;
        rs[CS + c] = r2;
        rs[2 * CS + c] = r3;
        ms[c] = m1;
        ms[CS + c] = m2;
        ms[2 * CS + c] = m3;
        // VELOCITY_CALC_io
        q[c] = cur.gu / r1;
        q[2 * CS + c] = cur.gw / r3;
        q[CS + c] = gv / r2;
        sdm = sd, s1m = s1, s3m = s3, d0m = cur.d0;
        gv = cur.gvp, g0v = cur.g0vp;
        cur = nxt;
    }
}

// DIVG_U_io on rows 1..M and the y wrap of its ghost rows
__global__ void __launch_bounds__(256) k_io_divu(IoGeo g, const double* __restrict__ Q, double* __restrict__ divu) {
#pragma clang fp contract(off)
    IO_CELL(g.M + 2, 0);
    const int J = JT;
    const int Jr = J == 0 ? M : (J == M + 1 ? 1 : J);
    const int c = at(IC, Jr, KC);
    const double DIVX = (Q[at(IP, Jr, KC)] - Q[c]) * g.dxi;
    const double DIVY = (Q[CS + at(IC, Jr + 1, KC)] - Q[CS + c]) * (ldc(g.dyfi, Jr - 1) * 1.0);
    const double DIVZ = (Q[2 * CS + at(IC, Jr, KP)] - Q[2 * CS + c]) * (g.dzi * 1.0);
    divu[at(IC, J, KC)] = (DIVX + DIVY + DIVZ) / 3.0;
}

template <int NT>
__device__ __forceinline__ double io_block_sum(double v, double* red) {
    const int t = threadIdx.x + 64 * threadIdx.y;
    red[t] = v;
    __syncthreads();
#pragma unroll
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (t < s) red[t] += red[t + s];
        __syncthreads();
    }
    return red[0];
}

// Values of one array around the cell held in registers, for the offsets
// the terms use: MU, MU0, DivU and PR at (0,0,0), (-1,0,0), (0,-1,0),
// (0,0,-1); MU_STG x-y at (0,0,0), (0,1,0), (1,0,0); x-z at (0,0,0), (0,0,1),
// (1,0,0); y-z at (0,0,0), (0,1,0), (0,0,1).
struct IoV4 {
    double c, im, jm, km;
    __device__ __forceinline__ double operator()(int di, int dj, int dk) const {
        return di < 0 ? im : dj < 0 ? jm : dk < 0 ? km : c;
    }
};
struct IoVS1 {
    double c, jp, ip;
    __device__ __forceinline__ double operator()(int di, int dj, int) const { return dj > 0 ? jp : di > 0 ? ip : c; }
};
struct IoVS2 {
    double c, kp, ip;
    __device__ __forceinline__ double operator()(int di, int, int dk) const { return dk > 0 ? kp : di > 0 ? ip : c; }
};
struct IoVS3 {
    double c, jp, kp;
    __device__ __forceinline__ double operator()(int, int dj, int dk) const { return dj > 0 ? jp : dk > 0 ? kp : c; }
};

// CONVECTION_X/Y/Z_io at the cell: hx, hy, hz.  hy is evaluated on row 1
// too (from in-bounds rows); the callers discard it there.
template <class A>
__device__ __forceinline__ void io_conv_a(const IoGeo& g, const A& G1, const A& G2, const A& G3, const A& Q1,
                                          const A& Q2, const A& Q3, int JC, double& hx, double& hy, double& hz) {
#pragma clang fp contract(off)
    const int JJ = JC, JJP = JC + 1;
    const double wff = ldc(g.wff, JJ - 1), wfb = ldc(g.wfb, JJ - 1), wffp = ldc(g.wff, JJP - 1),
                 wfbp = ldc(g.wfb, JJP - 1);
    {
        const double COE1 = g.dxi * 0.5 * 0.5, COE2 = ldc(g.dyfi, JJ - 1) * 0.5 * 1.0;
        const double COE32 = g.dzi * 0.5 * 0.5 * 1.0;
        const double H11F = (G1(1, 0, 0) + G1(0, 0, 0)) * (Q1(1, 0, 0) + Q1(0, 0, 0));
        const double H11B = (G1(0, 0, 0) + G1(-1, 0, 0)) * (Q1(0, 0, 0) + Q1(-1, 0, 0));
        const double H11 = (H11F - H11B) * COE1;
        const double H12F = (G2(0, 1, 0) + G2(-1, 1, 0)) * (wffp * Q1(0, 1, 0) + wfbp * Q1(0, 0, 0));
        const double H12B = (G2(0, 0, 0) + G2(-1, 0, 0)) * (wff * Q1(0, 0, 0) + wfb * Q1(0, -1, 0));
        const double H12 = (H12F - H12B) * COE2;
        const double H13F = (G3(0, 0, 1) + G3(-1, 0, 1)) * (Q1(0, 0, 1) + Q1(0, 0, 0));
        const double H13B = (G3(0, 0, 0) + G3(-1, 0, 0)) * (Q1(0, 0, 0) + Q1(0, 0, -1));
        const double H13 = (H13F - H13B) * COE32;
        hx = -(H11 + H12 + H13);
    }
    {
        const double COE1 = g.dxi * 0.5, COE2 = ldc(g.dyci, JJ - 1) * 0.5 * 0.5, COE31 = g.dzi * 0.5 * 1.0;
        const double H21F = (wff * G1(1, 0, 0) + wfb * G1(1, -1, 0)) * (Q2(1, 0, 0) + Q2(0, 0, 0));
        const double H21B = (wff * G1(0, 0, 0) + wfb * G1(0, -1, 0)) * (Q2(0, 0, 0) + Q2(-1, 0, 0));
        const double H21 = (H21F - H21B) * COE1;
        const double H22F = (G2(0, 1, 0) + G2(0, 0, 0)) * (Q2(0, 1, 0) + Q2(0, 0, 0));
        const double H22B = (G2(0, 0, 0) + G2(0, -1, 0)) * (Q2(0, 0, 0) + Q2(0, -1, 0));
        const double H22 = (H22F - H22B) * COE2;
        const double H23F = (wff * G3(0, 0, 1) + wfb * G3(0, -1, 1)) * (Q2(0, 0, 1) + Q2(0, 0, 0));
        const double H23B = (wff * G3(0, 0, 0) + wfb * G3(0, -1, 0)) * (Q2(0, 0, 0) + Q2(0, 0, -1));
        const double H23 = (H23F - H23B) * COE31;
        hy = -(H21 + H22 + H23);
    }
    {
        const double COE1 = g.dxi * 0.5 * 0.5, COE2 = ldc(g.dyfi, JJ - 1) * 0.5;
        const double COE31 = g.dzi * 0.5 * 0.5 * 1.0;
        const double H31F = (G1(1, 0, 0) + G1(1, 0, -1)) * (Q3(1, 0, 0) + Q3(0, 0, 0));
        const double H31B = (G1(0, 0, 0) + G1(0, 0, -1)) * (Q3(0, 0, 0) + Q3(-1, 0, 0));
        const double H31 = (H31F - H31B) * COE1;
        const double H32F = (G2(0, 1, 0) + G2(0, 1, -1)) * (wffp * Q3(0, 1, 0) + wfbp * Q3(0, 0, 0));
        const double H32B = (G2(0, 0, 0) + G2(0, 0, -1)) * (wff * Q3(0, 0, 0) + wfb * Q3(0, -1, 0));
        const double H32 = (H32F - H32B) * COE2;
        const double H33F = (G3(0, 0, 1) + G3(0, 0, 0)) * (Q3(0, 0, 1) + Q3(0, 0, 0));
        const double H33B = (G3(0, 0, 0) + G3(0, 0, -1)) * (Q3(0, 0, 0) + Q3(0, 0, -1));
        const double H33 = (H33F - H33B) * COE31;
        hz = -(H31 + H32 + H33);
    }
}

// The global-array form, as the RK-CN terms kernel calls it (c = (IC, JC, KC)).
__device__ __forceinline__ void io_conv(const IoGeo& g, const double* __restrict__ G, const double* __restrict__ Q,
                                        int IC, int IP, int IM, int JC, int KC, int KP, int KM, double& hx,
                                        double& hy, double& hz) {
    const int L = g.L, LP = L * (g.M + 2), CS = LP * g.N;
    const int c = IC + L * JC + LP * KC, ip = IP - IC, im = IM - IC, kp = (KP - KC) * LP, km = (KM - KC) * LP;
    io_conv_a(g, cell_acc(G, c, ip, im, kp, km, L), cell_acc(G + CS, c, ip, im, kp, km, L),
              cell_acc(G + 2 * CS, c, ip, im, kp, km, L), cell_acc(Q, c, ip, im, kp, km, L),
              cell_acc(Q + CS, c, ip, im, kp, km, L), cell_acc(Q + 2 * CS, c, ip, im, kp, km, L), JC, hx, hy, hz);
}

// Explicit terms of one cell (CONVECTION_*_io, VISCOUS_ALL_EXPLT_*_io) and
// RHS_MOM_EXPLICIT_io up to DPGRNS: h = the explicit terms, r = the RHS.
// G*, Q* by AQ; MU, MU0, DivU, MU_STG, PR by AS; E = EXPLT0 of the cell,
// rsg = RHO_STG of the GRAV component (unused for GRAV = 0).  Row 1 of v is
// evaluated with a clamped metric index and discarded by the callers.
template <int GRAV, class AQ, class AM, class AS1, class AS2, class AS3, class AP>
__device__ __forceinline__ void io_terms_cell(const IoGeo& g, const AQ& G1, const AQ& G2, const AQ& G3,
                                              const AQ& Q1, const AQ& Q2, const AQ& Q3, const AM& MU,
                                              const AM& MU0, const AM& DU, const AS1& MS1, const AS2& MS2,
                                              const AS3& MS3, const AP& PR, int JC, const double* E, double rsg,
                                              const IoParams& p, double tg, double tr, double dt, double adt,
                                              double* h, double* r) {
#pragma clang fp contract(off)
    const int JJ = JC, JJP = JC + 1, JJM = JC >= 2 ? JC - 1 : 1;
    const double CV = g.cvisc;
    double hx, hy, hz;
    io_conv_a(g, G1, G2, G3, Q1, Q2, Q3, JC, hx, hy, hz);
    // ---- VISCOUS_ALL_EXPLT_X_io ----
    {
        const double COE1 = 2.0 * g.dxi * CV, COE2 = ldc(g.dyfi, JJ - 1) * CV * 1.0, COE32 = g.dzi * CV * 1.0;
        double DUDX = (Q1(1, 0, 0) - Q1(0, 0, 0)) * g.dxi;
        const double TAU11F = 0.5 * (MU0(0, 0, 0) + MU(0, 0, 0)) * (DUDX - DU(0, 0, 0));
        DUDX = (Q1(0, 0, 0) - Q1(-1, 0, 0)) * g.dxi;
        const double TAU11B = 0.5 * (MU0(-1, 0, 0) + MU(-1, 0, 0)) * (DUDX - DU(-1, 0, 0));
        const double DTAU11DX = (TAU11F - TAU11B) * COE1;
        double DUDY = (Q1(0, 1, 0) - Q1(0, 0, 0)) * ldc(g.dyci, JJP - 1) / 1.0;
        double DVDX = (Q2(0, 1, 0) - Q2(-1, 1, 0)) * g.dxi;
        const double TAU12F = MS1(0, 1, 0) * (DUDY + DVDX);
        DUDY = (Q1(0, 0, 0) - Q1(0, -1, 0)) * ldc(g.dyci, JJ - 1) / 1.0;
        DVDX = (Q2(0, 0, 0) - Q2(-1, 0, 0)) * g.dxi;
        const double TAU12B = MS1(0, 0, 0) * (DUDY + DVDX);
        const double DTAU12DY = (TAU12F - TAU12B) * COE2;
        double DUDZ = (Q1(0, 0, 1) - Q1(0, 0, 0)) * g.dzi;
        double DWDX = (Q3(0, 0, 1) - Q3(-1, 0, 1)) * g.dxi;
        const double TAU13F = MS2(0, 0, 1) * (DUDZ + DWDX);
        DUDZ = (Q1(0, 0, 0) - Q1(0, 0, -1)) * g.dzi;
        DWDX = (Q3(0, 0, 0) - Q3(-1, 0, 0)) * g.dxi;
        const double TAU13B = MS2(0, 0, 0) * (DUDZ + DWDX);
        const double DTAU13DZ = (TAU13F - TAU13B) * COE32;
        hx = hx + DTAU11DX + DTAU12DY + DTAU13DZ;
    }
    // ---- VISCOUS_ALL_EXPLT_Y_io (rows 2..M) ----
    {
        const double COE1 = g.dxi * CV, COE3 = g.dzi * CV, COE2 = 2.0 * ldc(g.dyci, JJ - 1) * CV;
        double DVDY = (Q2(0, 1, 0) - Q2(0, 0, 0)) * ldc(g.dyfi, JJ - 1);
        const double TAU22F = 0.5 * (MU0(0, 0, 0) + MU(0, 0, 0)) * (DVDY - DU(0, 0, 0));
        DVDY = (Q2(0, 0, 0) - Q2(0, -1, 0)) * ldc(g.dyfi, JJM - 1);
        const double TAU22B = 0.5 * (MU0(0, -1, 0) + MU(0, -1, 0)) * (DVDY - DU(0, -1, 0));
        const double DTAU22DY = (TAU22F - TAU22B) * COE2;
        double DVDX = (Q2(1, 0, 0) - Q2(0, 0, 0)) * g.dxi;
        double DUDY = (Q1(1, 0, 0) - Q1(1, -1, 0)) * ldc(g.dyci, JJ - 1);
        const double TAU21F = MS1(1, 0, 0) * (DVDX + DUDY);
        DVDX = (Q2(0, 0, 0) - Q2(-1, 0, 0)) * g.dxi;
        DUDY = (Q1(0, 0, 0) - Q1(0, -1, 0)) * ldc(g.dyci, JJ - 1);
        const double TAU21B = MS1(0, 0, 0) * (DVDX + DUDY);
        const double DTAU21DX = (TAU21F - TAU21B) * COE1;
        double DVDZ = (Q2(0, 0, 1) - Q2(0, 0, 0)) * g.dzi;
        double DWDY = (Q3(0, 0, 1) - Q3(0, -1, 1)) * ldc(g.dyci, JJ - 1);
        const double TAU23F = MS3(0, 0, 1) * (DVDZ + DWDY);
        DVDZ = (Q2(0, 0, 0) - Q2(0, 0, -1)) * g.dzi;
        DWDY = (Q3(0, 0, 0) - Q3(0, -1, 0)) * ldc(g.dyci, JJ - 1);
        const double TAU23B = MS3(0, 0, 0) * (DVDZ + DWDY);
        const double DTAU23DZ = (TAU23F - TAU23B) * COE3;
        hy = hy + DTAU21DX + DTAU22DY + DTAU23DZ;
    }
    // ---- VISCOUS_ALL_EXPLT_Z_io ----
    {
        const double COE1 = g.dxi * CV, COE3 = 2.0 * g.dzi * CV, COE2 = ldc(g.dyfi, JJ - 1) * CV;
        double DWDZ = (Q3(0, 0, 1) - Q3(0, 0, 0)) * g.dzi;
        const double TAU33F = 0.5 * (MU0(0, 0, 0) + MU(0, 0, 0)) * (DWDZ - DU(0, 0, 0));
        DWDZ = (Q3(0, 0, 0) - Q3(0, 0, -1)) * g.dzi;
        const double TAU33B = 0.5 * (MU0(0, 0, -1) + MU(0, 0, -1)) * (DWDZ - DU(0, 0, -1));
        const double DTAU33DZ = (TAU33F - TAU33B) * COE3;
        double DWDX = (Q3(1, 0, 0) - Q3(0, 0, 0)) * g.dxi;
        double DUDZ = (Q1(1, 0, 0) - Q1(1, 0, -1)) * g.dzi;
        const double TAU31F = MS2(1, 0, 0) * (DWDX + DUDZ);
        DWDX = (Q3(0, 0, 0) - Q3(-1, 0, 0)) * g.dxi;
        DUDZ = (Q1(0, 0, 0) - Q1(0, 0, -1)) * g.dzi;
        const double TAU31B = MS2(0, 0, 0) * (DWDX + DUDZ);
        const double DTAU31DX = (TAU31F - TAU31B) * COE1;
        double DWDY = (Q3(0, 1, 0) - Q3(0, 0, 0)) * ldc(g.dyci, JJP - 1);
        double DVDZ = (Q2(0, 1, 0) - Q2(0, 1, -1)) * g.dzi;
        const double TAU32F = MS3(0, 1, 0) * (DWDY + DVDZ);
        DWDY = (Q3(0, 0, 0) - Q3(0, -1, 0)) * ldc(g.dyci, JJ - 1);
        DVDZ = (Q2(0, 0, 0) - Q2(0, 0, -1)) * g.dzi;
        const double TAU32B = MS3(0, 0, 0) * (DWDY + DVDZ);
        const double DTAU32DY = (TAU32F - TAU32B) * COE2;
        hz = hz + DTAU31DX + DTAU32DY + DTAU33DZ;
    }
    // ---- RHS_MOM_EXPLICIT_io: RK3 combine, pressure gradient, buoyancy ----
    const double pc = PR(0, 0, 0);
    const double rx = (tg * hx + tr * E[0]) * dt - (pc - PR(-1, 0, 0)) * (adt * g.dxi);
    const double ry = (tg * hy + tr * E[1]) * dt - (pc - PR(0, -1, 0)) * (adt * ldc(g.dyci, JC - 1) / 1.0);
    const double rz = (tg * hz + tr * E[2]) * dt - (pc - PR(0, 0, -1)) * (adt * g.dzi);
    double r1 = rx, r2 = ry, r3 = rz;
    if (GRAV == 1) r1 = r1 + rsg * (adt * p.f_a / 1.0);
    if (GRAV == 2) r2 = r2 + rsg * (adt * p.f_a / 1.0);
    if (GRAV == 3) r3 = r3 + rsg * (adt * p.f_a / 1.0);
    h[0] = hx;
    h[1] = hy;
    h[2] = hz;
    r[0] = r1;
    r[1] = r2;
    r[2] = r3;
}

// The outputs of one cell of the terms kernels: EXPLT0, then either inc =
// RHS (G0 == nullptr) or inc = G + RHS with G0 <- G (g* = G of the cell,
// gl* = G of the ghost row next to rows 1 and M).  Returns the cell's
// mass-flux integrand.
__device__ __forceinline__ double io_terms_store(const IoParams& p, const IoGeo& g, int JC, int e, int NC,
                                                 const double* h, const double* r, double* __restrict__ E,
                                                 double* __restrict__ inc, double* __restrict__ G0, int c, int cg,
                                                 int CS, const double* gv, const double* glv) {
#pragma clang fp contract(off)
    E[e] = h[0];
    E[2 * NC + e] = h[2];
    if (JC >= 2) E[NC + e] = h[1];
    if (G0) {
        inc[e] = gv[0] + r[0];
        inc[NC + e] = JC >= 2 ? gv[1] + r[1] : 0.0;
        inc[2 * NC + e] = gv[2] + r[2];
        G0[c] = gv[0];
        G0[CS + c] = gv[1];
        G0[2 * CS + c] = gv[2];
        if (JC == 1 || JC == g.M) {  // the ghost row next to this one
            G0[cg] = glv[0];
            G0[CS + cg] = glv[1];
            G0[2 * CS + cg] = glv[2];
        }
    } else {
        inc[e] = r[0];
        inc[2 * NC + e] = r[2];
        if (JC >= 2) inc[NC + e] = r[1];
    }
    const double rf = p.nflow == 1 ? r[0] : (p.nflow == 2 ? (JC >= 2 ? r[1] : 0.0) : r[2]);
    return rf / ldc(g.dyfi, JC - 1) / 1.0;
}

// Explicit terms and RHS_MOM_EXPLICIT_io up to DPGRNS, rows 1..M, one
// thread per cell in row tiles.  explt0 <- the explicit terms, part[block]
// <- sum of RHS/DYFI of the NFLOW component, and either (G0 == nullptr)
// inc[c] <- RHS of each component (row 1 of y untouched), or (G0 != nullptr)
// inc[c] <- G + RHS (0 for v on row 1) and G0 <- G on rows 0..M+1
// (SOLVERRK3_MOM_IO's G0 = G, in the pass that reads G anyway).  In the
// second form DPGRNS is taken off later, by k_io_update; DIVG_io's
// differences along the flow direction do not see that uniform shift.
template <int GRAV>
__global__ void __launch_bounds__(256) k_io_terms(IoGeo g, const double* __restrict__ G, const double* __restrict__ Q,
                                                  const double* __restrict__ MU, const double* __restrict__ MU0,
                                                  const double* __restrict__ DU, const double* __restrict__ MS,
                                                  const double* __restrict__ PR, const double* __restrict__ RS,
                                                  double* __restrict__ E, IoParams p, double tg, double tr, double dt,
                                                  double adt, double* __restrict__ inc, double* __restrict__ part,
                                                  double* __restrict__ G0) {
#pragma clang fp contract(off)
    __shared__ double red[256];
    const int L = g.L, M = g.M, N = g.N;
    int it, jt, kt, tile;
    if (!row_tile(L, N, M, &it, &jt, &kt, &tile)) return;   // padding block: whole block exits
    const int IC = it * 64 + threadIdx.x;
    const int KC = kt * 4 + threadIdx.y;
    const int JC = jt + 1;
    const bool live = IC < L && KC < N;
    double flow = 0.0;
    if (live) {
        const int LP = L * (M + 2), CS = LP * N, NC = L * M * N;
        const int ip = IC == L - 1 ? 1 - L : 1, im = IC == 0 ? L - 1 : -1;
        const int kp = KC == N - 1 ? (1 - N) * LP : LP, km = KC == 0 ? (N - 1) * LP : -LP;
        const int c = IC + L * JC + LP * KC;
        const int e = IC + L * (JC - 1) + L * M * KC;   // (L, M, N) index
        auto A = [&](const double* b) { return cell_acc(b, c, ip, im, kp, km, L); };
        const double Ev[3] = {E[e], E[NC + e], E[2 * NC + e]};
        const double rsg = GRAV == 0 ? 0.0 : RS[(GRAV - 1) * CS + c];
        double hv[3], rv[3];
        io_terms_cell<GRAV>(g, A(G), A(G + CS), A(G + 2 * CS), A(Q), A(Q + CS), A(Q + 2 * CS), A(MU), A(MU0),
                            A(DU), A(MS), A(MS + CS), A(MS + 2 * CS), A(PR), JC, Ev, rsg, p, tg, tr, dt, adt, hv, rv);
        const int cg = c + (JC == 1 ? -L : L);
        const double gv[3] = {G[c], G[CS + c], G[2 * CS + c]};
        double glv[3] = {0.0, 0.0, 0.0};
        if (G0 && (JC == 1 || JC == M)) {
            glv[0] = G[cg];
            glv[1] = G[CS + cg];
            glv[2] = G[2 * CS + cg];
        }
        flow = io_terms_store(p, g, JC, e, NC, hv, rv, E, inc, G0, c, cg, CS, gv, glv);
    }
    const double bs = io_block_sum<256>(flow, red);
    if (threadIdx.x == 0 && threadIdx.y == 0) part[tile] = bs;
}

// ---------------------------------------------------------------------------
// k_io_terms_m: k_io_terms marching along y.  A block owns IOKR = 2
// consecutive k rows of every x line and a chunk of rows j0 .. j1-1 (JC =
// j+1); thread (rb, i) of its 2L threads handles the cell (i, k0+rb) of each
// row.  LDS holds a ring of the rows JC-1, JC, JC+1 of G and Q (six
// components, IOKR + 2 k rows each: the periodic z neighbours as halo rows),
// so their stencil reads come from LDS and each value comes from global
// memory once per block; the ring's next row is loaded while the current one
// is computed.  MU, MU0, DivU, MU_STG, PR, RHO_STG and EXPLT0 are read from
// global memory as in k_io_terms.  The mass-flux partial of row j goes to
// part[kg*M + j].
// ---------------------------------------------------------------------------
constexpr int IOKR = 2, IOKP = IOKR + 2;

inline size_t io_terms_m_smem(int L) { return sizeof(double) * ((size_t)3 * 6 * IOKP * L + 16); }

template <int GRAV>
__global__ void __launch_bounds__(512) k_io_terms_m(IoGeo g, const double* __restrict__ G,
                                                    const double* __restrict__ Q, const double* __restrict__ MU,
                                                    const double* __restrict__ MU0, const double* __restrict__ DU,
                                                    const double* __restrict__ MS, const double* __restrict__ PR,
                                                    const double* __restrict__ RS, double* __restrict__ E,
                                                    IoParams p, double tg, double tr, double dt, double adt,
                                                    double* __restrict__ inc, double* __restrict__ part,
                                                    double* __restrict__ G0, int nyc, int jc, int lgL) {
#pragma clang fp contract(off)
    extern __shared__ double sm[];
    const int L = g.L, M = g.M, N = g.N;
    const int PL = IOKP * L, SL = 6 * PL;  // one component / one ring row
    double* ring = sm;
    double* red = sm + 3 * SL;
    const int nt = (N / IOKR) * nyc, per = (nt + 7) / 8;
    const int t = (blockIdx.x % 8) * per + blockIdx.x / 8;  // XCD x: tiles x*per .. (x+1)*per-1
    if (t >= nt) return;
    const int kg = t / nyc;
    const int j0 = (t - kg * nyc) * jc, j1 = min(M, j0 + jc);
    const int k0 = kg * IOKR;
    const int tid = threadIdx.x, rb = tid >> lgL, i = tid & (L - 1);
    const int KC = k0 + rb;
    const int LP = L * (M + 2), CS = LP * N, NC = L * M * N;
    // the cell's halo row: k0-1 (rb == 0) or k0+IOKR (rb == 1)
    const int kh = rb == 0 ? (k0 == 0 ? N - 1 : k0 - 1) : (k0 + IOKR == N ? 0 : k0 + IOKR);
    const int gc = i + LP * KC, gh = i + LP * kh;                   // global, row 0
    const int lc = (rb + 1) * L + i, lh = (rb == 0 ? 0 : IOKP - 1) * L + i;  // in a ring plane
    const int ip = i == L - 1 ? 1 - L : 1, im = i == 0 ? L - 1 : -1;
    const int kp = KC == N - 1 ? (1 - N) * LP : LP, km = KC == 0 ? (N - 1) * LP : -LP;
    const double *MS1 = MS, *MS2 = MS + CS, *MS3 = MS + 2 * CS;
    // The values from global memory a step needs besides the ring, loaded one
    // step ahead: row JC of MU, MU0, DivU, PR (cell, i-1, k-1), MU_STG (the
    // cell's x-neighbour and z-neighbour entries), EXPLT0 and RHO_STG, and
    // row JC+1 of MU_STG x-y and y-z; the j-1 entries of MU, MU0, DivU, PR and
    // the cell entries of MU_STG x-y, y-z carry over from the step before.
    struct Row {
        double mu[3], mu0[3], du[3], pr[3];  // cell, i-1, k-1
        double s1ip, s2[3], s3kp;            // MS1(i+1); MS2 cell, k+1, i+1; MS3(k+1)
        double s1jp, s3jp;                   // MS1, MS3 at row JC+1
        double e[3], rs;
    };
    auto load_row = [&](int JC, Row& r) {
        const int JL = min(JC, M), c = gc + L * JL, c1 = gc + L * min(JC + 1, M + 1);
        const int e = i + L * (JL - 1) + L * M * KC;
        r.mu[0] = MU[c], r.mu[1] = MU[c + im], r.mu[2] = MU[c + km];
        r.mu0[0] = MU0[c], r.mu0[1] = MU0[c + im], r.mu0[2] = MU0[c + km];
        r.du[0] = DU[c], r.du[1] = DU[c + im], r.du[2] = DU[c + km];
        r.pr[0] = PR[c], r.pr[1] = PR[c + im], r.pr[2] = PR[c + km];
        r.s1ip = MS1[c + ip];
        r.s2[0] = MS2[c], r.s2[1] = MS2[c + kp], r.s2[2] = MS2[c + ip];
        r.s3kp = MS3[c + kp];
        r.s1jp = MS1[c1], r.s3jp = MS3[c1];
        r.e[0] = E[e], r.e[1] = E[NC + e], r.e[2] = E[2 * NC + e];
        r.rs = GRAV == 0 ? 0.0 : RS[(GRAV - 1) * CS + c];
    };
    // fill the ring with rows j0 .. j0+2 (JC-1 .. JC+1 of the first step)
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        const int row = L * (j0 + s);
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const double* src = q < 3 ? G + q * CS : Q + (q - 3) * CS;
            const double a = src[gc + row], ah = src[gh + row];
            ring[s * SL + q * PL + lc] = a;
            ring[s * SL + q * PL + lh] = ah;
        }
    }
    Row cur;
    load_row(j0 + 1, cur);
    const int cm = gc + L * j0;  // row JC-1 of the first step
    double mujm = MU[cm], mu0jm = MU0[cm], dujm = DU[cm], prjm = PR[cm];
    double s1c = MS1[cm + L], s3c = MS3[cm + L];
    const int nw = blockDim.x >> 6;
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
        double pn[6], ph[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const double* src = q < 3 ? G + q * CS : Q + (q - 3) * CS;
            pn[q] = src[gc + rq];
            ph[q] = src[gh + rq];
        }
        const int c = gc + L * JC;
        const int e = i + L * (JC - 1) + L * M * KC;
        auto R = [&](int q) {
            return CellRing{ring + sM * SL + q * PL + lc, ring + s0 * SL + q * PL + lc, ring + sP * SL + q * PL + lc,
                          ip, im, L};
        };
        const IoV4 vmu{cur.mu[0], cur.mu[1], mujm, cur.mu[2]}, vmu0{cur.mu0[0], cur.mu0[1], mu0jm, cur.mu0[2]};
        const IoV4 vdu{cur.du[0], cur.du[1], dujm, cur.du[2]}, vpr{cur.pr[0], cur.pr[1], prjm, cur.pr[2]};
        const IoVS1 vs1{s1c, cur.s1jp, cur.s1ip};
        const IoVS2 vs2{cur.s2[0], cur.s2[1], cur.s2[2]};
        const IoVS3 vs3{s3c, cur.s3jp, cur.s3kp};
        double hv[3], rv[3];
        io_terms_cell<GRAV>(g, R(0), R(1), R(2), R(3), R(4), R(5), vmu, vmu0, vdu, vs1, vs2, vs3, vpr, JC, cur.e,
                            cur.rs, p, tg, tr, dt, adt, hv, rv);
        const int cg = c + (JC == 1 ? -L : L);
        const double gv[3] = {R(0).c[0], R(1).c[0], R(2).c[0]};
        const double glv[3] = {JC == 1 ? R(0).m[0] : R(0).p[0], JC == 1 ? R(1).m[0] : R(1).p[0],
                               JC == 1 ? R(2).m[0] : R(2).p[0]};
        double v = io_terms_store(p, g, JC, e, NC, hv, rv, E, inc, G0, c, cg, CS, gv, glv);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if ((tid & 63) == 0) red[tid >> 6] = v;
        __syncthreads();
        if (tid < 64) {  // wave 0: the wave sums by a shuffle tree
            double bs = tid < nw ? red[tid] : 0.0;
#pragma unroll
            for (int o = 8; o > 0; o >>= 1) bs += __shfl_xor(bs, o);
            if (tid == 0) part[kg * M + j] = bs;
        }
        // row JC+2 replaces row JC-1 in the ring; this row's values carry over
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            ring[sM * SL + q * PL + lc] = pn[q];
            ring[sM * SL + q * PL + lh] = ph[q];
        }
        mujm = cur.mu[0], mu0jm = cur.mu0[0], dujm = cur.du[0], prjm = cur.pr[0];
        s1c = cur.s1jp, s3c = cur.s3jp;
        cur = nxt;
        __syncthreads();
        const int s_ = sM;
        sM = s0;
        s0 = sP;
        sP = s_;
    }
}

// DPGRNS of the constant-mass-flux drive: the block sums of RHS/DYFI in a
// fixed order, divided by VOLM; for the pressure-gradient drive the
// constant the host passes.
// One block of 1024 threads; each thread keeps four independent partial sums
// (its strided values in turn), so that four loads are in flight per thread.
__global__ void __launch_bounds__(1024) k_io_dpgrns(const double* __restrict__ part, int nb, double volm,
                                                    int flowdrvtp, double cst, double* __restrict__ out) {
    __shared__ double red[1024];
    constexpr int T = 1024;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    int i = threadIdx.x;
    for (; i + 3 * T < nb; i += 4 * T) {
#pragma unroll
        for (int q = 0; q < 4; ++q) s[q] += part[i + q * T];
    }
    if (i < nb) s[0] += part[i];
    if (i + T < nb) s[1] += part[i + T];
    if (i + 2 * T < nb) s[2] += part[i + 2 * T];
    red[threadIdx.x] = (s[0] + s[1]) + (s[2] + s[3]);
    __syncthreads();
    for (int st = T / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = flowdrvtp == 1 ? red[0] / volm : cst;
}

// RHS - DPGRNS along NFLOW, MASSFLUX_CALC_IO, then INTFC_VARS3 + BC_WALL_G_io
// (every ghost row zero, v on the bottom wall face zero); rows 0..M+1.
__global__ void __launch_bounds__(256) k_io_massflux(IoGeo g, double* __restrict__ G, const double* __restrict__ inc,
                                                     const double* __restrict__ dpg, int nflow) {
#pragma clang fp contract(off)
    IO_CELL(g.M + 2, 0);
    const int J = JT;
    const int c = at(IC, J, KC);
    if (J == 0 || J == M + 1) {
        G[c] = 0.0;
        G[CS + c] = 0.0;
        G[2 * CS + c] = 0.0;
        return;
    }
    if (!inc) {   // the RK-CN branch updated G in its y solves: wall values only
        if (J == 1) G[CS + c] = 0.0;
        return;
    }
    const int NC = L * M * N;
    const int e = IC + L * (J - 1) + L * M * KC;
    const double dp = *dpg;
    G[c] = G[c] + (nflow == 1 ? inc[e] - dp : inc[e]);
    G[2 * CS + c] = G[2 * CS + c] + (nflow == 3 ? inc[2 * NC + e] - dp : inc[2 * NC + e]);
    if (J >= 2)
        G[CS + c] = G[CS + c] + (nflow == 2 ? inc[NC + e] - dp : inc[NC + e]);
    else
        G[CS + c] = 0.0;
}

// DIVG_io(NS): RHSLLPHI = (div G + DrhoDtP) / (DT TALP), rows 1..M.  With
// GN (k_io_terms' G + RHS, (L, M, N) layout) the divergence is taken of GN:
// v is zero on row 1 there and on the ghost row M+1 in G.
__global__ void __launch_bounds__(256) k_io_divg(IoGeo g, const double* __restrict__ G,
                                                 const double* __restrict__ GN, const double* __restrict__ drhodtp,
                                                 double coe0, double* __restrict__ f) {
#pragma clang fp contract(off)
    IO_CELL(g.M, 1);
    const int JC = JT;
    const int c = at(IC, JC, KC);
    double DIVX, DIVY, DIVZ;
    if (GN) {
        const int NC = L * M * N, e = IC + L * (JC - 1) + L * M * KC;
        const double vp = JC < M ? GN[NC + e + L] : 0.0;
        DIVX = (GN[IP + L * (JC - 1) + L * M * KC] - GN[e]) * (g.dxi / 1.0);
        DIVY = (vp - GN[NC + e]) * (ldc(g.dyfi, JC - 1) / 1.0);
        DIVZ = (GN[2 * NC + IC + L * (JC - 1) + L * M * KP] - GN[2 * NC + e]) * g.dzi;
    } else {
        DIVX = (G[at(IP, JC, KC)] - G[c]) * (g.dxi / 1.0);
        DIVY = (G[CS + at(IC, JC + 1, KC)] - G[CS + c]) * (ldc(g.dyfi, JC - 1) / 1.0);
        DIVZ = (G[2 * CS + at(IC, JC, KP)] - G[2 * CS + c]) * g.dzi;
    }
    const double DTRHO = drhodtp[c] / 1.0;
    const double DQCAP = DIVX + DIVY + DIVZ + DTRHO;
    f[IC + L * (JC - 1) + L * M * KC] = DQCAP * coe0;
}

// PRCALC_io (explicit) with BC_WALL_PR_io and MASSFLUX_UPDATE_IO, rows 1..M.
// With GN (k_io_terms' G + RHS) the new G is GN - DPGRNS along NFLOW minus
// the pressure correction, and the wall values of BC_WALL_G_io (zero ghost
// rows, zero v on row 1) are written here.
__global__ void __launch_bounds__(256) k_io_update(IoGeo g, double* __restrict__ G, double* __restrict__ PR,
                                                   const double* __restrict__ DPH, double dta,
                                                   const double* __restrict__ f, double be,
                                                   const double* __restrict__ GN, const double* __restrict__ dpg,
                                                   int nflow) {
#pragma clang fp contract(off)
    IO_CELL(g.M, 1);
    const int JC = JT;
    const int c = at(IC, JC, KC);
    if (GN) {
        const int NC = L * M * N, e = IC + L * (JC - 1) + L * M * KC;
        const double dp = *dpg;
        const double pn = PR[c] + DPH[c];
        PR[c] = pn;
        if (JC == 1) PR[at(IC, 0, KC)] = pn;
        if (JC == M) PR[at(IC, M + 1, KC)] = pn;
        const double u = nflow == 1 ? GN[e] - dp : GN[e];
        const double w = nflow == 3 ? GN[2 * NC + e] - dp : GN[2 * NC + e];
        G[c] = u - (DPH[c] - DPH[at(IM, JC, KC)]) * (dta * g.dxi);
        G[2 * CS + c] = w - (DPH[c] - DPH[at(IC, JC, KM)]) * (dta * g.dzi);
        G[CS + c] = JC >= 2 ? GN[NC + e] - (DPH[c] - DPH[at(IC, JC - 1, KC)]) * (dta * ldc(g.dyci, JC - 1) / 1.0)
                            : 0.0;
        if (JC == 1 || JC == M) {
            const int cg = at(IC, JC == 1 ? 0 : M + 1, KC);
            G[cg] = 0.0;
            G[CS + cg] = 0.0;
            G[2 * CS + cg] = 0.0;
        }
        return;
    }
    // PRCALC_io: PR + DPH, less BE * RHSLLPHI in the RK-CN branch
    const double pn = f ? PR[c] + DPH[c] - be * f[IC + L * (JC - 1) + L * M * KC] : PR[c] + DPH[c];
    PR[c] = pn;
    if (JC == 1) PR[at(IC, 0, KC)] = pn;
    if (JC == M) PR[at(IC, M + 1, KC)] = pn;
    const double DFX11 = DPH[c] - DPH[at(IM, JC, KC)];
    G[c] = G[c] - DFX11 * (dta * g.dxi);
    const double DFX33 = DPH[c] - DPH[at(IC, JC, KM)];
    G[2 * CS + c] = G[2 * CS + c] - DFX33 * (dta * g.dzi);
    if (JC >= 2) {
        const double DFX22 = DPH[c] - DPH[at(IC, JC - 1, KC)];
        G[CS + c] = G[CS + c] - DFX22 * (dta * ldc(g.dyci, JC - 1) / 1.0);
    }
}

// NaN-propagating maximum: a NaN operand wins and stays
__device__ __forceinline__ double io_nanmax(double a, double b) { return (b != b || b > a) ? b : a; }

// DIVGCK_io: block maxima of |div G + DrhoDtP| and |div Q| over rows 1..M
__global__ void __launch_bounds__(256) k_io_divgck(IoGeo g, const double* __restrict__ G, const double* __restrict__ Q,
                                                   const double* __restrict__ drhodtp, double* __restrict__ part) {
#pragma clang fp contract(off)
    __shared__ double r0[256], r1[256];
    const int L = g.L, M = g.M, N = g.N;
    const int LP = L * (M + 2), CS = LP * N;
    const long ncell = (long)L * M * N;
    double m0 = 0.0, m1 = 0.0;
    for (long e = blockIdx.x * 256 + threadIdx.x; e < ncell; e += (long)gridDim.x * 256) {
        const int IC = (int)(e % L), JC = (int)((e / L) % M) + 1, KC = (int)(e / ((long)L * M));
        const int IP = IC == L - 1 ? 0 : IC + 1, KP = KC == N - 1 ? 0 : KC + 1;
        const int c = IC + L * JC + LP * KC, ckp = IC + L * JC + LP * KP;
        const double COE2 = ldc(g.dyfi, JC - 1) * 1.0, COE3 = g.dzi * 1.0;
        double d = (G[c - IC + IP] - G[c]) * g.dxi + (G[CS + c + L] - G[CS + c]) * COE2 +
                   (G[2 * CS + ckp] - G[2 * CS + c]) * COE3 + drhodtp[c];
        m0 = io_nanmax(m0, fabs(d));
        d = (Q[c - IC + IP] - Q[c]) * g.dxi + (Q[CS + c + L] - Q[CS + c]) * COE2 + (Q[2 * CS + ckp] - Q[2 * CS + c]) * COE3;
        m1 = io_nanmax(m1, fabs(d));
    }
    r0[threadIdx.x] = m0;
    r1[threadIdx.x] = m1;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            r0[threadIdx.x] = io_nanmax(r0[threadIdx.x], r0[threadIdx.x + st]);
            r1[threadIdx.x] = io_nanmax(r1[threadIdx.x], r1[threadIdx.x + st]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = r0[0];
        part[2 * blockIdx.x + 1] = r1[0];
    }
}

__global__ void __launch_bounds__(256) k_io_max2(const double* __restrict__ part, int nb, double* __restrict__ out) {
    __shared__ double r0[256], r1[256];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < nb; i += 256) {
        a = io_nanmax(a, part[2 * i]);
        b = io_nanmax(b, part[2 * i + 1]);
    }
    r0[threadIdx.x] = a;
    r1[threadIdx.x] = b;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            r0[threadIdx.x] = io_nanmax(r0[threadIdx.x], r0[threadIdx.x + st]);
            r1[threadIdx.x] = io_nanmax(r1[threadIdx.x], r1[threadIdx.x + st]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = r0[0];
        out[1] = r1[0];
    }
}


// ---------------------------------------------------------------------------
// RK-CN branch (visthemflg = 1)
// ---------------------------------------------------------------------------

// VISCOUS_PAR_EXPLT_X/Y/Z_io (src/VISCOUS_X_EXPLICIT_io.f90:116-289,
// src/VISCOUS_Y_EXPLICIT_io.f90:264-423, src/VISCOUS_Z_EXPLICIT_io.f90:289-460)
// with the convection, and the RK-CN branch of RHS_MOM_EXPLICIT_io
// (src/RHS_MOM_EXPLICIT_io.f90:79-93) up to DPGRNS.  inc <- RHS (row 1 of v
// zero), explt0 <- convection + explicit stress part, divu <- the z density
// correction (what the reference leaves in DIVU_io), part <- block sums.
template <int GRAV>
__global__ void __launch_bounds__(256) k_io_terms_cn(IoGeo g, const double* __restrict__ G,
                                                     const double* __restrict__ Q, const double* __restrict__ MU,
                                                     const double* __restrict__ MS, const double* __restrict__ DR,
                                                     const double* __restrict__ PR, const double* __restrict__ RS,
                                                     double* __restrict__ E, double* __restrict__ DUo, IoParams p,
                                                     double tg, double tr, double dt, double adt,
                                                     double* __restrict__ inc, double* __restrict__ part) {
#pragma clang fp contract(off)
    __shared__ double red[256];
    const int L = g.L, M = g.M, N = g.N;
    int it, jt, kt, tile;
    if (!row_tile(L, N, M, &it, &jt, &kt, &tile)) return;
    const int IC = it * 64 + threadIdx.x;
    const int KC = kt * 4 + threadIdx.y;
    const int JC = jt + 1;
    const bool live = IC < L && KC < N;
    double flow = 0.0;
    if (live) {
        const int IP = IC == L - 1 ? 0 : IC + 1, IM = IC == 0 ? L - 1 : IC - 1;
        const int KP = KC == N - 1 ? 0 : KC + 1, KM = KC == 0 ? N - 1 : KC - 1;
        const int LP = L * (M + 2), CS = LP * N, NC = L * M * N;
        auto at = [&](int i, int j, int k) { return i + L * j + LP * k; };
        // row 1 of v is evaluated with a clamped metric index and discarded
        const int JM = JC - 1, JP = JC + 1, JJ = JC, JJP = JC + 1, JJM = JC >= 2 ? JC - 1 : 1;
        const double CV = g.cvisc;
        const double *G1 = G, *G2 = G + CS, *G3 = G + 2 * CS;
        const double *Q1 = Q, *Q2 = Q + CS, *Q3 = Q + 2 * CS;
        const double *MS1 = MS, *MS2 = MS + CS, *MS3 = MS + 2 * CS;
        const double *D1 = DR, *D2 = DR + CS, *D3 = DR + 2 * CS;
        const int c = at(IC, JC, KC);
        const int e = IC + L * (JC - 1) + L * M * KC;
        const double E1 = E[e], E2 = E[NC + e], E3 = E[2 * NC + e];
        double hx, hy, hz;
        io_conv(g, G, Q, IC, IP, IM, JC, KC, KP, KM, hx, hy, hz);
        double IM1, AD1, IM2, AD2, IM3, AD3;
        {   // x
            const double COE1 = -2.0 / 3.0 * g.dxi * CV, COE2 = ldc(g.dyfi, JJ - 1) * CV * 1.0, COE32 = g.dzi * CV * 1.0;
            const int cim = at(IM, JC, KC);
            double DUDX = (Q1[at(IP, JC, KC)] - Q1[c]) * g.dxi;
            double DVDY = (Q2[at(IC, JP, KC)] - Q2[c]) * ldc(g.dyfi, JJ - 1) * 1.0;
            double DWDZ = (Q3[at(IC, JC, KP)] - Q3[c]) * g.dzi * 1.0;
            const double F_EX = MU[c] * (DVDY + DWDZ), F_IM = MU[c] * DUDX;
            DUDX = (G1[at(IP, JC, KC)] * D1[at(IP, JC, KC)] - G1[c] * D1[c]) * g.dxi;
            const double F_AD = MU[c] * DUDX;
            DUDX = (Q1[c] - Q1[cim]) * g.dxi;
            DVDY = (Q2[at(IM, JP, KC)] - Q2[cim]) * ldc(g.dyfi, JJ - 1) * 1.0;
            DWDZ = (Q3[at(IM, JC, KP)] - Q3[cim]) * g.dzi * 1.0;
            const double B_EX = MU[cim] * (DVDY + DWDZ), B_IM = MU[cim] * DUDX;
            DUDX = (G1[c] * D1[c] - G1[cim] * D1[cim]) * g.dxi;
            const double B_AD = MU[cim] * DUDX;
            const double X_EX = (F_EX - B_EX) * COE1, X_IM = (F_IM - B_IM) * COE1 * (-2.0);
            const double X_AD = (F_AD - B_AD) * COE1 * (-2.0);
            double VP = MS1[at(IC, JP, KC)];
            double DVDX = (Q2[at(IC, JP, KC)] - Q2[at(IM, JP, KC)]) * g.dxi;
            double DUDY = (Q1[at(IC, JP, KC)] - Q1[c]) * ldc(g.dyci, JJP - 1) / 1.0;
            const double F2_EX = VP * DVDX, F2_IM = VP * DUDY;
            DUDY = (G1[at(IC, JP, KC)] * D1[at(IC, JP, KC)] - G1[c] * D1[c]) * ldc(g.dyci, JJP - 1) / 1.0;
            const double F2_AD = VP * DUDY;
            double VC = MS1[c];
            DUDY = (Q1[c] - Q1[at(IC, JM, KC)]) * ldc(g.dyci, JJ - 1) / 1.0;
            DVDX = (Q2[c] - Q2[cim]) * g.dxi;
            const double B2_EX = VC * DVDX, B2_IM = VC * DUDY;
            DUDY = (G1[c] * D1[c] - G1[at(IC, JM, KC)] * D1[at(IC, JM, KC)]) * ldc(g.dyci, JJ - 1) / 1.0;
            const double B2_AD = VC * DUDY;
            const double Y_EX = (F2_EX - B2_EX) * COE2, Y_IM = (F2_IM - B2_IM) * COE2, Y_AD = (F2_AD - B2_AD) * COE2;
            VP = MS2[at(IC, JC, KP)];
            double DUDZ = (Q1[at(IC, JC, KP)] - Q1[c]) * g.dzi;
            double DWDX = (Q3[at(IC, JC, KP)] - Q3[at(IM, JC, KP)]) * g.dxi;
            const double F3_EX = VP * DWDX, F3_IM = VP * DUDZ;
            DUDZ = (Q1[at(IC, JC, KP)] * D1[at(IC, JC, KP)] - Q1[c] * D1[c]) * g.dzi;
            const double F3_AD = VP * DUDZ;
            VC = MS2[c];
            DUDZ = (Q1[c] - Q1[at(IC, JC, KM)]) * g.dzi;
            DWDX = (Q3[c] - Q3[cim]) * g.dxi;
            const double B3_EX = VC * DWDX, B3_IM = VC * DUDZ;
            DUDZ = (Q1[c] * D1[c] - Q1[at(IC, JC, KM)] * D1[at(IC, JC, KM)]) * g.dzi;
            const double B3_AD = VC * DUDZ;
            const double Z_EX = (F3_EX - B3_EX) * COE32, Z_IM = (F3_IM - B3_IM) * COE32, Z_AD = (F3_AD - B3_AD) * COE32;
            hx = X_EX + Y_EX + Z_EX + hx;
            IM1 = X_IM + Y_IM + Z_IM;
            AD1 = X_AD + Y_AD + Z_AD;
        }
        {   // y (rows 2..M)
            const double COE1 = g.dxi * CV, COE3 = g.dzi * CV, COE2 = -2.0 / 3.0 * ldc(g.dyci, JJ - 1) * CV;
            const int cjm = at(IC, JM, KC);
            double DUDX = (Q1[at(IP, JC, KC)] - Q1[c]) * g.dxi;
            double DVDY = (Q2[at(IC, JP, KC)] - Q2[c]) * ldc(g.dyfi, JJ - 1);
            double DWDZ = (Q3[at(IC, JC, KP)] - Q3[c]) * g.dzi;
            const double F_EX = MU[c] * (DUDX + DWDZ), F_IM = MU[c] * DVDY;
            DVDY = (G2[at(IC, JP, KC)] * D2[at(IC, JP, KC)] - G2[c] * D2[c]) * ldc(g.dyfi, JJ - 1);
            const double F_AD = MU[c] * DVDY;
            DUDX = (Q1[at(IP, JM, KC)] - Q1[cjm]) * g.dxi;
            DVDY = (Q2[c] - Q2[cjm]) * ldc(g.dyfi, JJM - 1);
            DWDZ = (Q3[at(IC, JM, KP)] - Q3[cjm]) * g.dzi;
            const double B_EX = MU[cjm] * (DUDX + DWDZ), B_IM = MU[cjm] * DVDY;
            DVDY = (G2[c] * D2[c] - G2[cjm] * D2[cjm]) * ldc(g.dyfi, JJM - 1);
            const double B_AD = MU[cjm] * DVDY;
            const double Y_EX = (F_EX - B_EX) * COE2, Y_IM = (F_IM - B_IM) * COE2 * (-2.0);
            const double Y_AD = (F_AD - B_AD) * COE2 * (-2.0);
            double VP = MS1[at(IP, JC, KC)];
            double DVDX = (Q2[at(IP, JC, KC)] - Q2[c]) * g.dxi;
            double DUDY = (Q1[at(IP, JC, KC)] - Q1[at(IP, JM, KC)]) * ldc(g.dyci, JJ - 1);
            const double F1_EX = VP * DUDY, F1_IM = VP * DVDX;
            DVDX = (G2[at(IP, JC, KC)] * D2[at(IP, JC, KC)] - G2[c] * D2[c]) * g.dxi;
            const double F1_AD = VP * DVDX;
            double VC = MS1[c];
            DVDX = (Q2[c] - Q2[at(IM, JC, KC)]) * g.dxi;
            DUDY = (Q1[c] - Q1[cjm]) * ldc(g.dyci, JJ - 1);
            const double B1_EX = VC * DUDY, B1_IM = VC * DVDX;
            DVDX = (G2[c] * D2[c] - G2[at(IM, JC, KC)] * D2[at(IM, JC, KC)]) * g.dxi;
            const double B1_AD = VC * DVDX;
            const double X_EX = (F1_EX - B1_EX) * COE1, X_IM = (F1_IM - B1_IM) * COE1, X_AD = (F1_AD - B1_AD) * COE1;
            VP = MS3[at(IC, JC, KP)];
            double DVDZ = (Q2[at(IC, JC, KP)] - Q2[c]) * g.dzi;
            double DWDY = (Q3[at(IC, JC, KP)] - Q3[at(IC, JM, KP)]) * ldc(g.dyci, JJ - 1);
            const double F3_EX = VP * DWDY, F3_IM = VP * DVDZ;
            DVDZ = (G2[at(IC, JC, KP)] * D2[at(IC, JC, KP)] - G2[c] * D2[c]) * g.dzi;
            const double F3_AD = VP * DVDZ;
            VC = MS3[c];
            DVDZ = (Q2[c] - Q2[at(IC, JC, KM)]) * g.dzi;
            DWDY = (Q3[c] - Q3[cjm]) * ldc(g.dyci, JJ - 1);
            const double B3_EX = VC * DWDY, B3_IM = VC * DVDZ;
            DVDZ = (G2[c] * D2[c] - G2[at(IC, JC, KM)] * D2[at(IC, JC, KM)]) * g.dzi;
            const double B3_AD = VC * DVDZ;
            const double Z_EX = (F3_EX - B3_EX) * COE3, Z_IM = (F3_IM - B3_IM) * COE3, Z_AD = (F3_AD - B3_AD) * COE3;
            hy = X_EX + Y_EX + Z_EX + hy;
            IM2 = X_IM + Y_IM + Z_IM;
            AD2 = X_AD + Y_AD + Z_AD;
        }
        {   // z
            const double COE1 = g.dxi * CV, COE3 = -2.0 / 3.0 * g.dzi * CV, COE2 = ldc(g.dyfi, JJ - 1) * CV;
            const int ckm = at(IC, JC, KM);
            double DUDX = (Q1[at(IP, JC, KC)] - Q1[c]) * g.dxi;
            double DVDY = (Q2[at(IC, JP, KC)] - Q2[c]) * ldc(g.dyfi, JJ - 1);
            double DWDZ = (Q3[at(IC, JC, KP)] - Q3[c]) * g.dzi;
            const double F_EX = MU[c] * (DUDX + DVDY), F_IM = MU[c] * DWDZ;
            DWDZ = (G3[at(IC, JC, KP)] * D3[at(IC, JC, KP)] - G3[c] * D3[c]) * g.dzi;
            const double F_AD = MU[c] * DWDZ;
            DUDX = (Q1[at(IP, JC, KM)] - Q1[ckm]) * g.dxi;
            DVDY = (Q2[at(IC, JP, KM)] - Q2[ckm]) * ldc(g.dyfi, JJ - 1);
            DWDZ = (Q3[c] - Q3[ckm]) * g.dzi;
            const double B_EX = MU[ckm] * (DUDX + DVDY), B_IM = MU[ckm] * DWDZ;
            DWDZ = (G3[c] * D3[c] - G3[ckm] * D3[ckm]) * g.dzi;
            const double B_AD = MU[ckm] * DWDZ;
            const double Z_EX = (F_EX - B_EX) * COE3, Z_IM = (F_IM - B_IM) * COE3 * (-2.0);
            const double Z_AD = (F_AD - B_AD) * COE3 * (-2.0);
            double VP = MS2[at(IP, JC, KC)];
            double DWDX = (Q3[at(IP, JC, KC)] - Q3[c]) * g.dxi;
            double DUDZ = (Q1[at(IP, JC, KC)] - Q1[at(IP, JC, KM)]) * g.dzi;
            const double F1_EX = VP * DUDZ, F1_IM = VP * DWDX;
            DWDX = (G3[at(IP, JC, KC)] * D3[at(IP, JC, KC)] - G3[c] * D3[c]) * g.dxi;
            const double F1_AD = VP * DWDX;
            double VC = MS2[c];
            DWDX = (Q3[c] - Q3[at(IM, JC, KC)]) * g.dxi;
            DUDZ = (Q1[c] - Q1[ckm]) * g.dzi;
            const double B1_EX = VC * DUDZ, B1_IM = VC * DWDX;
            DWDX = (G3[c] * D3[c] - G3[at(IM, JC, KC)] * D3[at(IM, JC, KC)]) * g.dxi;
            const double B1_AD = VC * DWDX;
            const double X_EX = (F1_EX - B1_EX) * COE1, X_IM = (F1_IM - B1_IM) * COE1, X_AD = (F1_AD - B1_AD) * COE1;
            VP = MS3[at(IC, JP, KC)];
            double DWDY = (Q3[at(IC, JP, KC)] - Q3[c]) * ldc(g.dyci, JJP - 1);
            double DVDZ = (Q2[at(IC, JP, KC)] - Q2[at(IC, JP, KM)]) * g.dzi;
            const double F2_EX = VP * DVDZ, F2_IM = VP * DWDY;
            DWDY = (G3[at(IC, JP, KC)] * D3[at(IC, JP, KC)] - G3[c] * D3[c]) * ldc(g.dyci, JJP - 1);
            const double F2_AD = VP * DWDY;
            VC = MS3[c];
            DWDY = (Q3[c] - Q3[at(IC, JM, KC)]) * ldc(g.dyci, JJ - 1);
            DVDZ = (Q2[c] - Q2[ckm]) * g.dzi;
            const double B2_EX = VC * DVDZ, B2_IM = VC * DWDY;
            DWDY = (G3[c] * D3[c] - G3[at(IC, JM, KC)] * D3[at(IC, JM, KC)]) * ldc(g.dyci, JJ - 1);
            const double B2_AD = VC * DWDY;
            const double Y_EX = (F2_EX - B2_EX) * COE2, Y_IM = (F2_IM - B2_IM) * COE2, Y_AD = (F2_AD - B2_AD) * COE2;
            hz = X_EX + Y_EX + Z_EX + hz;
            IM3 = X_IM + Y_IM + Z_IM;
            AD3 = X_AD + Y_AD + Z_AD;
        }
        double r1 = (tg * hx + tr * E1) * dt;
        r1 = r1 + adt * IM1 + 0.5 * adt * AD1;
        r1 = r1 - (PR[c] - PR[at(IM, JC, KC)]) * (adt * g.dxi);
        double r2 = (tg * hy + tr * E2) * dt;
        r2 = r2 + adt * IM2 + 0.5 * adt * AD2;
        r2 = r2 - (PR[c] - PR[at(IC, JM, KC)]) * (adt * ldc(g.dyci, JC - 1) / 1.0);
        double r3 = (tg * hz + tr * E3) * dt;
        r3 = r3 + adt * IM3 + 0.5 * adt * AD3;
        r3 = r3 - (PR[c] - PR[at(IC, JC, KM)]) * (adt * g.dzi);
        if (GRAV == 1) r1 = r1 + RS[c] * (adt * p.f_a / 1.0);
        if (GRAV == 2) r2 = r2 + RS[CS + c] * (adt * p.f_a / 1.0);
        if (GRAV == 3) r3 = r3 + RS[2 * CS + c] * (adt * p.f_a / 1.0);
        E[e] = hx;
        E[2 * NC + e] = hz;
        inc[e] = r1;
        inc[2 * NC + e] = r3;
        inc[NC + e] = JC >= 2 ? r2 : 0.0;
        if (JC >= 2) E[NC + e] = hy;
        DUo[c] = AD3;
        const double rf = p.nflow == 1 ? r1 : (p.nflow == 2 ? (JC >= 2 ? r2 : 0.0) : r3);
        flow = rf / ldc(g.dyfi, JC - 1) / 1.0;
    }
    const double bs = io_block_sum<256>(flow, red);
    if (threadIdx.x == 0 && threadIdx.y == 0) part[tile] = bs;
}

struct IoCnCoef {
    double c1, c2;             // FACOE * 4/3 * (DXQI | DZQI | 1), FACOE * (DXQI | DZQI | 1)
    const double *amvr, *apvr; // AMVR, APVR (NND2, 3)
};

// MOMFA1_X_io / MOMFA2_Z_io / MOMFA3_Y_io (src/MOMFA_io.f90:51-377) for one
// component IDR: the coefficients of each line are formed on load from the
// properties, then the segmented solver; DIR 0 = x (cyclic, lines (j, k)),
// 2 = z (cyclic, lines (i, j)), 1 = y (non-cyclic, lines (i, k), zero
// boundary values).  The x pass subtracts DPGRNS from the NFLOW component
// (rows it applies to); the y pass adds the result to G from row N2I.
template <int MS, int DIR>
__global__ void __launch_bounds__(1024) k_io_cn(IoGeo g, int idr, int sub_dpg, const double* __restrict__ MU,
                                               const double* __restrict__ MSTG, const double* __restrict__ RS,
                                               double* __restrict__ R, double* __restrict__ Gc,
                                               const double* __restrict__ dpg, IoCnCoef cf, int LPB, int P) {
#pragma clang fp contract(off)
    extern __shared__ unsigned char smem[];
    const int L = g.L, M = g.M, N = g.N;
    const int LP = L * (M + 2), CS = LP * N;
    const int n = DIR == 0 ? L : (DIR == 2 ? N : M);
    const int nlines = DIR == 0 ? M * N : (DIR == 2 ? L * M : L * N);
    const int ll = threadIdx.x % LPB, seg = threadIdx.x / LPB;
    const int line = blockIdx.x * LPB + ll;
    const bool valid = line < nlines;
    int li = 0, lj = 1, lk = 0;   // the line's fixed indices (j is the ghost-aware row)
    if (valid) {
        if (DIR == 0) { lj = line % M + 1; lk = line / M; }
        else if (DIR == 2) { li = line % L; lj = line / L + 1; }
        else { li = line % L; lk = line / L; }
    }
    const double* RSc = RS + (size_t)(idr - 1) * CS;
    const double dp = sub_dpg ? *dpg : 0.0;
    // wall-row coefficients of u, w in y that TDMA_COEF sets aside
    // (src/TDMA_COEF.f90:86-96)
    const double amvr1 = 4.0 / (2.0 / ldc(g.dyci, 1) + 1.0 / ldc(g.dyfi, 0)) * (2.0 * ldc(g.dyfi, 0));
    const double apvrn = 4.0 / (2.0 / ldc(g.dyci, M - 1) + 1.0 / ldc(g.dyfi, M - 1)) * (2.0 * ldc(g.dyfi, M - 1));
    double a[MS], b[MS], c[MS], d[MS];
    // x and z: branch-free over the rows (rows past n form row n - 1's
    // coefficients, invalid lines those of line 0, and become identity rows
    // after), so the loads of all rows are issued before the first wait; y
    // skips them (its wall-row branches split the loads anyway, and the
    // clamped form measured slower there)
#pragma unroll
    for (int l = 0; l < MS; ++l) {
        if (DIR == 1 && !(valid && seg * MS + l < n)) {
            a[l] = 0.0; b[l] = 1.0; c[l] = 0.0; d[l] = 0.0;
            continue;
        }
        const int m = min(seg * MS + l, n - 1);
        const int IC = DIR == 0 ? m : li, JC = DIR == 1 ? m + 1 : lj, KC = DIR == 2 ? m : lk;
        const int cc = IC + L * JC + LP * KC;
        const int e = IC + L * (JC - 1) + L * M * KC;
        double AMU, APU, am, ap, ac;
        if (DIR == 0) {
            const int IM = IC == 0 ? L - 1 : IC - 1, IP = IC == L - 1 ? 0 : IC + 1;
            const int cm = cc - IC + IM, cp = cc - IC + IP;
            // u: MU at (IM, IC), coefficient c1; v, w: MU_STG(., 1 | 2) at (IC, IP), c2
            const bool u = idr == 1;
            const double* X = u ? MU : MSTG + (size_t)(idr - 2) * CS;
            const double k = u ? cf.c1 : cf.c2;
            AMU = -k * X[u ? cm : cc];
            APU = -k * X[u ? cc : cp];
            am = AMU / RSc[cm];
            ap = APU / RSc[cp];
            ac = 1.0 - (AMU + APU) / RSc[cc];
        } else if (DIR == 2) {
            const int KM = KC == 0 ? N - 1 : KC - 1, KP = KC == N - 1 ? 0 : KC + 1;
            const int cm = cc + LP * (KM - KC), cp = cc + LP * (KP - KC);
            // w: MU at (KM, KC), c1; u, v: MU_STG(., 2 | 3) at (KC, KP), c2
            const bool w = idr == 3;
            const double* X = w ? MU : MSTG + (size_t)(idr == 1 ? 1 : 2) * CS;
            const double k = w ? cf.c1 : cf.c2;
            AMU = -k * X[w ? cm : cc] * 1.0;
            APU = -k * X[w ? cc : cp] * 1.0;
            am = AMU / RSc[cm];
            ap = APU / RSc[cp];
            ac = 1.0 - (AMU + APU) / RSc[cc];
        } else {
            const int cm = cc - L, cp = cc + L;   // rows 0 and M + 1 are the ghost rows
            const double amv = cf.amvr[(idr - 1) * (M + 1) + JC - 1], apv = cf.apvr[(idr - 1) * (M + 1) + JC - 1];
            const double* FM = MSTG + (size_t)(idr - 1) * CS;
            if (idr == 1 || idr == 3) {
                if (JC == 1) {
                    const double MUQT = (MU[cc] + MU[cm]) * 0.5;
                    AMU = -cf.c2 * amvr1 * MUQT;
                    APU = -cf.c2 * apv * FM[cp];
                    am = 0.0;
                    ap = APU / RSc[cp];
                } else if (JC == M) {
                    const double MUQT = (MU[cc] + MU[cp]) * 0.5;
                    AMU = -cf.c2 * amv * FM[cc];
                    APU = -cf.c2 * apvrn * MUQT;
                    am = AMU / RSc[cm];
                    ap = 0.0;
                } else {
                    AMU = -cf.c2 * amv * FM[cc];
                    APU = -cf.c2 * apv * FM[cp];
                    am = AMU / RSc[cm];
                    ap = APU / RSc[cp];
                }
                ac = 1.0 - (AMU + APU) / RSc[cc];
            } else {
                AMU = -cf.c1 * amv * MU[cm];
                APU = -cf.c1 * apv * MU[cc];
                am = AMU / RSc[cm];
                ap = APU / RSc[JC == 1 ? cp : cc];
                ac = 1.0 - (AMU + APU) / RSc[cc];
            }
        }
        a[l] = am;
        b[l] = ac;
        c[l] = ap;
        d[l] = (DIR == 0 && (idr != 2 || JC >= 2)) ? R[e] - dp : R[e];
    }
#pragma unroll
    for (int l = 0; l < MS; ++l) {
        if (!(valid && seg * MS + l < n)) {
            a[l] = 0.0; b[l] = 1.0; c[l] = 0.0; d[l] = 0.0;
        }
    }
    if (DIR == 1) {   // zero boundary values: the end rows lose their outer coefficient
        if (seg == 0) a[0] = 0.0;
        const int jl = n - 1;
        if (jl / MS == seg) {
#pragma unroll
            for (int q = 0; q < MS; ++q)
                if (q == jl % MS) c[q] = 0.0;
        }
    }
    seg_solve<MS, double>(a, b, c, d, LPB, P, ll, seg, valid, DIR != 1, smem);
    if (!valid) return;
#pragma unroll
    for (int l = 0; l < MS; ++l) {
        const int m = seg * MS + l;
        if (m >= n) continue;
        const int IC = DIR == 0 ? m : li, JC = DIR == 1 ? m + 1 : lj, KC = DIR == 2 ? m : lk;
        const int e = IC + L * (JC - 1) + L * M * KC;
        if (DIR == 1) {
            if (JC >= (idr == 2 ? 2 : 1)) {
                const int cc = IC + L * JC + LP * KC;
                Gc[cc] = d[l] + Gc[cc];
            }
        } else {
            R[e] = d[l];
        }
    }
}

// MOMFA1_X_io through LDS: the x lines are rows of the arrays, so the block
// first copies its LPB rows of RHS, of the viscosity the component uses (MU
// for u, MU_STG for v and w) and of RHO_STG with coalesced loads into LDS
// (row stride L + 1: the lanes of a wave, LPB lines x consecutive segments,
// then fall on distinct banks), forms the coefficients from there exactly as
// k_io_cn<MS, 0> does, solves, and writes the rows back through LDS.
inline size_t io_cn_x_smem(int L, int LPB, int P) {
    return std::max(sizeof(double) * 3 * (size_t)LPB * (L + 1), seg_smem_bytes<double>(LPB, P));
}

template <int MS>
__global__ void __launch_bounds__(1024) k_io_cn_x(IoGeo g, int idr, int sub_dpg, const double* __restrict__ MU,
                                                 const double* __restrict__ MSTG, const double* __restrict__ RS,
                                                 double* __restrict__ R, const double* __restrict__ dpg, IoCnCoef cf,
                                                 int LPB, int P) {
#pragma clang fp contract(off)
    extern __shared__ unsigned char smem[];
    const int L = g.L, M = g.M, N = g.N;
    const int LP = L * (M + 2), CS = LP * N, RL = L + 1;
    const int nlines = M * N, line0 = blockIdx.x * LPB;
    double* sR = reinterpret_cast<double*>(smem);
    double* sX = sR + LPB * RL;
    double* sS = sX + LPB * RL;
    const double* X = idr == 1 ? MU : MSTG + (size_t)(idr - 2) * CS;  // MU for u, MU_STG(., 1 | 2) for v | w
    const double* RSc = RS + (size_t)(idr - 1) * CS;
    for (int e = threadIdx.x; e < LPB * L; e += blockDim.x) {
        const int ll = e / L, i = e - ll * L, line = line0 + ll;
        if (line < nlines) {
            const int lj = line % M + 1, lk = line / M;
            const int cc = i + L * lj + LP * lk, ee = i + L * (lj - 1) + L * M * lk;
            sR[ll * RL + i] = R[ee];
            sX[ll * RL + i] = X[cc];
            sS[ll * RL + i] = RSc[cc];
        }
    }
    __syncthreads();
    const int ll = threadIdx.x % LPB, seg = threadIdx.x / LPB;
    const int line = line0 + ll;
    const bool valid = line < nlines;
    const int lj = valid ? line % M + 1 : 1;
    const double dp = sub_dpg ? *dpg : 0.0;
    const double* xr = sX + ll * RL;
    const double* sr = sS + ll * RL;
    double a[MS], b[MS], c[MS], d[MS];
#pragma unroll
    for (int l = 0; l < MS; ++l) {
        const int m = seg * MS + l;
        a[l] = 0.0; b[l] = 1.0; c[l] = 0.0; d[l] = 0.0;
        if (!valid || m >= L) continue;
        const int IM = m == 0 ? L - 1 : m - 1, IP = m == L - 1 ? 0 : m + 1;
        double AMU, APU;
        if (idr == 1) {
            AMU = -cf.c1 * xr[IM];
            APU = -cf.c1 * xr[m];
        } else {
            AMU = -cf.c2 * xr[m];
            APU = -cf.c2 * xr[IP];
        }
        a[l] = AMU / sr[IM];
        c[l] = APU / sr[IP];
        b[l] = 1.0 - (AMU + APU) / sr[m];
        const double rv = sR[ll * RL + m];
        d[l] = (idr != 2 || lj >= 2) ? rv - dp : rv;
    }
    __syncthreads();   // the staged rows are read: seg_solve reuses the LDS
    seg_solve<MS, double>(a, b, c, d, LPB, P, ll, seg, valid, true, smem);
    __syncthreads();
    if (valid) {
#pragma unroll
        for (int l = 0; l < MS; ++l) {
            const int m = seg * MS + l;
            if (m < L) sR[ll * RL + m] = d[l];
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < LPB * L; e += blockDim.x) {
        const int l2 = e / L, i = e - l2 * L, ln = line0 + l2;
        if (ln < nlines) R[i + L * (ln % M) + L * M * (ln / M)] = sR[l2 * RL + i];
    }
}

static IoGeo io_geo(const chs_handle* h) {
    return IoGeo{h->L, h->M, h->N, h->dxi, h->dzi, h->cvisc, h->volm, h->dyfi, h->dyci, h->wff, h->wfb};
}

static int io_arrays(const chs_thermal_fields* tf, const chs_io_fields* f, IoArrays* a, const char* who) {
    CHS_CHECK(tf && f, CHS_E_ARG, std::string(who) + ": NULL argument");
    const double* req[] = {f->q, f->g, f->g0, f->pr, f->drhodtp, f->divu, f->rho_stg, f->mu_stg, f->explt0,
                           tf->density, tf->density0, tf->viscosity, tf->viscosity0};
    for (const double* r : req) CHS_CHECK(r != nullptr, CHS_E_ARG, std::string(who) + ": NULL field");
    *a = IoArrays{f->q, f->g, f->g0, f->pr, f->drhodtp, f->divu, f->rho_stg, f->mu_stg, f->explt0,
                  tf->density, tf->density0, tf->viscosity, tf->viscosity0};
    return CHS_OK;
}

static int io_check(const chs_handle* h, const chs_io_params* prm, int ns, const char* who) {
    CHS_CHECK(h && prm, CHS_E_ARG, std::string(who) + ": NULL argument");
    CHS_CHECK(ns >= 1 && ns <= 3, CHS_E_ARG, std::string(who) + ": NS must be 1..3");
    CHS_CHECK(h->M >= 2, CHS_E_ARG, std::string(who) + ": NCL2 >= 2 needed");
    CHS_CHECK(h->icase == 1, CHS_E_ARG, std::string(who) + ": the channel (ICASE = 1) only");
    CHS_CHECK(prm->thermal == 0 || prm->thermal == 1, CHS_E_ARG, std::string(who) + ": thermal must be 0 or 1");
    CHS_CHECK(prm->nflow >= 1 && prm->nflow <= 3, CHS_E_ARG, std::string(who) + ": NFLOW must be 1..3");
    CHS_CHECK(prm->flowdrvtp == 1 || prm->flowdrvtp == 2, CHS_E_ARG, std::string(who) + ": FLOWDRVTP must be 1 or 2");
    CHS_CHECK(prm->gravdir >= 0 && prm->gravdir <= 3, CHS_E_ARG, std::string(who) + ": GRAVDIR must be 0..3");
    CHS_CHECK(prm->visimplicit == 0 || prm->visimplicit == 1, CHS_E_ARG,
              std::string(who) + ": visimplicit must be 0 or 1");
    return CHS_OK;
}

static long io_tiles(const chs_handle* h, int rows) { return (long)((h->L + 63) / 64) * ((h->N + 3) / 4) * rows; }
static dim3 io_grid(const chs_handle* h, int rows) { return dim3((unsigned)(8 * ((io_tiles(h, rows) + 7) / 8))); }

int io_prepare(chs_handle* h, const chs_thermal_fields* tf, const chs_io_fields* f, const chs_io_params* prm,
               int ns) {
    CHS_TRY(io_check(h, prm, ns, "chs_io_prepare"));
    IoArrays a;
    CHS_TRY(io_arrays(tf, f, &a, "chs_io_prepare"));
    const hipStream_t s = h->stream;
    hipEvent_t ev;
    h->timer.begin("io_prepare", s, &ev);
    double* drhoi = nullptr;
    if (prm->visimplicit && prm->thermal) {
        CHS_CHECK(f->drhoi_stg != nullptr, CHS_E_ARG, "chs_io_prepare: drhoi_stg is needed with visimplicit = 1");
        drhoi = f->drhoi_stg;
    }
    const double ca = TGAM_IO[ns] + TROH_IO[ns], cb = TROH_IO[ns], cd = TGAM_IO[ns] - 2.0;
    const IoGeo g = io_geo(h);
    const long ncol = (long)((h->L + 63) / 64) * ((h->N + 3) / 4);
    // the marching form takes three launches: on grids of fewer than 4096
    // row tiles (64^3: 1056) one row-tile launch is faster (13.3 against
    // 20.4 us); at 128^3 (8320) the march wins (98.0 against 101.0 us)
    if (prm->thermal && h->M >= 3 && h->cn_mode != 2 && (h->cn_mode == 3 || ncol * (h->M + 2) >= 4096)) {
        // rows 2..M marching (about 8 blocks per CU), rows 0, 1 and M+1 in row tiles
        int nyc = (int)std::max(1L, std::min((long)(h->M - 1) / 4, (2048 + ncol - 1) / ncol));
        const int jc = (h->M - 1 + nyc - 1) / nyc;
        nyc = (h->M - 1 + jc - 1) / jc;
        CHS_LAUNCH(k_io_prepare_m, dim3((unsigned)(8 * ((ncol * nyc + 7) / 8))), dim3(64, 4), 0, s, g, a.rho,
                   a.rho0, a.mu, a.mu0, a.g, a.g0, a.drhodtp, a.rs, a.ms, a.q, drhoi, ca, cb, cd, nyc, jc);
        CHS_LAUNCH(k_io_prepare, io_grid(h, 2), dim3(64, 4), 0, s, g, a.rho, a.rho0, a.mu, a.mu0, a.g, a.g0,
                   a.drhodtp, a.rs, a.ms, a.q, drhoi, prm->thermal, ca, cb, cd, 2, 0);
        CHS_LAUNCH(k_io_prepare, io_grid(h, 1), dim3(64, 4), 0, s, g, a.rho, a.rho0, a.mu, a.mu0, a.g, a.g0,
                   a.drhodtp, a.rs, a.ms, a.q, drhoi, prm->thermal, ca, cb, cd, 1, h->M + 1);
    } else {
        CHS_LAUNCH(k_io_prepare, io_grid(h, h->M + 2), dim3(64, 4), 0, s, g, a.rho, a.rho0, a.mu, a.mu0, a.g,
                   a.g0, a.drhodtp, a.rs, a.ms, a.q, drhoi, prm->thermal, ca, cb, cd, h->M + 2, 0);
    }
    h->timer.end(ev, "io_prepare", s);
    return CHS_OK;
}

// launch one factorised line solve of the RK-CN branch
template <int DIR>
static int io_cn_launch(chs_handle* h, const IoGeo& g, int idr, int sub_dpg, const IoArrays& a, double* R,
                        double* Gc, const IoCnCoef& cf) {
    const int n = DIR == 0 ? h->L : (DIR == 2 ? h->N : h->M);
    const int nlines = DIR == 0 ? h->M * h->N : (DIR == 2 ? h->L * h->M : h->L * h->N);
    int MS = 0, P = 0, LPB = 0;
    // x lines are contiguous in memory: few lines per block (8) put a wave's
    // lanes on consecutive segments of the same lines, so its loads coalesce
    CHS_CHECK(tdma_seg_config(n, DIR != 1, &MS, &P, &LPB, 1024, DIR == 0), CHS_E_ARG,
              "chs_solverrk3_mom_io: no segmented solver layout for a line length of " + std::to_string(n));
    const dim3 grid((nlines + LPB - 1) / LPB), block(LPB * P);
    const size_t sm = seg_smem_bytes<double>(LPB, P);
    const hipStream_t s = h->stream;
    const char* nm = DIR == 0 ? "io_cn_x" : (DIR == 2 ? "io_cn_z" : "io_cn_y");
    hipEvent_t ev;
    h->timer.begin(nm, s, &ev);
    const size_t smx = io_cn_x_smem(h->L, LPB, P);
    if (DIR == 0 && h->cn_mode != 2 && smx <= 65536) {
#define IO_CN_X(MSV) \
    CHS_LAUNCH((k_io_cn_x<MSV>), grid, block, smx, s, g, idr, sub_dpg, a.mu, a.ms, a.rs, R, h->scalars, cf, LPB, P)
        switch (MS) {
            case 16: IO_CN_X(16); break;
            case 12: IO_CN_X(12); break;
            case 4: IO_CN_X(4); break;
            default: IO_CN_X(8); break;
        }
#undef IO_CN_X
    } else {
#define IO_CN(MSV) \
    CHS_LAUNCH((k_io_cn<MSV, DIR>), grid, block, sm, s, g, idr, sub_dpg, a.mu, a.ms, a.rs, R, Gc, h->scalars, cf, LPB, P)
        switch (MS) {
            case 16: IO_CN(16); break;
            case 12: IO_CN(12); break;
            case 4: IO_CN(4); break;
            default: IO_CN(8); break;
        }
#undef IO_CN
    }
    h->timer.end(ev, nm, s);
    return CHS_OK;
}

// SOLVERRK3_MOM_IO with visthemflg = 1 (src/SOLVERRK3_MOM_IO.f90:76-90): the
// split viscous terms and RHS for all three components (they read only G of
// their own component and the velocities of the substep's start, so one pass
// computes what the reference computes component by component), DPGRNS, then
// MOMFA_io per component, the wall values of G, the projection and PRCALC_io
// with its BE * RHSLLPHI term.
static int io_substep_cn(chs_handle* h, const chs_io_fields* f, const IoArrays& a, const IoParams& p, int ns,
                         double dt, double talp) {
    CHS_CHECK(f->drhoi_stg != nullptr, CHS_E_ARG, "chs_solverrk3_mom_io: drhoi_stg is needed with visimplicit = 1");
    const hipStream_t s = h->stream;
    const IoGeo g = io_geo(h);
    const dim3 gi = io_grid(h, h->M);
    const long nb = io_tiles(h, h->M);
    const size_t ncell = (size_t)h->L * h->M * h->N;
    hipEvent_t ev;
    h->timer.begin("io_terms_cn", s, &ev);
#define IO_TERMS_CN(GR)                                                                                          \
    CHS_LAUNCH(k_io_terms_cn<GR>, gi, dim3(64, 4), 0, s, g, a.g, a.q, a.mu, a.ms, f->drhoi_stg, a.pr, a.rs,      \
               a.explt0, a.divu, p, TGAM_IO[ns], TROH_IO[ns], dt, talp * dt, h->rhs, h->partials)
    switch (p.gravdir) {
        case 1: IO_TERMS_CN(1); break;
        case 2: IO_TERMS_CN(2); break;
        case 3: IO_TERMS_CN(3); break;
        default: IO_TERMS_CN(0); break;
    }
#undef IO_TERMS_CN
    h->timer.end(ev, "io_terms_cn", s);
    CHS_LAUNCH(k_io_dpgrns, dim3(1), dim3(1024), 0, s, h->partials, (int)nb, h->volm, p.flowdrvtp,
               -0.5 * p.cfgv * (talp * dt), h->scalars);
    // MOMFA_io: FACOE_io = alpha dt / (2 Re) (src/MOMFA_io.f90:29)
    const double FACOE = 0.5 * talp * dt * h->cvisc;
    const IoCnCoef cx{FACOE * 4.0 / 3.0 * h->dxqi, FACOE * h->dxqi, h->amvr, h->apvr};
    const IoCnCoef cz{FACOE * 4.0 / 3.0 * h->dzqi, FACOE * h->dzqi, h->amvr, h->apvr};
    const IoCnCoef cy{FACOE * 4.0 / 3.0, FACOE, h->amvr, h->apvr};
    const size_t CS = (size_t)h->L * (h->M + 2) * h->N;
    for (int idr = 1; idr <= 3; ++idr) {
        double* R = h->rhs + (size_t)(idr - 1) * ncell;
        double* Gc = a.g + (size_t)(idr - 1) * CS;
        CHS_TRY(io_cn_launch<0>(h, g, idr, p.nflow == idr, a, R, Gc, cx));
        CHS_TRY(io_cn_launch<2>(h, g, idr, 0, a, R, Gc, cz));
        CHS_TRY(io_cn_launch<1>(h, g, idr, 0, a, R, Gc, cy));
    }
    CHS_LAUNCH(k_io_massflux, io_grid(h, h->M + 2), dim3(64, 4), 0, s, g, a.g, nullptr, h->scalars, p.nflow);
    h->timer.begin("io_divg", s, &ev);
    CHS_LAUNCH(k_io_divg, gi, dim3(64, 4), 0, s, g, a.g, nullptr, a.drhodtp, 1.0 / dt / talp, h->rhsllphi);
    h->timer.end(ev, "io_divg", s);
    CHS_TRY(poisson_solve(h, h->rhsllphi, h->dph, 1, 0.5e-10 / (talp * dt)));
    h->timer.begin("io_update", s, &ev);
    CHS_LAUNCH(k_io_update, gi, dim3(64, 4), 0, s, g, a.g, a.pr, h->dph, dt * talp, h->rhsllphi,
               0.5 * talp * dt * h->cvisc, nullptr, nullptr, 0);
    h->timer.end(ev, "io_update", s);
    return CHS_OK;
}

int io_substep(chs_handle* h, const chs_thermal_fields* tf, const chs_io_fields* f, const chs_io_params* prm,
               int ns, double dt) {
    CHS_TRY(io_check(h, prm, ns, "chs_solverrk3_mom_io"));
    IoArrays a;
    CHS_TRY(io_arrays(tf, f, &a, "chs_solverrk3_mom_io"));
    const hipStream_t s = h->stream;
    const IoGeo g = io_geo(h);
    const dim3 gi = io_grid(h, h->M);
    const long nb = io_tiles(h, h->M);
    CHS_CHECK(nb <= 2L * h->npart, CHS_E_ARG, "chs_solverrk3_mom_io: grid too large for the block-sum buffer");
    const size_t ngh = (size_t)h->L * (h->M + 2) * h->N;
    const double talp = TGAM_IO[ns] + TROH_IO[ns];
    hipEvent_t ev;
    IoParams p{prm->thermal, prm->nflow, prm->flowdrvtp, prm->gravdir, prm->cfgv, prm->f_a};
    // G0 = G (src/SOLVERRK3_MOM_IO.f90:15): a copy in the RK-CN branch and for
    // NFLOW = 2; otherwise k_io_terms writes G0 and G + RHS ("fold"), the mass
    // flux pass disappears and k_io_update takes DPGRNS off
    const bool fold = !prm->visimplicit && prm->nflow != 2;
    if (!fold) CHS_HIP(hipMemcpyAsync(f->g0, f->g, 3 * ngh * sizeof(double), hipMemcpyDeviceToDevice, s));
    if (prm->visimplicit) return io_substep_cn(h, f, a, p, ns, dt, talp);   // no DIVG_U_io in that branch
    h->timer.begin("io_divu", s, &ev);
    CHS_LAUNCH(k_io_divu, io_grid(h, h->M + 2), dim3(64, 4), 0, s, g, f->q, f->divu);
    h->timer.end(ev, "io_divu", s);
    // the marching form needs L a power of two with 2L threads in whole
    // waves (at most 512) and its LDS ring on this device; chs_set_cn_solver
    // mode 2 selects the row-tile form
    const int L = h->L;
    const bool march = h->cn_mode != 2 && (L & (L - 1)) == 0 && L >= 32 && 2 * L <= 512 && h->N % IOKR == 0 &&
                       io_terms_m_smem(L) <= (size_t)h->lds_max && (long)(h->N / IOKR) * h->M <= 2L * h->npart &&
                       (h->cn_mode == 3 ||
                        march_chunk_rows(h->M, h->N / IOKR, io_terms_m_smem(L), h->lds_max, true) >= MARCH_MIN);
    long nbt = nb;
    h->timer.begin("io_terms", s, &ev);
    if (march) {
        const int nkg = h->N / IOKR;
        const size_t sm = io_terms_m_smem(L);
        // as many blocks per CU as the ring lets it hold: at 128^3 two
        // (heated substep 552.9 -> 535.4 us); one at L = 256
        const int jc = march_chunk_rows(h->M, nkg, sm, h->lds_max, true);
        const int nyc = (h->M + jc - 1) / jc;
        nbt = (long)nkg * h->M;
        const dim3 gm(8 * ((nkg * nyc + 7) / 8));
        const int lgL = __builtin_ctz((unsigned)L);
#define IO_TERMS_M(GR)                                                                                             \
    do {                                                                                                           \
        (void)hipFuncSetAttribute((const void*)k_io_terms_m<GR>, hipFuncAttributeMaxDynamicSharedMemorySize,     \
                                  (int)sm);                                                                        \
        CHS_LAUNCH(k_io_terms_m<GR>, gm, dim3(2 * L), sm, s, g, a.g, a.q, a.mu, a.mu0, a.divu, a.ms, a.pr, a.rs,     \
                   a.explt0, p, TGAM_IO[ns], TROH_IO[ns], dt, talp * dt, h->rhs, h->partials,                      \
                   fold ? f->g0 : nullptr, nyc, jc, lgL);                                                          \
    } while (0)
        switch (prm->gravdir) {
            case 1: IO_TERMS_M(1); break;
            case 2: IO_TERMS_M(2); break;
            case 3: IO_TERMS_M(3); break;
            default: IO_TERMS_M(0); break;
        }
#undef IO_TERMS_M
    } else {
#define IO_TERMS(GR)                                                                                              \
    CHS_LAUNCH(k_io_terms<GR>, gi, dim3(64, 4), 0, s, g, a.g, a.q, a.mu, a.mu0, a.divu, a.ms, a.pr, a.rs, a.explt0, \
               p, TGAM_IO[ns], TROH_IO[ns], dt, talp * dt, h->rhs, h->partials, fold ? f->g0 : nullptr)
        switch (prm->gravdir) {
            case 1: IO_TERMS(1); break;
            case 2: IO_TERMS(2); break;
            case 3: IO_TERMS(3); break;
            default: IO_TERMS(0); break;
        }
#undef IO_TERMS
    }
    h->timer.end(ev, "io_terms", s);
    CHS_LAUNCH(k_io_dpgrns, dim3(1), dim3(1024), 0, s, h->partials, (int)nbt, h->volm, prm->flowdrvtp,
               -0.5 * prm->cfgv * (talp * dt), h->scalars);
    if (!fold) {
        h->timer.begin("io_massflux", s, &ev);
        CHS_LAUNCH(k_io_massflux, io_grid(h, h->M + 2), dim3(64, 4), 0, s, g, f->g, h->rhs, h->scalars, prm->nflow);
        h->timer.end(ev, "io_massflux", s);
    }
    h->timer.begin("io_divg", s, &ev);
    CHS_LAUNCH(k_io_divg, gi, dim3(64, 4), 0, s, g, f->g, fold ? h->rhs : nullptr, f->drhodtp, 1.0 / dt / talp,
               h->rhsllphi);
    h->timer.end(ev, "io_divg", s);
    CHS_TRY(poisson_solve(h, h->rhsllphi, h->dph, 1, 0.5e-10 / (talp * dt)));
    h->timer.begin("io_update", s, &ev);
    CHS_LAUNCH(k_io_update, gi, dim3(64, 4), 0, s, g, f->g, f->pr, h->dph, dt * talp, nullptr, 0.0,
               fold ? h->rhs : nullptr, h->scalars, prm->nflow);
    h->timer.end(ev, "io_update", s);
    return CHS_OK;
}

int io_divgck(chs_handle* h, const chs_io_fields* f, double* div_g, double* div_q) {
    CHS_CHECK(h && f && f->g && f->q && f->drhodtp, CHS_E_ARG, "chs_divgck_io: NULL argument");
    const hipStream_t s = h->stream;
    const int nb = 1024;
    CHS_LAUNCH(k_io_divgck, dim3(nb), dim3(256), 0, s, io_geo(h), f->g, f->q, f->drhodtp, h->partials);
    CHS_LAUNCH(k_io_max2, dim3(1), dim3(256), 0, s, h->partials, nb, h->scalars + 1);
    double host[2];
    CHS_HIP(hipMemcpyAsync(host, h->scalars + 1, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    CHS_HIP(hipStreamSynchronize(s));
    if (div_g) *div_g = host[0];
    if (div_q) *div_q = host[1];
    return CHS_OK;
}

}  // namespace chs

extern "C" {

int chs_io_prepare(chs_handle* h, const chs_thermal_fields* tf, const chs_io_fields* f, const chs_io_params* p,
                   int ns) {
    return chs::io_prepare(h, tf, f, p, ns);
}

int chs_solverrk3_mom_io(chs_handle* h, const chs_thermal_fields* tf, const chs_io_fields* f, const chs_io_params* p,
                         int ns, double dt) {
    return chs::io_substep(h, tf, f, p, ns, dt);
}

int chs_divgck_io(chs_handle* h, const chs_io_fields* f, double* maxdiv_g, double* maxdiv_q) {
    return chs::io_divgck(h, f, maxdiv_g, maxdiv_q);
}

}  // extern "C"
