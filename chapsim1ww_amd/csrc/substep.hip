// One RK3 substep of the turbulence-generator channel on the device-resident
// state: SOLVERRK3_MOM_tg(NS) (src/SOLVERRK3_MOM_tg.f90:18-61).
//
//   k_rhs_x     CONVECTION_X/Y/Z_tg + RHS_CvLpGpS_tg for u, v, w, fused with
//               the x line solve of MOMFA1_X_tg (rows staged in LDS) and the
//               block sums of the constant-mass-flux integral
//   k_massflux  DPGRNS = sum / VOLM_tg (deterministic order)
//   k_z_solve   MOMFA2_Z_tg for u, v, w (u: minus DPGRNS pushed through the x solve)
//   k_y_solve   MOMFA3_Y_tg for u, v, w + Q += RHS (MOMFA_tg:39-45) + BC_WALL_Q_tg
//   k_divg      DIVG_tg
//   Poisson     FFT99_POIS3D_periodicxz + BC_WALL_DPH_tg (poisson.hip)
//   k_update    VELOUPDT_tg + BC_WALL_Q_tg + PRCALC_tg + BC_WALL_PR_tg
//
// Layouts are the reference's: Q(L, 0:M+1, N, 3), PR/DPH(L, 0:M+1, N),
// RHS(L, M, N) x 3 components, CONVH0(L, M, N, 3), RHSLLPHI(L, M, N).
// Indices below are 0-based: i in [0,L), j in [0,M) (Fortran J = j+1, stored
// at ghost-aware row j+1), k in [0,N).
#include <algorithm>
#include <cmath>
#include <list>
#include <mutex>
#include <vector>

#include "chs_common.hpp"
#include "tdma_dev.hpp"

namespace chs {

struct Geo {
    int L, M, N;
    int mc;   // the streamwise component NFLOW - 1 (0 or 2), which takes the flow-drive term
    double dxi, dzi, dxqi, dzqi;
    const double *dyfi, *dyci, *wff, *wfb, *amvr, *acvr, *apvr;
};

// 32-bit offsets: chs_create rejects grids whose 3-component arrays reach 2^31
// elements, so every index below fits an int (cheaper address arithmetic).
__device__ __forceinline__ int qidx(const Geo& g, int c, int i, int jg, int k) {
    return i + g.L * (jg + (g.M + 2) * (k + g.N * c));
}
__device__ __forceinline__ int gidx(const Geo& g, int i, int jg, int k) {
    return i + g.L * (jg + (g.M + 2) * k);
}
__device__ __forceinline__ int cidx(const Geo& g, int i, int j, int k) {
    return i + g.L * (j + g.M * k);
}

// q = e / d, r = e - q d for 0 <= e < 2^22: shifts when d = 2^lg (lg >= 0),
// else a float estimate with one correction either way (exact in that range)
__device__ __forceinline__ void divmod_row(int e, int d, int lg, float dinv, int& q, int& r) {
    if (lg >= 0) {
        q = e >> lg;
        r = e & ((1 << lg) - 1);
        return;
    }
    q = (int)((float)e * dinv);
    r = e - q * d;
    if (r < 0) {
        --q;
        r += d;
    } else if (r >= d) {
        ++q;
        r -= d;
    }
}

constexpr int KR = 4;        // k rows per k_rhs_x block (segmented x solve)
constexpr int KRS = 4;       // k rows per k_rhs_xw block
constexpr int XLPB = 16;     // line slots of the x solve (3*KR = 12 used)
constexpr int XMS = 8;       // rows per segment of the x solve

struct RkCoef {
    double tgam, troh, talp, dt;
};

// Neighbour offsets of a cell within its row plane: i+1, i-1 (periodic) and
// k+1, k-1 (periodic in a (L, 0:M+1, N) array; plain in an LDS plane that
// carries the z neighbours as halo rows).
struct Nb {
    int ip, im, kp, km;
};

// One component of Q (or PR) around a cell: v(dj, o) is the value at row
// J + dj (dj in -1..1, a constant at every call site) and in-row offset o.
// GCol reads the ghost-aware global array (rows L apart); RCol reads the LDS
// ring of k_rhs_xm, which holds rows J-1, J, J+1 as separate planes, in the
// kernel's compute type T.
struct GCol {
    const double* __restrict__ p;
    int L;
    __device__ __forceinline__ double operator()(int dj, int o) const { return p[dj * L + o]; }
};
template <class T>
struct RColT {
    const T *m, *c, *p;
    __device__ __forceinline__ T operator()(int dj, int o) const { return (dj < 0 ? m : dj > 0 ? p : c)[o]; }
};
using RCol = RColT<double>;

// y coefficients of the block's row j (uniform over the block), in the
// compute type T (double: the values themselves)
template <class T>
struct YCoefT {
    T dyfi, dyci, wf0, wb0, wf1, wb1;
};
using YCoef = YCoefT<double>;

template <class T = double>
__device__ __forceinline__ YCoefT<T> ycoef(const Geo& g, int j) {
    YCoefT<T> y;
    y.dyfi = T(ldc(g.dyfi, j));
    y.dyci = T(ldc(g.dyci, j));
    y.wf0 = T(ldc(g.wff, j));
    y.wb0 = T(ldc(g.wfb, j));
    y.wf1 = T(ldc(g.wff, j + 1));
    y.wb1 = T(ldc(g.wfb, j + 1));
    return y;
}

// Convection H_c (CONVECTION_X/Y/Z_tg: src/CONVECTION_X.f90:23-74,
// src/CONVECTION_Y.f90:22-127, src/CONVECTION_Z.f90:22-112; channel: RCCI1 =
// RNDI1 = RCCI2 = 1, XND2CL = YND2CL = ZND2CL = 1/2) at the cell (i, J, k)
// of U, V, W, evaluated in T (FP64 form: double; mixed form: float).
template <class A, class T = double>
__device__ __forceinline__ T conv_x(const Geo& g, const YCoefT<T>& y, const A& U, const A& V, const A& W,
                                    const Nb& n) {
    const T COE1 = T(0.5 * 0.5 * g.dxi);
    const T COE2 = y.dyfi * T(0.5) * T(1.0);
    const T COE3 = T(0.5 * 0.5 * g.dzi * 1.0);
    const T u_c = U(0, 0), u_p = U(0, n.ip), u_m = U(0, n.im);
    const T H11 = ((u_p + u_c) * (u_p + u_c) - (u_m + u_c) * (u_m + u_c)) * COE1;
    const T H12 = ((V(1, 0) + V(1, n.im)) * (y.wf1 * U(1, 0) + y.wb1 * u_c) -
                   (V(0, 0) + V(0, n.im)) * (y.wf0 * u_c + y.wb0 * U(-1, 0))) * COE2;
    const T H13 = ((W(0, n.kp) + W(0, n.im + n.kp)) * (U(0, n.kp) + u_c) -
                   (W(0, 0) + W(0, n.im)) * (u_c + U(0, n.km))) * COE3;
    return -(H11 + H12 + H13);
}

template <class A, class T = double>
__device__ __forceinline__ T conv_y(const Geo& g, const YCoefT<T>& y, const A& U, const A& V, const A& W,
                                    const Nb& n) {
    const T COE1 = T(g.dxi * 0.5);
    const T COE2 = y.dyci * T(0.5) * T(0.5);
    const T COE31 = T(g.dzi * 0.5 * 1.0);
    const T wf = y.wf0, wb = y.wb0;
    const T v_c = V(0, 0);
    const T H21 = ((wf * U(0, n.ip) + wb * U(-1, n.ip)) * (V(0, n.ip) + v_c) -
                   (wf * U(0, 0) + wb * U(-1, 0)) * (v_c + V(0, n.im))) * COE1;
    const T v_p = V(1, 0), v_m = V(-1, 0);
    const T H22 = ((v_p * T(1.0) + v_c * T(1.0)) * (v_p + v_c) - (v_c * T(1.0) + v_m * T(1.0)) * (v_c + v_m)) * COE2;
    const T H23 = ((wf * W(0, n.kp) * T(1.0) + wb * W(-1, n.kp) * T(1.0)) * (V(0, n.kp) + v_c) -
                   (wf * W(0, 0) * T(1.0) + wb * W(-1, 0) * T(1.0)) * (v_c + V(0, n.km))) * COE31;
    return -(H21 + H22 + H23);
}

template <class A, class T = double>
__device__ __forceinline__ T conv_z(const Geo& g, const YCoefT<T>& y, const A& U, const A& V, const A& W,
                                    const Nb& n) {
    const T COE1 = T(0.5 * 0.5 * g.dxi);
    const T COE2 = y.dyfi * T(0.5);
    const T COE3 = T(0.5 * 0.5 * g.dzi * 1.0);
    const T w_c = W(0, 0);
    const T H31 = ((U(0, n.ip) + U(0, n.ip + n.km)) * (W(0, n.ip) + w_c) -
                   (U(0, 0) + U(0, n.km)) * (w_c + W(0, n.im))) * COE1;
    const T H32 = ((V(1, 0) + V(1, n.km)) * (y.wf1 * W(1, 0) * T(1.0) + y.wb1 * w_c * T(1.0)) -
                   (V(0, 0) + V(0, n.km)) * (y.wf0 * w_c * T(1.0) + y.wb0 * W(-1, 0) * T(1.0))) * COE2;
    const T w_p = W(0, n.kp), w_m = W(0, n.km);
    const T H33 = ((w_p + w_c) * (w_p + w_c) - (w_c + w_m) * (w_c + w_m)) * COE3;
    return -(H31 + H32 + H33);
}

// Per-row coefficients of the RHS: the y viscous stencil of each component
// and the RK3 factors.
template <class T>
struct RowCoefT {
    YCoefT<T> y;
    T ap[3], ac[3], am[3];
};
using RowCoef = RowCoefT<double>;

template <class T = double>
__device__ __forceinline__ RowCoefT<T> row_coef(const Geo& g, int j) {
    RowCoefT<T> rc;
    rc.y = ycoef<T>(g, j);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        rc.ap[c] = T(ldc(g.apvr, c * (g.M + 1) + j));
        rc.ac[c] = T(ldc(g.acvr, c * (g.M + 1) + j));
        rc.am[c] = T(ldc(g.amvr, c * (g.M + 1) + j));
    }
    return rc;
}

// RHS_CvLpGpS_tg (src/RHS_CvLpGpS_tg.f90:117-176) at one cell for u, v, w:
// conv[c] = H_c (the new CONVH0), rhs[c] = the x-solve right-hand side.
// h0[c] is the cell's CONVH0 of the previous substep, P accesses PR.
// The pressure gradient is formed in FP64 from FP64 pressures (AP), whatever
// T: the pressure carries a large common level whose FP32 rounding would
// swamp its differences.
template <class A, class T = double, class AP = A>
__device__ __forceinline__ void rhs_cell(const Geo& g, const RkCoef& rk, double cvisc, const RowCoefT<T>& rc,
                                         const A& U, const A& V, const A& W, const AP& P, const Nb& n,
                                         const T* h0, T* conv, T* rhs) {
    const T COE1 = T(rk.talp * cvisc);
    const double COE2 = rk.talp * rk.dt;
    const T tgam = T(rk.tgam), troh = T(rk.troh), dt = T(rk.dt);
    const T dxqi = T(g.dxqi), dzqi = T(g.dzqi);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const A& X = c == 0 ? U : c == 1 ? V : W;
        conv[c] = c == 0 ? conv_x<A, T>(g, rc.y, U, V, W, n) : c == 1 ? conv_y<A, T>(g, rc.y, U, V, W, n)
                                                                       : conv_z<A, T>(g, rc.y, U, V, W, n);
        const T RHSC = tgam * conv[c] + troh * h0[c];
        const T qc = X(0, 0);
        const T RHSL = COE1 * ((X(0, n.ip) - T(2.0) * qc + X(0, n.im)) * dxqi +
                               (X(0, n.kp) - T(2.0) * qc + X(0, n.km)) * dzqi * T(1.0) +
                               (X(1, 0) * rc.ap[c] + qc * rc.ac[c] + X(-1, 0) * rc.am[c]));
        const T r = (RHSC + RHSL) * dt;
        double pgm;
        if (c == 0)
            pgm = (P(0, 0) - P(0, n.im)) * g.dxi * COE2;
        else if (c == 1)
            pgm = (P(0, 0) - P(-1, 0)) * (double(rc.y.dyci) / 1.0) * COE2;
        else
            pgm = (P(0, 0) - P(0, n.km)) * g.dzi * COE2;
        rhs[c] = r - T(pgm);
    }
}

// Position of padded element p of an LDS row.  KH == 0: plain.  KH > 0 (rows
// of L + 2*KH entries, KH even): element p sits in sub-row p % 4 at p / 4, so
// the lanes of a wave reading elements 4*lane + t (t fixed) hit consecutive
// words.
template <int KH>
__device__ __forceinline__ int row_pos(int p, int Q4) {
    return KH > 0 ? (p & 3) * Q4 + (p >> 2) : p;
}

// Phase 1 of the fused RHS (RHS_CvLpGpS_tg, src/RHS_CvLpGpS_tg.f90:117-176)
// for the kr rows (j, k0 .. k0+kr-1) of u, v, w: CONVH0 <- H, and the x-solve
// right-hand side into LDS row c*kr + r (stride LS) at padded position KH + i;
// with KH > 0 the KH entries either side are the periodic halo.  Returns this
// thread's part of the mass-flux integrand sum RHS_u / DYFI.
template <int KH>
__device__ __forceinline__ double rhs_rows(const Geo& g, const RkCoef& rk, double cvisc,
                                           const double* __restrict__ Q, const double* __restrict__ PR,
                                           double* __restrict__ H0, double* sR, int LS, int kr, int lgL, int j,
                                           int k0) {
    const int L = g.L, M = g.M, N = g.N;
    const int LP = L * (M + 2);  // one k plane of a ghost-aware array
    const int CS = LP * N;       // component stride of Q
    const int LMN = L * M * N;
    const int J = j + 1;
    const int Q4 = LS >> 2;
    const RowCoef rc = row_coef(g, j);
    double msum = 0.0;
    const float Linv = 1.0f / (float)L;
    for (int e = threadIdx.x; e < kr * L; e += blockDim.x) {
        int r, i;
        divmod_row(e, L, lgL, Linv, r, i);
        const int k = k0 + r;
        Nb n;
        n.ip = i == L - 1 ? 1 - L : 1;
        n.im = i == 0 ? L - 1 : -1;
        n.kp = k == N - 1 ? (1 - N) * LP : LP;
        n.km = k == 0 ? (N - 1) * LP : -LP;
        const int o = i + L * J + LP * k;
        const GCol U{Q + o, L}, V{Q + o + CS, L}, W{Q + o + 2 * CS, L}, P{PR + o, L};
        const int ci = i + L * j + L * M * k;
        // All three components are evaluated before anything is stored, so
        // their loads issue as one batch.  v on the bottom wall face (j = 0)
        // is not advanced (NII = 2): its values are computed from in-bounds
        // rows and neither CONVH0 nor its LDS row (never solved) keeps them.
        double h0[3], conv[3], rhs[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) h0[c] = H0[ci + c * LMN];
        rhs_cell(g, rk, cvisc, rc, U, V, W, P, n, h0, conv, rhs);
        msum += (g.mc == 0 ? rhs[0] : rhs[2]) / rc.y.dyfi / 1.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (c != 1 || j != 0) H0[ci + c * LMN] = conv[c];
            double* row = sR + (c * kr + r) * LS;
            row[row_pos<KH>(KH + i, Q4)] = rhs[c];
            if (KH > 0) {
                if (i < KH) row[row_pos<KH>(KH + i + L, Q4)] = rhs[c];
                if (i >= L - KH) row[row_pos<KH>(KH + i - L, Q4)] = rhs[c];
            }
        }
    }
    return msum;
}

// Block sum in a fixed order: a shuffle tree per wave, then the wave sums in
// wave order by thread 0 (the result is valid there).  blockDim % 64 == 0.
__device__ __forceinline__ double block_sum(double v, double* red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int q = 0; q < (int)(blockDim.x >> 6); ++q) s += red[q];
    return s;
}

// ---------------------------------------------------------------------------
// k_rhs_x: block = (j, KR consecutive k).  Phase 1 computes RHS_c for the 3
// components of its rows into LDS and writes CONVH0; phase 2 solves the x
// lines (cyclic, constant coefficients) from LDS by the segmented Thomas
// method and writes RHS.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) k_rhs_x(Geo g, RkCoef rk, double cvisc, const double* __restrict__ Q,
                                               const double* __restrict__ PR, double* __restrict__ H0,
                                               double* __restrict__ RHS, double* __restrict__ part, double xa,
                                               double xb, double xc, int lgL) {
    extern __shared__ unsigned char smem[];
    const int L = g.L;
    const int LS = L + 1;  // padded LDS row stride (bank spread for the segment reads)
    // LDS: sR (the 3*KR RHS rows) | red (block sum); the reduced systems of
    // the x solves reuse the front once sR has been read into registers
    double* sR = reinterpret_cast<double*>(smem);
    const int nkb = g.N / KR;
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const bool active = t < g.M * nkb;
    // k-group-major: consecutive tiles (same XCD) walk j, so the j-1/j+1 rows
    // a tile reads were loaded by its neighbours moments before (L2 hits)
    const int kg = active ? t / g.M : 0;
    const int j = active ? t - kg * g.M : 0;
    const int k0 = kg * KR;
    const int LMN = L * g.M * g.N;
    double msum = 0.0;
    if (active) msum = rhs_rows<0>(g, rk, cvisc, Q, PR, H0, sR, LS, KR, lgL, j, k0);
    // block sum of the mass-flux integrand (deterministic tree)
    {
        double* red = reinterpret_cast<double*>(smem + sizeof(double) * 3 * KR * LS);
        red[threadIdx.x] = msum;
        __syncthreads();
        int half = 1;
        while (2 * half < (int)blockDim.x) half *= 2;
        for (int s = half; s > 0; s >>= 1) {   // any block size (2 L threads)
            if ((int)threadIdx.x < s && threadIdx.x + s < blockDim.x) red[threadIdx.x] += red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0 && active) part[t] = red[0];
        __syncthreads();
    }
    // x solves of the 3*KR rows, XMS rows per segment, L/XMS segments
    const int P = L / XMS;
    const int ll = threadIdx.x % XLPB, seg = threadIdx.x / XLPB;
    const int c = ll / KR, r = ll - c * KR;
    const bool valid = active && ll < 3 * KR && !(c == 1 && j == 0);
    double a[XMS], b[XMS], cc[XMS], d[XMS];
#pragma unroll
    for (int l = 0; l < XMS; ++l) {
        a[l] = xa;
        b[l] = xb;
        cc[l] = xc;
        d[l] = valid ? sR[(c * KR + r) * LS + seg * XMS + l] : 0.0;
    }
    __syncthreads();  // sR fully read: its space now holds the reduced systems
    seg_solve<XMS, double>(a, b, cc, d, XLPB, P, ll, seg, valid, true, smem);
    if (valid) {
        const int base = c * LMN + L * (j + g.M * (k0 + r));
#pragma unroll
        for (int l = 0; l < XMS; ++l) RHS[base + seg * XMS + l] = d[l];
    }
}

// ---------------------------------------------------------------------------
// k_rhs_xw: k_rhs_x with the x solve of MOMFA1_X_tg as the truncated inverse
// of its constant cyclic system: a periodic (2KX+1)-point stencil (weights
// zero-padded to the window).  Rows sit in LDS with KX halo entries each side
// (row_pos layout); each thread forms 4 consecutive outputs from one register
// window of 2KX+4 entries.  Block = 256 threads, rows (j, KR consecutive k).
// ---------------------------------------------------------------------------
template <int KX, int KR>
__global__ void __launch_bounds__(256) k_rhs_xw(Geo g, RkCoef rk, double cvisc, const double* __restrict__ Q,
                                               const double* __restrict__ PR, double* __restrict__ H0,
                                               double* __restrict__ RHS, double* __restrict__ part,
                                               const double* __restrict__ wx, int lgL) {
    extern __shared__ double sw[];
    constexpr int W = 2 * KX + 1;
    const int L = g.L;
    const int LS = L + 2 * KX;
    const int Q4 = LS >> 2;
    double* red = sw + 3 * KR * LS;
    const int nkb = g.N / KR;
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const bool active = t < g.M * nkb;
    const int kg = active ? t / g.M : 0;
    const int j = active ? t - kg * g.M : 0;
    const int k0 = kg * KR;
    double msum = 0.0;
    if (active) msum = rhs_rows<KX>(g, rk, cvisc, Q, PR, H0, sw, LS, KR, lgL, j, k0);
    const double bs = block_sum(msum, red);  // its barrier also publishes the LDS rows
    if (threadIdx.x == 0 && active) part[t] = bs;
    if (!active) return;
    double w[W];
#pragma unroll
    for (int m = 0; m < W; ++m) w[m] = wx[m];
    const int L4 = L / 4, lg4 = lgL >= 0 ? lgL - 2 : -1;
    const float L4inv = 1.0f / (float)L4;
    const int LMN = L * g.M * g.N;
    for (int e = threadIdx.x; e < 3 * KR * L4; e += blockDim.x) {
        int cr, g4;
        divmod_row(e, L4, lg4, L4inv, cr, g4);
        const int c = cr / KR, r = cr - c * KR;
        if (c == 1 && j == 0) continue;
        // padded elements 4*g4 + t, t = 0 .. W+2: sub-row t % 4, entry g4 + t / 4
        const double* row = sw + cr * LS + g4;
        double x[W + 3];
#pragma unroll
        for (int tt = 0; tt < W + 3; ++tt) x[tt] = row[(tt & 3) * Q4 + (tt >> 2)];
        double o[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            double acc = 0.0;
#pragma unroll
            for (int m = 0; m < W; ++m) acc += w[m] * x[m + u];
            o[u] = acc;
        }
        double* dst = RHS + (c * LMN + 4 * g4 + L * (j + g.M * (k0 + r)));
        reinterpret_cast<double2*>(dst)[0] = make_double2(o[0], o[1]);
        reinterpret_cast<double2*>(dst)[1] = make_double2(o[2], o[3]);
    }
}

// ---------------------------------------------------------------------------
// k_rhs_xm: k_rhs_xw marching along y.  A block owns KRS = 4 consecutive k
// rows of every x line and a chunk of consecutive y rows j0 .. j1-1; thread
// (rb, i) of its (KRS/CPT)*L threads handles the CPT cells (i, k0 + rb +
// u*KRS/CPT), u < CPT, of each row.  LDS holds a ring of the rows J-1, J, J+1 of u, v, w (KRS + 2 k
// rows each: the periodic z neighbours as halo) and row J of PR (KRS + 1 k
// rows), so each value of Q and PR comes from global memory once per block
// instead of once per stencil use; the rows a step adds to the ring are
// loaded PD steps ahead, in registers.  Per row: RHS of the three components
// from the ring into the x-solve rows, barrier, the x stencil and the RHS
// stores, the prefetched rows into the ring, barrier.  The mass-flux partial
// of row j goes to part[kg*M + j], the slot k_rhs_xw's tile (j, kg) writes.
// ---------------------------------------------------------------------------
constexpr int XM_KP = KRS + 2;  // k rows of a ring plane
constexpr int XM_PD = 1;        // steps the ring's next rows are loaded ahead (2 with 2 cells: no faster)
constexpr int XM_CPT = 1;       // cells of a row per thread (1 or 2; 2 measured no faster)

// LDS of k_rhs_xm in compute type T: the ring and the x-solve rows in T,
// then (8-byte aligned) the PR plane in FP64 and 16 doubles of block sums
inline size_t rhs_xm_smem(int L, int KX, size_t tsize = sizeof(double)) {
    size_t b = tsize * ((size_t)9 * XM_KP * L + (size_t)3 * KRS * (L + 2 * KX));
    return ((b + 7) & ~(size_t)7) + sizeof(double) * ((size_t)(KRS + 1) * L + 16);
}

// four consecutive values (16-byte aligned for float, 32 for double)
__device__ __forceinline__ void store4(double* dst, const double* v) {
    reinterpret_cast<double2*>(dst)[0] = make_double2(v[0], v[1]);
    reinterpret_cast<double2*>(dst)[1] = make_double2(v[2], v[3]);
}
__device__ __forceinline__ void store4(float* dst, const float* v) {
    reinterpret_cast<float4*>(dst)[0] = make_float4(v[0], v[1], v[2], v[3]);
}

// T: the compute type of the cells' right-hand sides, the LDS ring and the x
// solve (double: the FP64 form; float: the mixed form's FP32 stencils, whose
// RHS and CONVH0 are stored in FP32).  Q and PR stay FP64 in HBM.
template <int KX, int PD, int CPT, class T = double>
__global__ void __launch_bounds__(CPT == 1 ? 1024 : 512) k_rhs_xm(Geo g, RkCoef rk, double cvisc, const double* __restrict__ Q,
                                               const double* __restrict__ PR, T* __restrict__ H0,
                                               T* __restrict__ RHS, double* __restrict__ part,
                                               const double* __restrict__ wx, int lgL, int nyc, int jc) {
    static_assert(KRS % CPT == 0 && KRS / CPT >= 2, "CPT cells per thread, a low and a high halo row");
    constexpr int RB = KRS / CPT;  // thread rows
    extern __shared__ double smd[];
    T* sm = reinterpret_cast<T*>(smd);
    // the pressure values stay FP64 (the gradient of a large common level)
    using TP = double;
    constexpr int KR = KRS, KP = XM_KP, W = 2 * KX + 1;
    const int L = g.L, M = g.M, N = g.N;
    const int PL = KP * L;            // one component of one ring row
    const int SL = 3 * PL;            // one ring row (u, v, w)
    T* ring = sm;                     // 3 ring rows
    const int LS = L + 2 * KX, Q4 = LS >> 2;
    T* xr = sm + 3 * SL;              // 3*KR x-solve rows (row_pos layout)
    TP* pln = reinterpret_cast<TP*>(smd + (sizeof(T) * (size_t)(9 * KP * L + 3 * KR * LS) + 7) / 8);   // PR row J, k0-1 .. k0+KR-1
    double* red = smd + (sizeof(T) * (size_t)(9 * KP * L + 3 * KR * LS) + 7) / 8 + (KR + 1) * L;
    const int nt = (N / KR) * nyc;
    const int per = (nt + 7) / 8;
    const int t = (blockIdx.x % 8) * per + blockIdx.x / 8;  // XCD x: tiles x*per .. (x+1)*per-1
    if (t >= nt) return;
    const int kg = t / nyc;
    const int j0 = (t - kg * nyc) * jc, j1 = min(M, j0 + jc);
    const int k0 = kg * KR;
    const int tid = threadIdx.x;
    int rb, i;
    divmod_row(tid, L, lgL, 1.0f / (float)L, rb, i);
    const int LP = L * (M + 2), CS = LP * N, LMN = L * M * N;
    // The thread's cells (k rows rb + u*RB) and one halo row: k0-1 (rb == 0),
    // k0+KR (rb == RB-1), or, for the rows between, its first cell once more
    // (same value, same place); PR needs only the k0-1 halo and the other
    // threads move their first cell's PR twice likewise.  So every load and
    // LDS store is unconditional and a step's loads issue as one batch.
    const bool lo = rb == 0, hi = rb == RB - 1;
    int gc[CPT], lc[CPT], kc[CPT];
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
        kc[u] = k0 + rb + RB * u;
        gc[u] = i + LP * kc[u];             // global, row 0
        lc[u] = (rb + RB * u + 1) * L + i;  // in a ring plane
    }
    const int kh = lo ? (k0 == 0 ? N - 1 : k0 - 1) : hi ? (k0 + KR == N ? 0 : k0 + KR) : kc[0];
    const int gh = i + LP * kh, lh = lo ? i : hi ? (KP - 1) * L + i : lc[0];
    const int gph = lo ? gh : gc[0], lph = lo ? i : lc[0];
    Nb n;
    n.ip = i == L - 1 ? 1 - L : 1;
    n.im = i == 0 ? L - 1 : -1;
    n.kp = L;
    n.km = -L;
    // fill the ring with rows j0 .. j0+2 and PR row j0+1
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        const int row = L * (j0 + s);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            T a[CPT];
#pragma unroll
            for (int u = 0; u < CPT; ++u) a[u] = T(Q[gc[u] + row + c * CS]);
            const T ah = T(Q[gh + row + c * CS]);
#pragma unroll
            for (int u = 0; u < CPT; ++u) ring[s * SL + c * PL + lc[u]] = a[u];
            ring[s * SL + c * PL + lh] = ah;
        }
    }
    {
        const int row = L * (j0 + 1);
        TP a[CPT];
#pragma unroll
        for (int u = 0; u < CPT; ++u) a[u] = TP(PR[gc[u] + row]);
        const TP ah = TP(PR[gph + row]);
#pragma unroll
        for (int u = 0; u < CPT; ++u) pln[lc[u]] = a[u];
        pln[lph] = ah;
    }
    TP pm[CPT];  // own PR at row J-1
    T h0[CPT][3];    // own CONVH0 of row j
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
        pm[u] = TP(PR[gc[u] + L * j0]);
#pragma unroll
        for (int c = 0; c < 3; ++c) h0[u][c] = T(H0[i + L * j0 + L * M * kc[u] + c * LMN]);
    }
    const int L4 = L / 4, lg4 = lgL >= 0 ? lgL - 2 : -1;
    const float L4inv = 1.0f / (float)L4;
    const int nw = blockDim.x >> 6;
    // Nothing loaded above may still be in flight when the loop starts:
    // otherwise every step would open with a wait for all outstanding memory
    // operations (the previous step's stores included).  s_waitcnt vmcnt(0).
    __builtin_amdgcn_s_waitcnt(0xF70);
    __syncthreads();
    // What step j puts into the ring at its end (Q row J+2, PR row J+1) and
    // the CONVH0 row j+1 of the next step.  Rows past the chunk's end are
    // clamped into the array: loaded, never used.
    struct Pre {
        T q[CPT][3], qh[3], h[CPT][3];
        TP p[CPT], ph;
    };
    // CONVH0 is loaded first: the wait for the Q rows at the ring store then
    // covers it, so that using it in the next step waits for nothing (the
    // vector memory counter is in order: a later wait would also wait for
    // the stores issued since).
    auto load_pre = [&](int jj, Pre& x) {
        const int rq = L * min(jj + 3, M + 1), rp = L * min(jj + 2, M + 1), rh = L * min(jj + 1, M - 1);
#pragma unroll
        for (int u = 0; u < CPT; ++u) {
#pragma unroll
            for (int c = 0; c < 3; ++c) x.h[u][c] = T(H0[i + rh + L * M * kc[u] + c * LMN]);
            x.p[u] = TP(PR[gc[u] + rp]);
        }
        x.ph = TP(PR[gph + rp]);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
#pragma unroll
            for (int u = 0; u < CPT; ++u) x.q[u][c] = T(Q[gc[u] + rq + c * CS]);
            x.qh[c] = T(Q[gh + rq + c * CS]);
        }
    };
    Pre pa;
    if (PD == 2) load_pre(j0, pa);
    int sM = 0, s0 = 1, sP = 2;
    for (int j = j0; j < j1; ++j) {
        Pre pb;
        if (PD == 2)
            load_pre(j + 1, pb);
        else
            load_pre(j, pa);
        // RHS of row j from the ring
        const RowCoefT<T> rc = row_coef<T>(g, j);
        const double dyfj = ldc(g.dyfi, j);
        double v = 0.0;
        TP pc[CPT];
#pragma unroll
        for (int u = 0; u < CPT; ++u) {
            const int o = lc[u];
            const RColT<T> U{ring + sM * SL + o, ring + s0 * SL + o, ring + sP * SL + o};
            const RColT<T> V{U.m + PL, U.c + PL, U.p + PL};
            const RColT<T> Wc{V.m + PL, V.c + PL, V.p + PL};
            pc[u] = pln[o];
            const RColT<TP> P{&pm[u], pln + o, nullptr};
            T conv[3], rhs[3];
            rhs_cell<RColT<T>, T, RColT<TP>>(g, rk, cvisc, rc, U, V, Wc, P, n, h0[u], conv, rhs);
            const int ci = i + L * j + L * M * kc[u];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                if (c != 1 || j != 0) H0[ci + c * LMN] = conv[c];
                T* row = xr + (c * KR + rb + RB * u) * LS;
                row[row_pos<KX>(KX + i, Q4)] = rhs[c];
                if (i < KX) row[row_pos<KX>(KX + i + L, Q4)] = rhs[c];
                if (i >= L - KX) row[row_pos<KX>(KX + i - L, Q4)] = rhs[c];
            }
            v += double(g.mc == 0 ? rhs[0] : rhs[2]) / dyfj / 1.0;
        }
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
        // row J+2 replaces row J-1 in the ring, PR row J+1 replaces row J
        // (both free since the barrier), before the RHS stores are issued
#pragma unroll
        for (int c = 0; c < 3; ++c) {
#pragma unroll
            for (int u = 0; u < CPT; ++u) {
                ring[sM * SL + c * PL + lc[u]] = pa.q[u][c];
                h0[u][c] = pa.h[u][c];
            }
            ring[sM * SL + c * PL + lh] = pa.qh[c];
        }
#pragma unroll
        for (int u = 0; u < CPT; ++u) {
            pln[lc[u]] = pa.p[u];
            pm[u] = pc[u];
        }
        pln[lph] = pa.ph;
        // x stencil of the 3*KR rows, 4 outputs per work item
        for (int e = tid; e < 3 * KR * L4; e += blockDim.x) {
            int cr, g4;
            divmod_row(e, L4, lg4, L4inv, cr, g4);
            const int c = cr / KR, rr = cr - c * KR;
            if (c == 1 && j == 0) continue;
            const T* xrow = xr + cr * LS + g4;
            T x[W + 3];
#pragma unroll
            for (int tt = 0; tt < W + 3; ++tt) x[tt] = xrow[(tt & 3) * Q4 + (tt >> 2)];
            T o4[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                T acc = T(0.0);
#pragma unroll
                for (int m = 0; m < W; ++m) acc += T(ldc(wx, m)) * x[m + u];
                o4[u] = acc;
            }
            T* dst = RHS + (c * LMN + 4 * g4 + L * (j + M * (k0 + rr)));
            store4(dst, o4);
        }
        if (PD == 2) pa = pb;
        __syncthreads();
        const int s_ = sM;
        sM = s0;
        s0 = sP;
        sP = s_;
    }
}

// CONVH0 between its FP64 and FP32 stores (the mixed form's FP32 stencils
// keep it in FP32 between substeps)
__global__ void k_narrow(const double* __restrict__ a, float* __restrict__ b, size_t n) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        b[e] = float(a[e]);
}
__global__ void k_widen(const float* __restrict__ b, double* __restrict__ a, size_t n) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        a[e] = double(b[e]);
}

// Makes the FP32 (f32) or the FP64 store hold the current CONVH0.
int convh0_store(chs_handle* h, bool f32) {
    if (f32 == h->h0_f32) return CHS_OK;
    if (f32 && !h->h0f) {
        set_error("substep: the FP32 CONVH0 store of the mixed form is missing");
        return CHS_E_ARG;
    }
    const size_t n = 3 * (size_t)h->L * h->M * h->N;
    if (f32)
        CHS_LAUNCH(k_narrow, dim3(2048), dim3(256), 0, h->stream, h->convh0, h->h0f, n);
    else
        CHS_LAUNCH(k_widen, dim3(2048), dim3(256), 0, h->stream, h->h0f, h->convh0, n);
    h->h0_f32 = f32;
    return CHS_OK;
}

__global__ void k_set_scalar(double* __restrict__ out, double v) {
    if (threadIdx.x == 0) out[0] = v;
}

// DPGRNS = (sum of block partials) / VOLM_tg, summed in a fixed order.
// One block of 1024 threads, four independent partial sums per thread (its
// strided values in turn) so that four loads are in flight per thread.
__global__ void __launch_bounds__(1024) k_massflux(const double* __restrict__ part, int nb, double volm,
                                                   double* __restrict__ out) {
    __shared__ double red[1024];
    constexpr int T = 1024;
    double s4[4] = {0.0, 0.0, 0.0, 0.0};
    int i = threadIdx.x;
    for (; i + 3 * T < nb; i += 4 * T) {
#pragma unroll
        for (int q = 0; q < 4; ++q) s4[q] += part[i + q * T];
    }
    if (i < nb) s4[0] += part[i];
    if (i + T < nb) s4[1] += part[i + T];
    if (i + 2 * T < nb) s4[2] += part[i + 2 * T];
    red[threadIdx.x] = (s4[0] + s4[1]) + (s4[2] + s4[3]);
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0] / volm;
}

// ---------------------------------------------------------------------------
// k_z_solve: cyclic z lines (MOMFA2_Z_tg), lines (c, i, j), rows k.
// The mass-flux correction of u is applied here: x-solve(R - D) =
// x-solve(R) - D * xone, xone = 1/(a+b+c) of the constant x system.
// ---------------------------------------------------------------------------
template <int MS>
__global__ void __launch_bounds__(1024) k_z_solve(Geo g, double* __restrict__ RHS, const double* __restrict__ dpg,
                                                  double xone, double za, double zb, double zc, int LPB, int P,
                                                  int use_dpg) {
    extern __shared__ unsigned char smem[];
    const int LM = g.L * g.M;
    const int nlines = 3 * LM;
    const int ll = threadIdx.x % LPB, seg = threadIdx.x / LPB;
    const int bt = blockIdx.x;
    const int line = bt * LPB + ll;
    const int c = line / LM;
    const int rem = line - c * LM;
    const int j = rem / g.L;
    const bool valid = line < nlines && !(c == 1 && j == 0);
    const size_t LMN = (size_t)LM * g.N;
    const size_t base = valid ? (size_t)c * LMN + rem : 0;
    const double D = (use_dpg && c == g.mc) ? dpg[0] * xone : 0.0;
    double a[MS], b[MS], cc[MS], d[MS];
#pragma unroll
    for (int l = 0; l < MS; ++l) {
        const int k = seg * MS + l;
        a[l] = za;
        b[l] = zb;
        cc[l] = zc;
        d[l] = valid ? RHS[base + (size_t)k * LM] - D : 0.0;
    }
    seg_solve<MS, double>(a, b, cc, d, LPB, P, ll, seg, valid, true, smem);
    if (valid) {
#pragma unroll
        for (int l = 0; l < MS; ++l) RHS[base + (size_t)(seg * MS + l) * LM] = d[l];
    }
}

// ---------------------------------------------------------------------------
// k_y_solve: y lines (MOMFA3_Y_tg) for u, then v, then w of the same (i,k)
// lines, all rows j = 1..M; then Q += RHS from N2I (MOMFA_tg:39-45) and the
// wall ghost values (BC_WALL_Q_tg).  v's row-1 right-hand side is the RHS_tg
// value the reference leaves there: the u increment on j = 1.
// ---------------------------------------------------------------------------
template <int MS>
__global__ void __launch_bounds__(768) k_y_solve(Geo g, double facoe, const double* __restrict__ RHS,
                                                  double* __restrict__ Q, int LPB, int P) {
    extern __shared__ unsigned char smem[];
    const int L = g.L, M = g.M;
    const int nlines = L * g.N;
    const int ll = threadIdx.x % LPB, seg = threadIdx.x / LPB;
    const int line = blockIdx.x * LPB + ll;
    const bool valid = line < nlines;
    const int k = valid ? line / L : 0, i = valid ? line - k * L : 0;
    const size_t LMN = (size_t)L * M * g.N;
    const size_t base = (size_t)i + (size_t)L * M * k;
    double du_row1 = 0.0;
    for (int c = 0; c < 3; ++c) {
        double a[MS], b[MS], cc[MS], d[MS];
#pragma unroll
        for (int l = 0; l < MS; ++l) {
            const int j = seg * MS + l;
            if (j < M) {
                const int o = c * (M + 1) + j;
                b[l] = 1.0 - facoe * g.acvr[o];
                cc[l] = -facoe * g.apvr[o];
                a[l] = -facoe * g.amvr[o];
                d[l] = valid ? RHS[c * LMN + base + (size_t)j * L] : 0.0;
            } else {
                a[l] = 0.0;
                b[l] = 1.0;
                cc[l] = 0.0;
                d[l] = 0.0;
            }
        }
        if (seg == 0) {
            a[0] = 0.0;  // BCJ = 0
            if (c == 1) d[0] = du_row1;
        }
        {
            const int jl = M - 1;
            if (jl / MS == seg) {
#pragma unroll
                for (int l = 0; l < MS; ++l)
                    if (l == jl % MS) cc[l] = 0.0;
            }
        }
        seg_solve<MS, double>(a, b, cc, d, LPB, P, ll, seg, valid, false, smem);
        if (c == 0 && seg == 0) du_row1 = d[0];
        if (valid) {
            const int n2i = c == 1 ? 1 : 0;
#pragma unroll
            for (int l = 0; l < MS; ++l) {
                const int j = seg * MS + l;
                if (j < M && j >= n2i) {
                    const size_t qo = qidx(g, c, i, j + 1, k);
                    Q[qo] = d[l] + Q[qo];
                }
            }
            if (seg == 0) {
                Q[qidx(g, c, i, 0, k)] = 0.0;
                if (c == 1) Q[qidx(g, 1, i, 1, k)] = 0.0;
                Q[qidx(g, c, i, M + 1, k)] = 0.0;
            }
        }
        __syncthreads();  // the LDS of seg_solve is reused by the next component
    }
}

// ---------------------------------------------------------------------------
// CN solves as truncated inverses.  The factorised Crank-Nicolson systems of
// MOMFA_tg are strongly diagonally dominant at DNS time steps (off-diagonal /
// diagonal ~ 1e-4..1e-2), so every row of their inverses decays below 1e-17 of
// its sum within a few rows.  The host computes the exact inverse rows in
// extended precision and truncates them where the dropped weights sum to
// < 1e-17 of the row's absolute sum; the solves then become short stencils.
// Shapes or time steps where that needs more than KMAX rows use the
// segmented Thomas kernels above instead.
// ---------------------------------------------------------------------------
constexpr int KMAX = 16;

// y: Delta_c(j) = sum_m W_c[j][m] R_c(j+m), rows clipped to [0, M); for v the
// row-1 right-hand side is the u increment on j = 1 (what RHS_tg holds there in
// the reference).  Then Q += Delta from N2I (MOMFA_tg:39-45) and BC_WALL_Q_tg.
// y, register form: one thread per (i, k) column and chunk of jc rows (YJC,
// shorter where the grid is small); the
// window R(j-KW..j+KW) slides along j in registers and every lane of a wave
// (64 consecutive i) applies the same row weights W_c[j][.] (scalar loads).
// wy rows are zero-padded to 2KW+1 entries (zero outside [0, M)).
constexpr int YJC = 64;     // rows per thread chunk of the y window (measured at 256x192x256: 32 -> 64 rows, 257 -> 240 us)
constexpr int YNS = 4;      // prefetch slots (rows fetched ahead) of the y window (2 -> 4: 200 -> 192 us; 6, 8: no faster)
constexpr int WU = 4;       // rows fetched ahead per step of the z window (2: 140 -> 153 us; slots
                            // refilled a step ahead as in the y window: 133 -> 141 us)
template <int KW, class TR = double>
// One wave per block (chunk blockIdx.z): its LDS holds only the chunk's
// weight rows, so more blocks fit a CU.  TR: the storage type of the z-solved
// right-hand side (float in the mixed form); the y elimination is FP64.
__global__ void __launch_bounds__(64) k_y_window(Geo g, const TR* __restrict__ R2, double* __restrict__ Q,
                                                 const double* __restrict__ wy, int jc) {
    const int L = g.L, M = g.M, N = g.N;
    const int ntx = (L + 63) / 64;
    const int nch = (M + jc - 1) / jc;
    const int lane = threadIdx.x, ch = blockIdx.z;
    const int it = blockIdx.x % ntx, k = blockIdx.x / ntx;
    extern __shared__ double sw[];   // the chunk's row weights [j1 - j0][W], from row j0
    const bool valid = it * 64 + lane < L;
    const int i = valid ? it * 64 + lane : 0;
    const int j0 = ch * jc, jn = min(M, j0 + jc), j1 = valid ? jn : j0;   // invalid lanes do no rows
    const size_t LM = (size_t)L * M;
    constexpr int W = 2 * KW + 1;
    double du0 = 0.0;
    const int c = blockIdx.y;
    {
        const TR* Rc = R2 + (size_t)c * LM * N + (size_t)k * LM + i;
        for (int e = threadIdx.x; e < (jn - j0) * W; e += blockDim.x) sw[e] = wy[((size_t)c * M + j0) * W + e];
        __syncthreads();
        if (!valid) return;
        if (c == 1 && ch == 0) {
            // v's row-1 right-hand side: the u increment on j = 1 (row 0 of u's solve)
            const TR* Ru = R2 + (size_t)k * LM + i;
#pragma unroll
            for (int m = KW; m < W; ++m) du0 += wy[m] * double(Ru[(size_t)(m - KW) * L]);
        }
        double win[W];
#pragma unroll
        for (int m = 0; m < W; ++m) {
            const int jj = j0 - KW + m;
            double v = 0.0;
            if (jj >= 0 && jj < M) v = (c == 1 && jj == 0) ? du0 : double(Rc[(size_t)jj * L]);
            win[m] = v;
        }
        // YNS fixed prefetch slots, slot u serving the steps j = j0 + u mod
        // YNS (no register rotation, which made the compiler wait on the
        // pending loads): the window row entering at step j and Q's row j,
        // loaded YNS steps ahead with clamped indices (no branch between the
        // loads); rows past M enter the window as zeros
        double rs[YNS], qs[YNS];
#pragma unroll
        for (int u = 0; u < YNS; ++u) {
            rs[u] = double(Rc[(size_t)min(j0 + KW + 1 + u, M - 1) * L]);
            qs[u] = Q[qidx(g, c, i, min(j0 + u, M - 1) + 1, k)];
        }
        auto step = [&](int j, double& r, double& q) {
            const double* w = sw + (size_t)(j - j0) * W;
            double acc = 0.0;
#pragma unroll
            for (int m = 0; m < W; ++m) acc += w[m] * win[m];
            if (!(c == 1 && j == 0)) Q[qidx(g, c, i, j + 1, k)] = acc + q;
#pragma unroll
            for (int m = 0; m < W - 1; ++m) win[m] = win[m + 1];
            win[W - 1] = j + KW + 1 < M ? r : 0.0;
            r = double(Rc[(size_t)min(j + KW + 1 + YNS, M - 1) * L]);
            q = Q[qidx(g, c, i, min(j + YNS, M - 1) + 1, k)];
        };
        int j = j0;
        for (; j + YNS - 1 < j1; j += YNS) {
#pragma unroll
            for (int u = 0; u < YNS; ++u) step(j + u, rs[u], qs[u]);
        }
#pragma unroll
        for (int u = 0; u < YNS - 1; ++u)
            if (j + u < j1) step(j + u, rs[u], qs[u]);
        if (ch == 0) {
            Q[qidx(g, c, i, 0, k)] = 0.0;
            if (c == 1) Q[qidx(g, 1, i, 1, k)] = 0.0;
        }
        if (ch == nch - 1) Q[qidx(g, c, i, M + 1, k)] = 0.0;
    }
}

// z, register form: one thread per (i, j) column and chunk of zkc rows of k
// (cyclic), the window slides along k; weights are the same for every lane.
// zkc is ZKC or a multiple of it (longer chunks refill the window less often)
// and divides N; the host checks both.
// u subtracts D = DPGRNS * xone on load (the mass-flux correction pushed
// through the x solve).  Reads RHS, writes RHS2.
constexpr int ZKC = 64;
static_assert(ZKC % WU == 0, "k_z_window steps a chunk WU rows at a time: a partial step would run past the chunk");
// T: storage and arithmetic type (float in the mixed form).
template <int KW, class T = double>
__global__ void __launch_bounds__(256) k_z_window(Geo g, const T* __restrict__ R, T* __restrict__ R2,
                                                  const double* __restrict__ dpg, double xone, int use_dpg,
                                                  const double* __restrict__ wz, int zkc) {
    const int L = g.L, M = g.M, N = g.N;
    const int ntx = (L + 63) / 64;
    const int nch = N / zkc;
    int b = blockIdx.x;
    const int lane = threadIdx.x & 63, ch = threadIdx.x >> 6;     // blockDim = 64 * 4
    const int it = b % ntx; b /= ntx;
    const int j = b % M;
    const int cc = b / M;                                          // component
    const int chunk = blockIdx.y * 4 + ch;
    const int i = it * 64 + lane;
    if (i >= L || chunk >= nch || (cc == 1 && j == 0)) return;
    const size_t LM = (size_t)L * M;
    const T* Rc = R + (size_t)cc * LM * N + (size_t)j * L + i;
    T* Oc = R2 + (size_t)cc * LM * N + (size_t)j * L + i;
    const T D = T((use_dpg && cc == g.mc) ? dpg[0] * xone : 0.0);
    constexpr int W = 2 * KW + 1;
    T w[W];
#pragma unroll
    for (int m = 0; m < W; ++m) w[m] = T(wz[m]);
    const int k0 = chunk * zkc;
    T win[W];
#pragma unroll
    for (int m = 0; m < W; ++m) {
        int kk = k0 - KW + m;
        kk = kk < 0 ? kk + N : (kk >= N ? kk - N : kk);
        win[m] = Rc[(size_t)kk * LM] - D;
    }
    for (int kb = k0; kb < k0 + zkc; kb += WU) {
        T rn[WU];
#pragma unroll
        for (int u = 0; u < WU; ++u) {
            int kn = kb + u + KW + 1;
            kn = kn >= N ? kn - N : kn;
            rn[u] = Rc[(size_t)kn * LM] - D;
        }
#pragma unroll
        for (int u = 0; u < WU; ++u) {
            T acc = T(0.0);
#pragma unroll
            for (int m = 0; m < W; ++m) acc += w[m] * win[m];
            Oc[(size_t)(kb + u) * LM] = acc;
#pragma unroll
            for (int m = 0; m < W - 1; ++m) win[m] = win[m + 1];
            win[W - 1] = rn[u];
        }
    }
}

// DIVG_tg (src/DIVG.f90:21-57): f = div(Q) / (DT * TALP).  One thread per
// column of DVJ rows of j in tiles of 64 i x 4 k (row_tile over the row
// groups: each XCD walks its tiles in j order, so the next group's first v
// row is in its L2); v's row J+1 of one cell is row J of the next, and every
// load is issued before the first store (rows past M are clamped into the
// array and not stored).  32-bit offsets.
constexpr int DVJ = 4;   // rows per thread (measured at 256x192x256: 1 -> 4, 89 -> 83 us)
__global__ void __launch_bounds__(256) k_divg(Geo g, double coe0, const double* __restrict__ Q, double* __restrict__ F) {
    int it, jg, kt, tile;
    const int M = g.M;
    if (!row_tile(g.L, g.N, (M + DVJ - 1) / DVJ, &it, &jg, &kt, &tile)) return;
    const int L = g.L, N = g.N;
    const int i = it * 64 + (threadIdx.x & 63), k = kt * 4 + (threadIdx.x >> 6);
    if (i >= L || k >= N) return;
    const int LP = L * (M + 2), CS = LP * N;
    const int ip = i == L - 1 ? 1 - L : 1, kp = k == N - 1 ? (1 - N) * LP : LP;
    const int j0 = jg * DVJ;
    const double* U = Q + (i + L * (j0 + 1) + LP * k);
    const double* V = U + CS;
    const double* W = V + CS;
    double u0[DVJ], u1[DVJ], w0[DVJ], w1[DVJ], v[DVJ + 1];
#pragma unroll
    for (int r = 0; r < DVJ; ++r) {
        const int o = L * (min(j0 + r, M - 1) - j0);
        u0[r] = U[o];
        u1[r] = U[o + ip];
        w0[r] = W[o];
        w1[r] = W[o + kp];
    }
    // v rows j0 .. j0 + DVJ, clamped to row M (stored at M + 1, the ghost
    // row), so the last row of a short group still finds its J + 1
#pragma unroll
    for (int r = 0; r <= DVJ; ++r) v[r] = V[L * (min(j0 + r, M) - j0)];
#pragma unroll
    for (int r = 0; r < DVJ; ++r) {
        const int j = j0 + r;
        if (j >= M) break;
        const double COE1 = g.dxi / 1.0, COE2 = ldc(g.dyfi, j) / 1.0;
        const double D = (u1[r] - u0[r]) * COE1 + (v[r + 1] - v[r]) * COE2 + (w1[r] - w0[r]) * g.dzi;
        F[i + L * (j + M * k)] = D * coe0;
    }
}

// VELOUPDT_tg (src/VELO_CORRECTION.f90:37-75) + PRCALC_tg
// (src/PRESSURE_CORRECTION.f90:11-20) + BC_WALL_Q_tg / BC_WALL_PR_tg, in the
// row-group tiles of k_divg: UPJ rows of j per thread, DPH's row J-1 of one
// cell being row J of the previous.  Every load is issued before any store
// (rows past M are clamped into the array and not stored; v's wall row is
// written back unchanged, so no store sits under a per-row branch).
constexpr int UPJ = 1;   // 2 or 4 rows per thread measured no faster (189 -> 189, 192 us)

__device__ __forceinline__ double nanmax(double a, double b) { return a != a ? a : (b != b ? b : fmax(a, b)); }

// CHK (the mixed-precision form): the same update, and the Poisson residual
// r = f - L_h phi of the refined solution from the DPH values the update
// loads plus phi(i+1), phi(k+1) and phi(J+1) (wall rows: the Neumann rule of
// the solve), reduced per tile to part[2 tile] = max|r|, part[2 tile + 1] =
// max|f| (NaN propagates); k_max2_tiles then reduces the tiles.
// alpha dt max|r| is the projected max|div u|.
struct ResidCoef {
    const double *amph, *acph, *apph;
    double* part;
};

template <bool CHK>
__global__ void __launch_bounds__(256) k_update(Geo g, double cu, double cw, double cv, double be,
                                                const double* __restrict__ DPH, const double* __restrict__ F,
                                                double* __restrict__ Q, double* __restrict__ PR, ResidCoef rc) {
    int it, jg, kt, tile;
    const int M = g.M;
    if (!row_tile(g.L, g.N, (M + UPJ - 1) / UPJ, &it, &jg, &kt, &tile)) return;
    const int L = g.L, N = g.N;
    const int i0 = it * 64 + (threadIdx.x & 63), k0 = kt * 4 + (threadIdx.x >> 6);
    const bool valid = i0 < L && k0 < N;
    if (!CHK && !valid) return;
    // CHK: lanes outside the grid evaluate a valid cell, store nothing and
    // join the block's reduction with zeros
    const int i = min(i0, L - 1), k = min(k0, N - 1);
    const int LP = L * (M + 2), CS = LP * N;
    const int im = i == 0 ? L - 1 : -1, km = k == 0 ? (N - 1) * LP : -LP;
    const int ip = i == L - 1 ? 1 - L : 1, kp = k == N - 1 ? (1 - N) * LP : LP;
    const int j0 = jg * UPJ;
    const int g0 = i + L * (j0 + 1) + LP * k;  // (i, J0, k); J0 - 1 >= 0 is always in the array
    double ph[UPJ + 2], dxm[UPJ], dkm[UPJ], dxp[UPJ], dkp[UPJ], qu[UPJ], qv[UPJ], qw[UPJ], f[UPJ], pr[UPJ];
    ph[0] = DPH[g0 - L];
#pragma unroll
    for (int r = 0; r < UPJ; ++r) {
        const int jc = min(j0 + r, M - 1);
        const int go = g0 + L * (jc - j0);
        ph[r + 1] = DPH[go];
        dxm[r] = DPH[go + im];
        dkm[r] = DPH[go + km];
        if (CHK) {
            dxp[r] = DPH[go + ip];
            dkp[r] = DPH[go + kp];
        }
        qu[r] = Q[go];
        qv[r] = Q[go + CS];
        qw[r] = Q[go + 2 * CS];
        f[r] = F[i + L * (jc + M * k)];
        pr[r] = PR[go];
    }
    if (CHK) ph[UPJ + 1] = DPH[g0 + L * (min(j0 + UPJ, M) - j0)];
    double mr = 0.0, mf = 0.0;
#pragma unroll
    for (int r = 0; r < UPJ; ++r) {
        const int j = j0 + r;
        if (j >= M) break;
        if (CHK) {
            const double pc = ph[r + 1];
            const double pjp = j == M - 1 ? pc : ph[r + 2];
            const double pjm = j == 0 ? pc : ph[r];
            const double lx = (dxp[r] - 2.0 * pc + dxm[r]) * g.dxqi;
            const double ly = pjp * ldc(rc.apph, j) + pc * ldc(rc.acph, j) + pjm * ldc(rc.amph, j);
            const double lz = (dkp[r] - 2.0 * pc + dkm[r]) * g.dzqi;
            if (valid) {
                mr = nanmax(mr, fabs(f[r] - (lx + ly + lz)));
                mf = nanmax(mf, fabs(f[r]));
            }
            if (!valid) continue;
        }
        const int go = g0 + L * r;
        const double cvj = cv * ldc(g.dyci, j) / 1.0;
        Q[go] = qu[r] - (ph[r + 1] - dxm[r]) * cu;
        Q[go + 2 * CS] = qw[r] - (ph[r + 1] - dkm[r]) * cw;
        Q[go + CS] = j >= 1 ? qv[r] - (ph[r + 1] - ph[r]) * cvj : qv[r];
        const double p = pr[r] + ph[r + 1] - be * f[r];
        PR[go] = p;
        if (j == 0) PR[go - L] = p;
        if (j == M - 1) PR[go + L] = p;
    }
    if (CHK) {
        __shared__ double red[2][4];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            mr = nanmax(mr, __shfl_xor(mr, o));
            mf = nanmax(mf, __shfl_xor(mf, o));
        }
        if ((threadIdx.x & 63) == 0) {
            red[0][threadIdx.x >> 6] = mr;
            red[1][threadIdx.x >> 6] = mf;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double a = red[0][0], b = red[1][0];
#pragma unroll
            for (int w = 1; w < 4; ++w) {
                a = nanmax(a, red[0][w]);
                b = nanmax(b, red[1][w]);
            }
            rc.part[2 * tile] = a;
            rc.part[2 * tile + 1] = b;
        }
    }
}

// max |div u| (DIVGCK_tg, src/DIVGCK.f90:33-46) and max CFL rate (CFL_tg,
// src/CFL.f90:39-53) in one sweep: part[2b] = div, part[2b+1] = cfl.
__global__ void __launch_bounds__(256) k_divgck_cfl(Geo g, const double* __restrict__ Q, double* __restrict__ part) {
    __shared__ double red[2][256];
    const size_t ncell = (size_t)g.L * g.M * g.N;
    double md = 0.0, mc = 0.0;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < ncell; t += (size_t)gridDim.x * blockDim.x) {
        const int i = t % g.L;
        const int j = (t / g.L) % g.M;
        const int k = t / ((size_t)g.L * g.M);
        const int ip = i == g.L - 1 ? 0 : i + 1, kp = k == g.N - 1 ? 0 : k + 1;
        const int J = j + 1;
        const double u0 = Q[qidx(g, 0, i, J, k)], u1 = Q[qidx(g, 0, ip, J, k)];
        const double v0 = Q[qidx(g, 1, i, J, k)], v1 = Q[qidx(g, 1, i, J + 1, k)];
        const double w0 = Q[qidx(g, 2, i, J, k)], w1 = Q[qidx(g, 2, i, J, kp)];
        const double D = (u1 - u0) * g.dxi + (v1 - v0) * g.dyfi[j] * 1.0 + (w1 - w0) * g.dzi * 1.0;
        const double C = (fabs((u0 + u1) * g.dxi) + fabs((v0 + v1) * g.dyfi[j] * 1.0) + fabs((w0 + w1) * g.dzi * 1.0)) * 0.5;
        // NaN-propagating maxima: a NaN anywhere makes the result NaN
        md = nanmax(md, fabs(D));
        mc = nanmax(mc, C);
    }
    red[0][threadIdx.x] = md;
    red[1][threadIdx.x] = mc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            red[0][threadIdx.x] = nanmax(red[0][threadIdx.x], red[0][threadIdx.x + s]);
            red[1][threadIdx.x] = nanmax(red[1][threadIdx.x], red[1][threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = red[0][0];
        part[2 * blockIdx.x + 1] = red[1][0];
    }
}

// xz-mean statistics of PP_MEAN_ZX_FLOW_TG (src/processing/PP_MEAN_ZX.f90:
// 104-237), first and second moments: one block per row j, out[13 j + m] as the
// C ABI documents (cell-centred u, v, w sums, p, their products, u p).
__global__ void __launch_bounds__(256) k_mean_xz(Geo g, const double* __restrict__ Q, const double* __restrict__ PR,
                                                 double* __restrict__ out) {
    __shared__ double red[13][256];
    const int L = g.L, N = g.N, j = blockIdx.x, J = j + 1;
    const int LP = L * (g.M + 2), CS = LP * N;
    double acc[13];
#pragma unroll
    for (int m = 0; m < 13; ++m) acc[m] = 0.0;
    for (int e = threadIdx.x; e < L * N; e += blockDim.x) {
        const int k = e / L, i = e - k * L;
        const int ip = i == L - 1 ? 0 : i + 1, kp = k == N - 1 ? 0 : k + 1;
        const int o = i + L * J + LP * k;
        double U[3];
        U[0] = Q[o] + Q[o - i + ip];
        U[1] = Q[CS + o] + Q[CS + o + L];
        U[2] = Q[2 * CS + o] + Q[2 * CS + o + LP * (kp - k)];
        const double P = PR[o];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            acc[m] += U[m];
            acc[10 + m] += U[m] * P;
        }
        acc[3] += P;
        int l = 4;
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int n = m; n < 3; ++n) acc[l++] += U[m] * U[n];
    }
#pragma unroll
    for (int m = 0; m < 13; ++m) red[m][threadIdx.x] = acc[m];
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
#pragma unroll
            for (int m = 0; m < 13; ++m) red[m][threadIdx.x] += red[m][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x < 13) {
        const int m = threadIdx.x;
        const double VL = 1.0 / (double)(L * N);
        const double c = m == 3 ? VL : (m >= 4 && m < 10) ? VL * (0.5 * 0.5) : VL * 0.5;
        out[13 * j + m] = red[m][0] * c;
    }
}

__global__ void k_reduce_max2_nan(const double* __restrict__ part, int nb, double* __restrict__ out) {
    __shared__ double red[2][256];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        a = nanmax(a, part[2 * i]);
        b = nanmax(b, part[2 * i + 1]);
    }
    red[0][threadIdx.x] = a;
    red[1][threadIdx.x] = b;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            red[0][threadIdx.x] = nanmax(red[0][threadIdx.x], red[0][threadIdx.x + s]);
            red[1][threadIdx.x] = nanmax(red[1][threadIdx.x], red[1][threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = red[0][0];
        out[1] = red[1][0];
    }
}

// The mixed substep's check: the maxima (NaN-propagating) of the nb pairs
// (max|r|, max|f|) of this substep to acc[1], acc[2], and acc[0] = the largest
// max|r| / bound of every substep since the host last took an error from it
// (a NaN in r or f makes it NaN), so that a verdict the host does not read in
// time (it runs ahead of the device) is kept for the next one it reads.
__global__ void k_check_acc(const double* __restrict__ part, int nb, double inv_bound, double* __restrict__ acc) {
    __shared__ double red[2][256];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        a = nanmax(a, part[2 * i]);
        b = nanmax(b, part[2 * i + 1]);
    }
    red[0][threadIdx.x] = a;
    red[1][threadIdx.x] = b;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            red[0][threadIdx.x] = nanmax(red[0][threadIdx.x], red[0][threadIdx.x + s]);
            red[1][threadIdx.x] = nanmax(red[1][threadIdx.x], red[1][threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double r = red[0][0], f = red[1][0];
        acc[0] = nanmax(acc[0], f != f ? f : r * inv_bound);
        acc[1] = r;
        acc[2] = f;
    }
}

// Maxima (NaN-propagating) of n pairs part[2t], part[2t+1]: every block of
// 256 threads reduces a contiguous range of pairs, four independent loads in
// flight per thread, to out2[2 blockIdx.x + 0/1] (a fixed order: the result
// does not depend on timing); k_reduce_max2_nan then reduces the blocks.
__global__ void __launch_bounds__(256) k_max2_tiles(const double* __restrict__ part, int n, double* __restrict__ out2) {
    __shared__ double red[2][256];
    const int per = (n + gridDim.x - 1) / gridDim.x;
    const int t0 = blockIdx.x * per, t1 = min(n, t0 + per);
    double a[4] = {0.0, 0.0, 0.0, 0.0}, b[4] = {0.0, 0.0, 0.0, 0.0};
    int t = t0 + threadIdx.x;
    for (; t + 768 < t1; t += 1024) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            a[q] = nanmax(a[q], part[2 * (t + 256 * q)]);
            b[q] = nanmax(b[q], part[2 * (t + 256 * q) + 1]);
        }
    }
    for (; t < t1; t += 256) {
        a[0] = nanmax(a[0], part[2 * t]);
        b[0] = nanmax(b[0], part[2 * t + 1]);
    }
    red[0][threadIdx.x] = nanmax(nanmax(a[0], a[1]), nanmax(a[2], a[3]));
    red[1][threadIdx.x] = nanmax(nanmax(b[0], b[1]), nanmax(b[2], b[3]));
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            red[0][threadIdx.x] = nanmax(red[0][threadIdx.x], red[0][threadIdx.x + s]);
            red[1][threadIdx.x] = nanmax(red[1][threadIdx.x], red[1][threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out2[2 * blockIdx.x] = red[0][0];
        out2[2 * blockIdx.x + 1] = red[1][0];
    }
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------
bool tdma_seg_config(int n, bool cyclic, int* MS, int* P, int* LPB, int maxthr, bool lpb_small = false);
int poisson_solve(chs_handle* h, const double* f, double* dph, int ghost, double abs_tol);
int poisson_solve_unchecked(chs_handle* h, const double* f, double* dph, int ghost);

// The verdict of the last mixed-precision substep's convergence check, once
// its values have reached the host: wait = false leaves it pending while the
// GPU is still busy.  A NaN or a residual above the tolerance is reported by
// the call that reads it (CHS_E_NOTCONV); the state is then not to be trusted.
// device scalars of the mixed substep's check: the accumulated max|r| /
// bound, then max|r| and max|f| of the latest checked substep
constexpr int CHK_SLOT = 10;

int mixed_check(chs_handle* h, bool wait) {
    if (!h->chk_pending) return CHS_OK;
    if (wait) {
        CHS_HIP(hipEventSynchronize(h->chk_ev));
    } else {
        const hipError_t q = hipEventQuery(h->chk_ev);
        if (q == hipErrorNotReady) return CHS_OK;
        CHS_HIP(q);
    }
    h->chk_pending = false;
    // chk_host: the largest max|r| / bound since the last reported error,
    // then max|r| and max|f| of the latest checked substep
    const double worst = h->chk_host[0], r = h->chk_host[1], fm = h->chk_host[2];
    h->last_refine_resid = fm > 0.0 ? r / fm : 0.0;
    if (worst != worst || worst > 1.0) {
        // reported once: the accumulated maximum starts again from zero
        CHS_LAUNCH(k_set_scalar, dim3(1), dim3(64), 0, h->stream, h->scalars + CHK_SLOT, 0.0);
        if (worst != worst) {
            set_error("substep: NaN in the mixed-precision Poisson residual of an earlier substep");
        } else {
            set_error("substep: the mixed-precision Poisson solve of an earlier substep left max|f - L phi| at " +
                      std::to_string(worst) + " times its bound 0.5e-10 / (alpha dt)");
        }
        return CHS_E_NOTCONV;
    }
    return CHS_OK;
}

static Geo geo_of(const chs_handle* h) {
    Geo g;
    g.L = h->L; g.M = h->M; g.N = h->N;
    g.mc = h->nflow - 1;
    g.dxi = h->dxi; g.dzi = h->dzi; g.dxqi = h->dxqi; g.dzqi = h->dzqi;
    g.dyfi = h->dyfi; g.dyci = h->dyci; g.wff = h->wff; g.wfb = h->wfb;
    g.amvr = h->amvr; g.acvr = h->acvr; g.apvr = h->apvr;
    return g;
}

static const double TGAM[4] = {1.0, 8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
static const double TROH[4] = {0.0, 0.0, -17.0 / 60.0, -5.0 / 12.0};

using ld = long double;

// Thomas solve in extended precision (a[0] and c[n-1] ignored).
static void thomas_ld(int n, const ld* a, const ld* b, const ld* c, ld* x) {
    std::vector<ld> h(n), gg(n);
    h[0] = c[0] / b[0];
    gg[0] = x[0] / b[0];
    for (int i = 1; i < n; ++i) {
        const ld den = b[i] - a[i] * h[i - 1];
        h[i] = c[i] / den;
        gg[i] = (x[i] - a[i] * gg[i - 1]) / den;
    }
    x[n - 1] = gg[n - 1];
    for (int i = n - 2; i >= 0; --i) x[i] = gg[i] - h[i] * x[i + 1];
}

// Smallest K whose dropped tail sum_{|m|>K} |row_m| is <= tol * sum |row|
static int trunc_width(const std::vector<ld>& row, int centre, bool cyclic) {
    const int n = (int)row.size();
    ld tot = 0;
    for (ld v : row) tot += fabsl(v);
    for (int K = 0; K <= KMAX; ++K) {
        ld tail = 0;
        for (int m = 0; m < n; ++m) {
            int d = m - centre;
            if (cyclic) {
                d = ((d % n) + n) % n;
                if (d > n / 2) d -= n;
            }
            if (d > K || d < -K) tail += fabsl(row[m]);
        }
        if (tail <= 1e-17L * tot) return K;
    }
    return KMAX + 1;
}

// Inverse of the constant cyclic system (a, b, c) of size n, row 0: the
// periodic stencil.  Returns K (> KMAX when too wide) and fills w[0..2K].
static int cyclic_stencil(int n, double a, double b, double c, std::vector<double>& w) {
    // A v = e_0 by the Sherman-Morrison form of the cyclic system
    std::vector<ld> A(n, (ld)a), B(n, (ld)b), Cc(n, (ld)c), x(n, 0), u(n, 0);
    // cyclic -> tridiagonal with corrections: gamma = -b0
    const ld gam = -B[0];
    std::vector<ld> bb(B);
    bb[0] = B[0] - gam;
    bb[n - 1] = B[n - 1] - (ld)c * (ld)a / gam;
    x[0] = 1;
    thomas_ld(n, A.data(), bb.data(), Cc.data(), x.data());
    u[0] = gam;
    u[n - 1] = (ld)c;
    thomas_ld(n, A.data(), bb.data(), Cc.data(), u.data());
    const ld fact = (x[0] + (ld)a * x[n - 1] / gam) / (1 + u[0] + (ld)a * u[n - 1] / gam);
    std::vector<ld> v(n);
    for (int i = 0; i < n; ++i) v[i] = x[i] - fact * u[i];
    // symmetric circulant (a == c): row 0 of the inverse = column 0
    const int K = trunc_width(v, 0, true);
    if (K > KMAX) return K;
    w.assign(2 * K + 1, 0.0);
    for (int m = -K; m <= K; ++m) w[m + K] = (double)v[((m % n) + n) % n];
    return K;
}

struct CnStencils {
    double facoe = -1.0;
    int Kx = -1, Kz = -1, Ky = -1, KW = 4, KWz = 4, KWx = 4;
    bool ok = false;
    double* dev = nullptr;   // wx[2KWx+1] | wz[2KWz+1] | wy[3][M][2KW+1], zero-padded windows
    size_t ox = 0, oz = 0, oy = 0;
    std::vector<double> host;   // the uploaded values (kept alive for the stream-ordered copy)
    hipEvent_t used = nullptr;  // recorded after the last substep that read dev
};

// The stencils of the last few FACOE values of one handle.  Entries are
// evicted oldest first; an evicted entry's device memory is freed once its
// last-use event has completed (never while a queued substep may read it).
struct CnCache {
    std::list<CnStencils> live;   // list: entries keep their addresses
    std::vector<std::pair<double*, hipEvent_t>> retired;
};

static void reap_retired(CnCache& c, bool wait) {
    for (size_t i = 0; i < c.retired.size();) {
        auto& r = c.retired[i];
        const bool done = r.second == nullptr || (wait ? hipEventSynchronize(r.second) == hipSuccess
                                                       : hipEventQuery(r.second) == hipSuccess);
        if (!done) {
            ++i;
            continue;
        }
        if (r.first) (void)hipFree(r.first);
        if (r.second) (void)hipEventDestroy(r.second);
        c.retired.erase(c.retired.begin() + i);
    }
}

void release_stencils(chs_handle* h) {
    std::lock_guard<std::mutex> lk(h->cn_mu);
    CnCache* c = h->cn_cache;
    if (!c) return;
    for (auto& e : c->live) c->retired.push_back({e.dev, e.used});
    c->live.clear();
    reap_retired(*c, true);
    delete c;
    h->cn_cache = nullptr;
}

// Stencils for FACOE (the caller holds h->cn_mu): from the handle's cache, or
// computed on the host from the host copies of AMVR/ACVR/APVR and uploaded on
// the handle's stream.
static int cn_stencils_locked(chs_handle* h, double facoe, CnStencils** out) {
    if (!h->cn_cache) h->cn_cache = new CnCache();
    CnCache& cache = *h->cn_cache;
    for (auto& c : cache.live)
        if (c.facoe == facoe) {
            *out = &c;
            return CHS_OK;
        }
    reap_retired(cache, false);
    CnStencils st;
    st.facoe = facoe;
    const int M = h->M;
    const double COE = -facoe * h->dxqi;
    const double xa = COE, xc = COE, xb = 1.0 - xc - xa;
    const double RMC2 = 1.0 * (-facoe * h->dzqi);
    const double za = RMC2, zc = RMC2, zb = 1.0 - zc - za;
    std::vector<double> wx, wz;
    st.Kx = cyclic_stencil(h->L, xa, xb, xc, wx);
    st.Kz = cyclic_stencil(h->N, za, zb, zc, wz);
    // y: full inverse of each component's non-cyclic system (rows j = 0..M-1)
    std::vector<std::vector<ld>> inv(3 * (size_t)M);
    const std::vector<double>&amvr = h->h_amvr, &acvr = h->h_acvr, &apvr = h->h_apvr;
    int Ky = 0;
    for (int c = 0; c < 3; ++c) {
        std::vector<ld> a(M), b(M), cc(M), col(M);
        for (int j = 0; j < M; ++j) {
            const int o = c * (M + 1) + j;
            b[j] = (ld)(1.0 - facoe * acvr[o]);
            cc[j] = (ld)(-facoe * apvr[o]);
            a[j] = (ld)(-facoe * amvr[o]);
        }
        a[0] = 0;
        cc[M - 1] = 0;
        for (int j = 0; j < M; ++j) inv[c * M + j].assign(M, 0);
        for (int m = 0; m < M; ++m) {  // column m of the inverse
            std::fill(col.begin(), col.end(), (ld)0);
            col[m] = 1;
            thomas_ld(M, a.data(), b.data(), cc.data(), col.data());
            for (int j = 0; j < M; ++j) inv[c * M + j][m] = col[j];
        }
        for (int j = 0; j < M; ++j) Ky = std::max(Ky, trunc_width(inv[c * M + j], j, false));
    }
    st.Ky = Ky;
    auto window = [](int K) { return K <= 4 ? 4 : K <= 8 ? 8 : K <= 12 ? 12 : 16; };
    st.KW = window(Ky);
    st.KWz = window(st.Kz);
    st.KWx = window(st.Kx);
    // a periodic window must not wrap onto itself: 2*KW + 1 <= the line length
    st.ok = st.Kx <= KMAX && st.Kz <= KMAX && st.Ky <= KMAX && 2 * st.KWx + 1 <= h->L &&
            2 * st.KWz + 1 <= h->N;
    if (st.ok) {
        const int W = 2 * st.KW + 1;
        Ky = st.KW;
        const size_t nwx = 2 * (size_t)st.KWx + 1, nwz = 2 * (size_t)st.KWz + 1;
        std::vector<double>& buf = st.host;
        buf.assign(nwx + nwz + 3 * (size_t)M * W, 0.0);
        st.ox = 0;
        st.oz = nwx;
        st.oy = nwx + nwz;
        std::copy(wx.begin(), wx.end(), buf.begin() + (st.KWx - st.Kx));
        std::copy(wz.begin(), wz.end(), buf.begin() + st.oz + (st.KWz - st.Kz));
        for (int c = 0; c < 3; ++c)
            for (int j = 0; j < M; ++j)
                for (int m = -Ky; m <= Ky; ++m) {
                    const int jm = j + m;
                    if (jm >= 0 && jm < M && m >= -st.Ky && m <= st.Ky)
                        buf[st.oy + ((size_t)c * M + j) * W + (m + Ky)] = (double)inv[c * M + j][jm];
                }
        CHS_HIP(hipMalloc((void**)&st.dev, buf.size() * sizeof(double)));
        hipError_t e = hipEventCreateWithFlags(&st.used, hipEventDisableTiming);
        if (e == hipSuccess)
            e = hipMemcpyAsync(st.dev, buf.data(), buf.size() * sizeof(double), hipMemcpyHostToDevice, h->stream);
        if (e != hipSuccess) {
            (void)hipStreamSynchronize(h->stream);
            (void)hipFree(st.dev);
            if (st.used) (void)hipEventDestroy(st.used);
            set_error(std::string("CN stencil upload: ") + hipGetErrorString(e));
            return CHS_E_HIP;
        }
    }
    if (cache.live.size() >= 8) {
        CnStencils& old = cache.live.front();
        cache.retired.push_back({old.dev, old.used});
        cache.live.pop_front();
    }
    cache.live.push_back(std::move(st));
    *out = &cache.live.back();
    return CHS_OK;
}

// The x solve of k_rhs_x uses L/XMS segments of one block (XLPB*L/XMS
// threads); the z rows come in groups of KR.  L needs no power of two: the
// RHS kernels split thread indices by a division when it is not one.
bool substep_supported(const chs_handle* h) {
    const int P = h->L / XMS;
    return h->L % XMS == 0 && P >= 2 && XLPB * P <= 1024 && h->N % KR == 0;
}

// last substep's CN solve path: 1 = truncated-inverse stencils, 0 = segmented Thomas
int cn_last_path(chs_handle* h, double facoe, int* kx, int* kz, int* ky) {
    std::lock_guard<std::mutex> lk(h->cn_mu);
    CnStencils* cn = nullptr;
    CHS_TRY(cn_stencils_locked(h, facoe, &cn));
    *kx = cn->Kx;
    *kz = cn->Kz;
    *ky = cn->Ky;
    return cn->ok && h->cn_mode != 1 && h->N % ZKC == 0 && (h->M + YJC - 1) / YJC <= 8 && h->M > 16 ? 1 : 0;
}

// measured at 64^3 / 128^3 / 256x192x256 (FP64 substep): 2048 waves
// 94.8 -> 75.1 / 259.5 -> 250.7 / 1071 -> 1073 us; 4096 the same, 8192 slower
// on the large grid (1086 us)
constexpr long WIN_WAVES = 2048;
// Chunk lengths of the y and z windows.  A window marches the rows of its
// chunk one after the other, so on a grid too small to give every SIMD a few
// waves the march is latency-bound: halve the chunk (more waves, each
// re-reading the window's 2KW rows once more) until the launch has
// WIN_WAVES waves or the chunk is 8 rows.
static int ywin_chunk(const chs_handle* h) {
    const long per = 3L * h->N * ((h->L + 63) / 64);   // waves per chunk of every column
    int jc = YJC;
    while (jc > 8 && per * ((h->M + jc - 1) / jc) < WIN_WAVES) jc /= 2;
    return jc;
}
// Rows per y chunk of the marching RHS kernel: about one block per CU, or
// (FP64 ring: fill) one for each block the LDS lets a CU hold (the ring
// takes most of a CU's LDS at L = 256: one; 128: two); chunks of >= 4 rows,
// none empty.  Measured (substep): FP64 at 128^3, 256 blocks 259.5 us, 512
// blocks 245 us; the FP32 ring at 128^3 302 -> 304 us and at 256x192x256
// 1069 -> 1092 us with 512 blocks, so it keeps one block per CU.
static int xm_chunk(const chs_handle* h, size_t smem, bool fill) {
    return march_chunk_rows(h->M, h->N / KRS, smem, h->lds_max, fill);
}
// zkc divides N and is a multiple of WU (N is a multiple of ZKC on this path)
static int zwin_chunk(const chs_handle* h) {
    const long per = 3L * h->M * ((h->L + 63) / 64);
    int zkc = h->N % (2 * ZKC) == 0 ? 2 * ZKC : ZKC;
    while (zkc > 8 && per * (h->N / zkc) < WIN_WAVES) zkc /= 2;
    return zkc;
}

int rk3_substep(chs_handle* h, int ns, double dt) {
    CHS_TRY(mixed_check(h, false));
    const Geo g = geo_of(h);
    const hipStream_t s = h->stream;
    RkCoef rk;
    rk.tgam = TGAM[ns];
    rk.troh = TROH[ns];
    rk.talp = TGAM[ns] + TROH[ns];
    rk.dt = dt;
    const double facoe = 0.5 * rk.talp * dt * h->cvisc;  // FACOE_TG (src/MOMFA_tg.f90:29)
    hipEvent_t ev;
    // the handle's stencil cache stays locked until the substep's kernels
    // that read the stencils are queued and their last-use event recorded
    std::unique_lock<std::mutex> lk(h->cn_mu);
    CnStencils* cn = nullptr;
    CHS_TRY(cn_stencils_locked(h, facoe, &cn));
    // stencil path needs: N in whole z tiles, at most 4 y chunks of YJC rows
    // (blocks of 64*chunks threads), and chunks longer than the window
    const bool use_st = cn->ok && h->rhs2 != nullptr && h->cn_mode != 1 && h->N % ZKC == 0 &&
                        (h->M + YJC - 1) / YJC <= 8 && h->M > 16;
    bool f32 = false;
    // 1. RHS + x solve
    {
        const double COE = -facoe * h->dxqi;
        const double xa = COE, xc = COE, xb = 1.0 - xc - xa;
        // log2 L when L is a power of two (index shifts in the RHS kernels), else -1
        const int lgL = (h->L & (h->L - 1)) == 0 ? __builtin_ctz((unsigned)h->L) : -1;
        // 4 k rows per row-tile block: measured at 256x192x256, 8 rows (fewer
        // k halo rows per output) ran 1.3% slower (0.452 vs 0.446 ms)
        const int krs = KRS;
        const int kr = use_st ? krs : KR;
        const int nblk = h->M * (h->N / kr);
        const dim3 grx(((nblk + 7) / 8) * 8);
        // y-marching form: XM_CPT cells of each row per thread (whole waves)
        // and its LDS ring on this device
        // (cn_mode 3: wherever it fits; the mixed form marches for its FP32
        // stencils; FP64 otherwise where the chunks have MARCH_MIN rows)
        const bool use_xm = use_st && (h->cn_mode == 0 || h->cn_mode == 3) && (KRS / XM_CPT) * h->L % 64 == 0 &&
                            (KRS / XM_CPT) * h->L <= (XM_CPT == 1 ? 1024 : 512) &&
                            rhs_xm_smem(h->L, cn->KWx) <= (size_t)h->lds_max &&
                            (h->cn_mode == 3 || h->precision == CHS_MIXED ||
                             xm_chunk(h, rhs_xm_smem(h->L, cn->KWx), true) >= MARCH_MIN);
        // the mixed form's FP32 stencils: RHS, the x solve and the z solve in
        // FP32 (RHS and RHS2 stored as float in the first half of the FP64
        // buffers), CONVH0 kept in FP32 (h0f); the y elimination stays FP64
        f32 = h->precision == CHS_MIXED && use_xm;
        CHS_TRY(convh0_store(h, f32));
        h->timer.begin("rhs_x", s, &ev);
        if (f32) {
            const size_t sm = rhs_xm_smem(h->L, cn->KWx, sizeof(float));
            const int jc = xm_chunk(h, sm, false);
            const int nkb = h->N / KRS, nyc = (h->M + jc - 1) / jc;
            const dim3 gm(8 * ((nkb * nyc + 7) / 8));
            const double* wx = cn->dev + cn->ox;
            float* rf = reinterpret_cast<float*>(h->rhs);
#define CHS_RHSXM32(KX)                                                                                          \
    do {                                                                                                         \
        (void)hipFuncSetAttribute((const void*)k_rhs_xm<KX, XM_PD, XM_CPT, float>,                                \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);                          \
        CHS_LAUNCH((k_rhs_xm<KX, XM_PD, XM_CPT, float>), gm, dim3((KRS / XM_CPT) * h->L), sm, s, g, rk, h->cvisc, \
                   h->q, h->pr, h->h0f, rf, h->partials, wx, lgL, nyc, jc);                                      \
    } while (0)
            switch (cn->KWx) {
                case 4: CHS_RHSXM32(4); break;
                case 8: CHS_RHSXM32(8); break;
                case 12: CHS_RHSXM32(12); break;
                default: CHS_RHSXM32(16); break;
            }
#undef CHS_RHSXM32
        } else if (use_xm) {
            const size_t sm = rhs_xm_smem(h->L, cn->KWx);
            const int jc = xm_chunk(h, sm, true);
            const int nkb = h->N / KRS, nyc = (h->M + jc - 1) / jc;
            const dim3 gm(8 * ((nkb * nyc + 7) / 8));
            const double* wx = cn->dev + cn->ox;
#define CHS_RHSXM(KX)                                                                                  \
    do {                                                                                               \
        (void)hipFuncSetAttribute((const void*)k_rhs_xm<KX, XM_PD, XM_CPT>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                                  (int)sm);                                                            \
        CHS_LAUNCH((k_rhs_xm<KX, XM_PD, XM_CPT>), gm, dim3((KRS / XM_CPT) * h->L), sm, s, g, rk, h->cvisc, h->q, h->pr, h->convh0,  \
                   h->rhs, h->partials, wx, lgL, nyc, jc);                                             \
    } while (0)
            switch (cn->KWx) {
                case 4: CHS_RHSXM(4); break;
                case 8: CHS_RHSXM(8); break;
                case 12: CHS_RHSXM(12); break;
                default: CHS_RHSXM(16); break;
            }
#undef CHS_RHSXM
        } else if (use_st) {
            const size_t sm = sizeof(double) * (3 * (size_t)krs * (size_t)(h->L + 2 * cn->KWx) + 4);
            const double* wx = cn->dev + cn->ox;
#define CHS_RHSXW(KX, KR) CHS_LAUNCH((k_rhs_xw<KX, KR>), grx, dim3(256), sm, s, g, rk, h->cvisc, h->q, h->pr, h->convh0, h->rhs, h->partials, wx, lgL)
            switch (cn->KWx) {
                case 4: CHS_RHSXW(4, KRS); break;
                case 8: CHS_RHSXW(8, KRS); break;
                case 12: CHS_RHSXW(12, KRS); break;
                default: CHS_RHSXW(16, KRS); break;
            }
#undef CHS_RHSXW
        } else {
            const int P = h->L / XMS;
            const int nthr = XLPB * P;
            const size_t sm = std::max(sizeof(double) * (3 * KR * (size_t)(h->L + 1) + (size_t)nthr),
                                       seg_smem_bytes<double>(XLPB, P));
            CHS_LAUNCH(k_rhs_x, grx, dim3(nthr), sm, s, g, rk, h->cvisc, h->q, h->pr, h->convh0, h->rhs, h->partials,
                       xa, xb, xc, lgL);
        }
        h->timer.end(ev, "rhs_x", s);
        h->timer.begin("massflux", s, &ev);
        if (h->flowdrvtp == 1)
            CHS_LAUNCH(k_massflux, dim3(1), dim3(1024), 0, s, h->partials, nblk, h->volm, h->scalars);
        else   // constant pressure gradient (RHS_CvLpGpS_tg.f90:200-201), written by a kernel (no memset node)
            CHS_LAUNCH(k_set_scalar, dim3(1), dim3(64), 0, s, h->scalars, -0.50 * h->cfgv * (rk.talp * dt));
        h->timer.end(ev, "massflux", s);
        // 2. z solves
        const double RMC2 = 1.0 * (-facoe * h->dzqi);
        const double za = RMC2, zc = RMC2, zb = 1.0 - zc - za;
        const double xone = 1.0 / (xa + xb + xc);
        const int usedpg0 = 1;
        if (use_st) {
            // 128-row chunks where N allows (measured at 256x192x256: 139 -> 134 us)
            const int zkc = zwin_chunk(h);
            const dim3 gzw(3 * h->M * ((h->L + 63) / 64), (h->N / zkc + 3) / 4);
            const double* wz = cn->dev + cn->oz;
            h->timer.begin("z_solve", s, &ev);
            if (f32) {
                const float* rf = reinterpret_cast<const float*>(h->rhs);
                float* r2f = reinterpret_cast<float*>(h->rhs2);
                switch (cn->KWz) {
                    case 4: CHS_LAUNCH((k_z_window<4, float>), gzw, dim3(256), 0, s, g, rf, r2f, h->scalars, xone, usedpg0, wz, zkc); break;
                    case 8: CHS_LAUNCH((k_z_window<8, float>), gzw, dim3(256), 0, s, g, rf, r2f, h->scalars, xone, usedpg0, wz, zkc); break;
                    case 12: CHS_LAUNCH((k_z_window<12, float>), gzw, dim3(256), 0, s, g, rf, r2f, h->scalars, xone, usedpg0, wz, zkc); break;
                    default: CHS_LAUNCH((k_z_window<16, float>), gzw, dim3(256), 0, s, g, rf, r2f, h->scalars, xone, usedpg0, wz, zkc); break;
                }
            } else {
                switch (cn->KWz) {
                    case 4: CHS_LAUNCH(k_z_window<4>, gzw, dim3(256), 0, s, g, h->rhs, h->rhs2, h->scalars, xone, usedpg0, wz, zkc); break;
                    case 8: CHS_LAUNCH(k_z_window<8>, gzw, dim3(256), 0, s, g, h->rhs, h->rhs2, h->scalars, xone, usedpg0, wz, zkc); break;
                    case 12: CHS_LAUNCH(k_z_window<12>, gzw, dim3(256), 0, s, g, h->rhs, h->rhs2, h->scalars, xone, usedpg0, wz, zkc); break;
                    default: CHS_LAUNCH(k_z_window<16>, gzw, dim3(256), 0, s, g, h->rhs, h->rhs2, h->scalars, xone, usedpg0, wz, zkc); break;
                }
            }
            h->timer.end(ev, "z_solve", s);
        }
        int MS, Pz, LPB;
        if (!tdma_seg_config(h->N, true, &MS, &Pz, &LPB, 1024)) {
            set_error("substep: unsupported NCL3");
            return CHS_E_ARG;
        }
        const int nlines = 3 * h->L * h->M;
        const dim3 gz((nlines + LPB - 1) / LPB), bz(LPB * Pz);
        const size_t smz = seg_smem_bytes<double>(LPB, Pz);
        const int usedpg = 1;
        if (!use_st) {
        h->timer.begin("z_solve", s, &ev);
        switch (MS) {
            case 16: CHS_LAUNCH(k_z_solve<16>, gz, bz, smz, s, g, h->rhs, h->scalars, xone, za, zb, zc, LPB, Pz, usedpg); break;
            case 12: CHS_LAUNCH(k_z_solve<12>, gz, bz, smz, s, g, h->rhs, h->scalars, xone, za, zb, zc, LPB, Pz, usedpg); break;
            case 4: CHS_LAUNCH(k_z_solve<4>, gz, bz, smz, s, g, h->rhs, h->scalars, xone, za, zb, zc, LPB, Pz, usedpg); break;
            default: CHS_LAUNCH(k_z_solve<8>, gz, bz, smz, s, g, h->rhs, h->scalars, xone, za, zb, zc, LPB, Pz, usedpg); break;
        }
        h->timer.end(ev, "z_solve", s);
        }
    }
    // 3. y solves + Q update + BC_WALL_Q
    if (use_st) {
        const int yjc = ywin_chunk(h);
        const int nch = (h->M + yjc - 1) / yjc;
        const dim3 gy(h->N * ((h->L + 63) / 64), 3, nch), by(64);
        const double* wy = cn->dev + cn->oy;
        const size_t smw = sizeof(double) * (size_t)std::min(h->M, yjc) * (2 * cn->KW + 1);
        h->timer.begin("y_solve", s, &ev);
        if (f32) {
            const float* r2f = reinterpret_cast<const float*>(h->rhs2);
            switch (cn->KW) {
                case 4: CHS_LAUNCH((k_y_window<4, float>), gy, by, smw, s, g, r2f, h->q, wy, yjc); break;
                case 8: CHS_LAUNCH((k_y_window<8, float>), gy, by, smw, s, g, r2f, h->q, wy, yjc); break;
                case 12: CHS_LAUNCH((k_y_window<12, float>), gy, by, smw, s, g, r2f, h->q, wy, yjc); break;
                default: CHS_LAUNCH((k_y_window<16, float>), gy, by, smw, s, g, r2f, h->q, wy, yjc); break;
            }
        } else {
            switch (cn->KW) {
                case 4: CHS_LAUNCH(k_y_window<4>, gy, by, smw, s, g, h->rhs2, h->q, wy, yjc); break;
                case 8: CHS_LAUNCH(k_y_window<8>, gy, by, smw, s, g, h->rhs2, h->q, wy, yjc); break;
                case 12: CHS_LAUNCH(k_y_window<12>, gy, by, smw, s, g, h->rhs2, h->q, wy, yjc); break;
                default: CHS_LAUNCH(k_y_window<16>, gy, by, smw, s, g, h->rhs2, h->q, wy, yjc); break;
            }
        }
        h->timer.end(ev, "y_solve", s);
    } else {
        int MS, P, LPB;
        if (!tdma_seg_config(h->M, false, &MS, &P, &LPB, 768)) {
            set_error("substep: unsupported NCL2");
            return CHS_E_ARG;
        }
        lk.unlock();
        const int nlines = h->L * h->N;
        const dim3 gy((nlines + LPB - 1) / LPB), by(LPB * P);
        const size_t smy = seg_smem_bytes<double>(LPB, P);
        h->timer.begin("y_solve", s, &ev);
        switch (MS) {
            case 16: CHS_LAUNCH(k_y_solve<16>, gy, by, smy, s, g, facoe, h->rhs, h->q, LPB, P); break;
            case 12: CHS_LAUNCH(k_y_solve<12>, gy, by, smy, s, g, facoe, h->rhs, h->q, LPB, P); break;
            case 4: CHS_LAUNCH(k_y_solve<4>, gy, by, smy, s, g, facoe, h->rhs, h->q, LPB, P); break;
            default: CHS_LAUNCH(k_y_solve<8>, gy, by, smy, s, g, facoe, h->rhs, h->q, LPB, P); break;
        }
        h->timer.end(ev, "y_solve", s);
    }
    if (lk.owns_lock()) {
        if (cn->used) CHS_HIP(hipEventRecord(cn->used, s));
        lk.unlock();
    }
    // 4. DIVG
    h->timer.begin("divg", s, &ev);
    CHS_LAUNCH(k_divg, dim3(row_tile_blocks(h->L, h->N, (h->M + DVJ - 1) / DVJ)), dim3(256), 0, s, g,
               1.0 / (dt * rk.talp), h->q, h->rhsllphi);
    h->timer.end(ev, "divg", s);
    // 5. Poisson + BC_WALL_DPH.  FP64: one pass.  Mixed, NS >= 1: the FP32
    //    pass and one FP64-residual refinement (an FP32 pass leaves ~1e-7 of
    //    max|f|, one refinement ~1e-12); the projection below checks that the
    //    projected divergence alpha*dt*max|f - L phi| stays below half the
    //    1e-10 target, and the host reads the verdict without waiting.
    // NS = 0 starts from a field of arbitrary divergence: refine there with a
    // check before each further pass (poisson_solve), as many passes as needed
    const int ntile = (int)(((h->L + 63) / 64) * ((h->N + 3) / 4) * (long)((h->M + UPJ - 1) / UPJ));
    constexpr int NB2 = 256;   // first-stage blocks of the tile reduction
    const bool fused_chk = h->precision == CHS_MIXED && h->refine_max > 0 && ns != 0 &&
                           ntile + NB2 <= h->npart;
    if (fused_chk)
        CHS_TRY(poisson_solve_unchecked(h, h->rhsllphi, h->dph, 1));
    else
        CHS_TRY(poisson_solve(h, h->rhsllphi, h->dph, 1, 0.5e-10 / (rk.talp * dt)));
    // 6. projection + pressure update
    const ResidCoef rcf{h->amph, h->acph, h->apph, h->partials};
    const dim3 gu(row_tile_blocks(h->L, h->N, (h->M + UPJ - 1) / UPJ));
    h->timer.begin("update", s, &ev);
    if (fused_chk)
        CHS_LAUNCH(k_update<true>, gu, dim3(256), 0, s, g, dt * rk.talp * h->dxi, dt * rk.talp * h->dzi, dt * rk.talp,
                   0.5 * rk.talp * dt * h->cvisc, h->dph, h->rhsllphi, h->q, h->pr, rcf);
    else
        CHS_LAUNCH(k_update<false>, gu, dim3(256), 0, s, g, dt * rk.talp * h->dxi, dt * rk.talp * h->dzi, dt * rk.talp,
                   0.5 * rk.talp * dt * h->cvisc, h->dph, h->rhsllphi, h->q, h->pr, rcf);
    h->timer.end(ev, "update", s);
    if (fused_chk) {
        double* p2 = h->partials + 2 * (size_t)ntile;   // after the tile pairs
        CHS_LAUNCH(k_max2_tiles, dim3(NB2), dim3(256), 0, s, h->partials, ntile, p2);
        h->chk_tol = 0.5e-10 / (rk.talp * dt);
        CHS_LAUNCH(k_check_acc, dim3(1), dim3(256), 0, s, p2, NB2, 1.0 / h->chk_tol, h->scalars + CHK_SLOT);
        CHS_HIP(hipMemcpyAsync(h->chk_host, h->scalars + CHK_SLOT, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
        CHS_HIP(hipEventRecord(h->chk_ev, s));
        h->chk_pending = true;
    }
    return CHS_OK;
}

int divgck_cfl(chs_handle* h, double* div_out, double* cfl_out) {
    CHS_TRY(mixed_check(h, true));
    const Geo g = geo_of(h);
    const hipStream_t s = h->stream;
    const int nb = 1024;
    hipEvent_t ev;
    h->timer.begin("divgck_cfl", s, &ev);
    CHS_LAUNCH(k_divgck_cfl, dim3(nb), dim3(256), 0, s, g, h->q, h->partials);
    CHS_LAUNCH(k_reduce_max2_nan, dim3(1), dim3(256), 0, s, h->partials, nb, h->scalars + 1);
    h->timer.end(ev, "divgck_cfl", s);
    double host[2];
    CHS_HIP(hipMemcpyAsync(host, h->scalars + 1, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    CHS_HIP(hipStreamSynchronize(s));
    if (div_out) *div_out = host[0];
    if (cfl_out) *cfl_out = host[1];
    return CHS_OK;
}

int mean_xz(chs_handle* h, double* out_host) {
    const Geo g = geo_of(h);
    const hipStream_t s = h->stream;
    // kept on the device for the time means (chs_mean_t)
    CHS_LAUNCH(k_mean_xz, dim3(h->M), dim3(256), 0, s, g, h->q, h->pr, h->stats);
    CHS_HIP(hipMemcpyAsync(out_host, h->stats, 13 * (size_t)h->M * sizeof(double), hipMemcpyDeviceToHost, s));
    CHS_HIP(hipStreamSynchronize(s));
    return CHS_OK;
}

}  // namespace chs
