// Batched tridiagonal line solver for gfx950: segmented (partitioned) Thomas.
//
// The reference solves every line with one sequential Thomas sweep
// (src/TDMAIJJ_nonCYC.f90:67-123) or the two-vector cyclic variant
// (src/TDMAIJI_CYC.f90:67-157).  On a GPU one thread per line leaves most of
// the chip idle and needs O(n) storage per line, so here each line is split
// into P segments of MS rows, one thread per segment, its rows in registers:
//
//  1. each thread eliminates its interior rows 1..MS-2 with Thomas, writing
//     every interior unknown as x_l = alpha_l + beta_l x_0 + gamma_l x_{MS-1}
//     in terms of its own first and last unknowns;
//  2. substituting into the segment's first and last rows leaves a reduced
//     tridiagonal system of 2P unknowns (x_0, x_{MS-1} of every segment),
//     solved in LDS by one thread per line -- Thomas, or for cyclic lines the
//     reference's two-vector method;
//  3. each thread reconstructs its interior unknowns.
//
// A workgroup holds LPB lines x P segments; lane i of a wave owns line i, so a
// wave's loads of row j of 64 lines are one coalesced 512-byte access in the
// "IJJ" layout (lines adjacent in memory, rows strided).
#pragma once

#include "chs_common.hpp"

namespace chs {

// Step 1: local elimination of one segment (MS >= 3).
//   in : a, b, c (real), d (real or complex) for local rows 0..MS-1
//   out: d[1..MS-2] = alpha, be = beta, ga = gamma, and the two reduced rows
template <int MS, typename V>
__device__ __forceinline__ void seg_eliminate(const double (&a)[MS], const double (&b)[MS], const double (&c)[MS],
                                              V (&d)[MS], double (&be)[MS], double (&ga)[MS], double& A0,
                                              double& B0, double& C0, V& D0, double& A1, double& B1, double& C1,
                                              V& D1) {
    static_assert(MS >= 3, "segments need at least one interior row");
    double H[MS];
    {
        const double r = 1.0 / b[1];
        H[1] = c[1] * r;
        d[1] = vscale(d[1], r);
        be[1] = -a[1] * r;
    }
#pragma unroll
    for (int l = 2; l <= MS - 2; ++l) {
        const double r = 1.0 / (b[l] - a[l] * H[l - 1]);
        H[l] = c[l] * r;
        d[l] = vscale(vaxpy(d[l], a[l], d[l - 1]), r);
        be[l] = -(a[l] * be[l - 1]) * r;
    }
    // the last interior row couples to x_{MS-1}: that term becomes gamma
    ga[MS - 2] = -H[MS - 2];
#pragma unroll
    for (int l = MS - 3; l >= 1; --l) {
        d[l] = vaxpy(d[l], H[l], d[l + 1]);
        be[l] = be[l] - H[l] * be[l + 1];
        ga[l] = -H[l] * ga[l + 1];
    }
    A0 = a[0];
    B0 = b[0] + c[0] * be[1];
    C0 = c[0] * ga[1];
    D0 = vaxpy(d[0], c[0], d[1]);
    A1 = a[MS - 1] * be[MS - 2];
    B1 = b[MS - 1] + a[MS - 1] * ga[MS - 2];
    C1 = c[MS - 1];
    D1 = vaxpy(d[MS - 1], a[MS - 1], d[MS - 2]);
}

// Step 3: interior unknowns from the segment's end values.
template <int MS, typename V>
__device__ __forceinline__ void seg_reconstruct(V (&d)[MS], const double (&be)[MS], const double (&ga)[MS], V F, V E) {
#pragma unroll
    for (int l = 1; l <= MS - 2; ++l) d[l] = vfma(vfma(d[l], be[l], F), ga[l], E);
    d[0] = F;
    d[MS - 1] = E;
}

// Step 2a: Thomas on the reduced system; rows r at index r*st of each array.
// A[0] and C[R-1] are ignored.  Solution overwrites D.
template <typename V>
__device__ inline void reduced_thomas(const double* A, const double* B, double* C, V* D, int R, int st) {
    double h = C[0] / B[0];
    V g = vscale(D[0], 1.0 / B[0]);
    C[0] = h;
    D[0] = g;
    for (int r = 1; r < R; ++r) {
        const double a = A[r * st];
        const double pp = 1.0 / (B[r * st] - h * a);
        h = C[r * st] * pp;
        g = vscale(vaxpy(D[r * st], a, g), pp);
        C[r * st] = h;
        D[r * st] = g;
    }
    V x = g;
    for (int r = R - 2; r >= 0; --r) {
        x = vaxpy(D[r * st], C[r * st], x);
        D[r * st] = x;
    }
}

// Step 2b: cyclic reduced system (A[0] couples row 0 to row R-1, C[R-1] row
// R-1 to row 0), by the two-vector method of TDMAIJI_CYC.  The second vector
// G2 overwrites A: row i of A is dead once G2[i] is formed (A[n-1] is kept in
// a register).
template <typename V>
__device__ inline void reduced_cyclic(double* A, const double* B, double* C, V* D, int R, int st) {
    double* G2 = A;
    const int n = R;
    const double An1 = A[(n - 1) * st], Bn1 = B[(n - 1) * st], Cn1 = C[(n - 1) * st];
    const V Dn1 = D[(n - 1) * st];
    double h = C[0] / B[0];
    V g1 = vscale(D[0], 1.0 / B[0]);
    double g2 = -A[0] / B[0];
    C[0] = h;
    D[0] = g1;
    G2[0] = g2;
    for (int i = 1; i <= n - 3; ++i) {
        const double a = A[i * st];
        const double pp = 1.0 / (B[i * st] - h * a);
        h = C[i * st] * pp;
        g1 = vscale(vaxpy(D[i * st], a, g1), pp);
        g2 = -(g2 * a) * pp;
        C[i * st] = h;
        D[i * st] = g1;
        G2[i * st] = g2;
    }
    {
        const int i = n - 2;
        const double a = A[i * st];
        const double pp = 1.0 / (B[i * st] - h * a);
        g1 = vscale(vaxpy(D[i * st], a, g1), pp);
        g2 = (-C[i * st] - g2 * a) * pp;
        C[i * st] = 0.0;
        D[i * st] = g1;
        G2[i * st] = g2;
    }
    for (int i = n - 3; i >= 0; --i) {
        const double hh = C[i * st];
        D[i * st] = vaxpy(D[i * st], hh, D[(i + 1) * st]);
        G2[i * st] = G2[i * st] - hh * G2[(i + 1) * st];
    }
    // x_{n-1}
    const V num = vaxpy(vaxpy(Dn1, Cn1, D[0]), An1, D[(n - 2) * st]);
    const double den = Bn1 + Cn1 * G2[0] + An1 * G2[(n - 2) * st];
    const V xn = vscale(num, 1.0 / den);
    for (int i = n - 2; i >= 0; --i) D[i * st] = vfma(D[i * st], G2[i * st], xn);
    D[(n - 1) * st] = xn;
}

// Step 2c: non-cyclic reduced system by parallel cyclic reduction: every
// segment thread keeps its two rows in registers; each of the ceil(log2 R)
// steps eliminates the couplings at distance s from both neighbours, after
// which every row stands alone.  Replaces the serial sweep of one thread per
// line (2R dependent steps) by ~log2 R steps over all threads of the block.
// All threads must call it (barriers).  Returns this thread's two unknowns.
template <typename V>
__device__ __forceinline__ void reduced_pcr(double* sA, double* sB, double* sC, V* sD, int R, int LPB, int ll,
                                            int seg, V& x0, V& x1) {
    double a[2], b[2], c[2];
    V d[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int m = (2 * seg + q) * LPB + ll;
        a[q] = sA[m];
        b[q] = sB[m];
        c[q] = sC[m];
        d[q] = sD[m];
    }
    for (int st = 1; st < R; st <<= 1) {
        double na[2], nb[2], nc[2];
        V nd[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int r = 2 * seg + q;
            na[q] = 0.0;
            nc[q] = 0.0;
            nb[q] = b[q];
            nd[q] = d[q];
            if (r - st >= 0) {
                const int m = (r - st) * LPB + ll;
                const double al = -a[q] / sB[m];
                na[q] = al * sA[m];
                nb[q] = nb[q] + al * sC[m];
                nd[q] = vfma(nd[q], al, sD[m]);
            }
            if (r + st < R) {
                const int m = (r + st) * LPB + ll;
                const double ga = -c[q] / sB[m];
                nc[q] = ga * sC[m];
                nb[q] = nb[q] + ga * sA[m];
                nd[q] = vfma(nd[q], ga, sD[m]);
            }
        }
        __syncthreads();   // every thread has read the rows of this step
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int m = (2 * seg + q) * LPB + ll;
            a[q] = na[q];
            b[q] = nb[q];
            c[q] = nc[q];
            d[q] = nd[q];
            sA[m] = na[q];
            sB[m] = nb[q];
            sC[m] = nc[q];
            sD[m] = nd[q];
        }
        __syncthreads();
    }
    x0 = vscale(d[0], 1.0 / b[0]);
    x1 = vscale(d[1], 1.0 / b[1]);
}

// LDS bytes needed by seg_solve for LPB lines and P segments.
template <typename V>
__host__ __device__ constexpr size_t seg_smem_bytes(int LPB, int P) {
    return (size_t)2 * P * LPB * (3 * sizeof(double) + sizeof(V));
}

// Steps 1-3 for the block's lines.  Every thread of the block must call it
// (it contains barriers).  smem: seg_smem_bytes<V>(LPB, P) bytes.
// On entry a, b, c, d hold the thread's MS rows; on return d holds the solution.
template <int MS, typename V>
__device__ __forceinline__ void seg_solve(double (&a)[MS], double (&b)[MS], double (&c)[MS], V (&d)[MS], int LPB,
                                          int P, int ll, int seg, bool valid, bool cyclic, unsigned char* smem) {
    const int R = 2 * P;
    double* sA = reinterpret_cast<double*>(smem);
    double* sB = sA + (size_t)R * LPB;
    double* sC = sB + (size_t)R * LPB;
    V* sD = reinterpret_cast<V*>(sC + (size_t)R * LPB);
    double be[MS], ga[MS];
    double A0, B0, C0, A1, B1, C1;
    V D0, D1;
    seg_eliminate<MS, V>(a, b, c, d, be, ga, A0, B0, C0, D0, A1, B1, C1, D1);
    const int r0 = 2 * seg, r1 = 2 * seg + 1;
    sA[r0 * LPB + ll] = A0; sB[r0 * LPB + ll] = B0; sC[r0 * LPB + ll] = C0; sD[r0 * LPB + ll] = D0;
    sA[r1 * LPB + ll] = A1; sB[r1 * LPB + ll] = B1; sC[r1 * LPB + ll] = C1; sD[r1 * LPB + ll] = D1;
    __syncthreads();
    V F, E;
    if (!cyclic) {
        reduced_pcr<V>(sA, sB, sC, sD, R, LPB, ll, seg, F, E);
    } else {
        if (seg == 0 && valid) reduced_cyclic<V>(sA + ll, sB + ll, sC + ll, sD + ll, R, LPB);
        __syncthreads();
        F = sD[r0 * LPB + ll];
        E = sD[r1 * LPB + ll];
    }
    seg_reconstruct<MS, V>(d, be, ga, F, E);
}

}  // namespace chs
