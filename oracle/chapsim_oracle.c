/*
 * CPU oracle for the CHAPSim per-RK3-substep solver core.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C, FP64 restatement of the
 * reference Fortran routines named below, loop for loop, so that the HIP
 * library can be checked against it.  Only tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py may load it; the product path never
 * calls it and has no CPU fallback.
 *
 * Index conventions: arrays are Fortran column-major, x fastest.  Macros below
 * take the Fortran (1-based, ghost-aware) indices and map them to C offsets:
 *   Q_tg(NCL1, 0:M+1, NCL3, 3), PR/DPH(NCL1, 0:M+1, NCL3)
 *   RHS/RHSLLPHI/Qtmp(NCL1, M, NCL3), CONVH0(NCL1, M, NCL3, 3)
 *
 * Parity status of each part (how it is pinned) is listed in the test suite
 * headers; every function here has at least one independent check.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int L, M, N;            /* NCL1, NCL2, NCL3 */
    int icase;              /* 1 = channel (walls in y), 4 = 3-periodic box */
    double dx, dz, dxi, dzi, dxqi, dzqi, cvisc, volm;
    double *ynd;            /* [M+1]  YND(1..NND2)                         */
    double *ycc;            /* [M]    YCC(1..NCL2)                         */
    double *dyfi;           /* [M]    DYFI(1..NCL2)                        */
    double *dyci;           /* [M+1]  DYCI(1..NND2)                        */
    double *wff, *wfb;      /* [M+1]  YCL2ND_WFF/WFB(1..NND2)              */
    double *amph, *acph, *apph;   /* [M]                                   */
    double *amvr, *acvr, *apvr;   /* [3*(M+1)]  (JC, IDR) -> (IDR-1)*(M+1)+JC-1 */
    double *ak1;            /* [L]    AK1(1..L)                            */
    double *ak3;            /* [N]    AK3(1..N)                            */
} orc_grid;

#define ICHANL 1
#define IBOX3P 4

/* ---------------------------------------------------------------------------
 * Grid, metrics and coefficient set-up.
 * COORDJR (src/COORDJR.f90:1-235, channel branch with ISTR2=1),
 * AREA_INOUTLET (src/COORDJR.f90:355-399), MESH_YWEIGHT_FACTORS
 * (src/COORDJR.f90:427-470), CONSPARA (src/CONSPARA.f90:15-110),
 * TDMA_COEF (src/TDMA_COEF.f90:30-116, channel branch),
 * FFT99_ROOT (src/possion3d_FFT99.f90:124-182).
 * ------------------------------------------------------------------------- */
void orc_grid_setup(orc_grid *g, double hx, double hz, double str2, double ren)
{
    const int L = g->L, M = g->M, N = g->N, NND2 = M + 1;
    const double PI = 2.0 * asin(1.0);
    /* y nodes and cell centres: tanh stretching (src/COORDJR.f90:28-38) */
    for (int J = 1; J <= NND2; ++J) {
        double X2 = (double)(J - 1) / (double)(NND2 - 1);
        g->ynd[J - 1] = tanh(str2 * (X2 - 0.5)) / tanh(str2 * 0.5);
    }
    for (int J = 1; J <= M; ++J) {
        double X2 = ((double)(J - 1) + 0.5) / (double)(NND2 - 1);
        g->ycc[J - 1] = tanh(str2 * (X2 - 0.5)) / tanh(str2 * 0.5);
    }
    for (int J = 1; J <= M; ++J) g->dyfi[J - 1] = 1.0 / (g->ynd[J] - g->ynd[J - 1]);
    for (int J = 2; J <= NND2 - 1; ++J)
        g->dyci[J - 1] = 1.0 / ((g->ynd[J] - g->ynd[J - 2]) * 0.5);
    g->dyci[0] = 2.0 / (g->ynd[1] - g->ynd[0]);
    g->dyci[NND2 - 1] = 2.0 / (g->ynd[NND2 - 1] - g->ynd[M - 1]);
    /* y interpolation weights (src/COORDJR.f90:434-446) */
    for (int JC = 2; JC <= M; ++JC) {
        int JM = JC - 1;
        g->wff[JC - 1] = (g->ynd[JC - 1] - g->ycc[JM - 1]) / (g->ycc[JC - 1] - g->ycc[JM - 1]);
        g->wfb[JC - 1] = (g->ycc[JC - 1] - g->ynd[JC - 1]) / (g->ycc[JC - 1] - g->ycc[JM - 1]);
    }
    g->wfb[0] = 1.0; g->wff[0] = 0.0;
    g->wfb[NND2 - 1] = 0.0; g->wff[NND2 - 1] = 1.0;
    /* CONSPARA */
    g->dx = hx / (double)L; g->dxi = 1.0 / g->dx; g->dxqi = g->dxi * g->dxi;
    g->dz = hz / (double)N; g->dzi = 1.0 / g->dz; g->dzqi = g->dzi * g->dzi;
    g->cvisc = 1.0 / ren;
    /* VOLM_tg: sum over cells of 1/RCCI1/DYFI, loop order K,J,I */
    double v = 0.0;
    for (int KC = 1; KC <= N; ++KC)
        for (int JC = 1; JC <= M; ++JC)
            for (int IC = 1; IC <= L; ++IC) v = v + 1.0 / 1.0 / g->dyfi[JC - 1];
    g->volm = v;
    /* TDMA_COEF, channel: Poisson operator with Neumann folded in */
    for (int JC = 1; JC <= M; ++JC) {
        int JP = JC + 1;
        g->amph[JC - 1] = g->dyfi[JC - 1] * g->dyci[JC - 1];
        g->apph[JC - 1] = g->dyfi[JC - 1] * g->dyci[JP - 1];
        if (JC == 1) g->amph[JC - 1] = g->dyfi[JC - 1] * g->dyci[JC - 1] * 0.5;
        if (JP == NND2) g->apph[JC - 1] = g->dyfi[JC - 1] * g->dyci[JP - 1] * 0.5;
    }
    for (int JC = 1; JC <= M; ++JC) g->acph[JC - 1] = -(g->amph[JC - 1] + g->apph[JC - 1]);
    g->acph[0] = g->amph[0] + g->acph[0];
    g->amph[0] = 0.0;
    g->acph[M - 1] = g->apph[M - 1] + g->acph[M - 1];
    g->apph[M - 1] = 0.0;
#define AV(arr, JC, IDR) (g->arr[((IDR)-1) * (M + 1) + (JC)-1])
    for (int i = 0; i < 3 * (M + 1); ++i) g->amvr[i] = g->acvr[i] = g->apvr[i] = 0.0;
    for (int IDR = 1; IDR <= 3; ++IDR) {
        if (IDR == 1 || IDR == 3) {
            for (int JC = 2; JC <= M - 1; ++JC) {
                int JP = JC + 1;
                AV(amvr, JC, IDR) = g->dyfi[JC - 1] * g->dyci[JC - 1];
                AV(apvr, JC, IDR) = g->dyfi[JC - 1] * g->dyci[JP - 1];
                AV(acvr, JC, IDR) = -(AV(amvr, JC, IDR) + AV(apvr, JC, IDR));
            }
            double UCAJ = 4.0 / (2.0 / g->dyci[1] + 1.0 / g->dyfi[0]);
            AV(amvr, 1, IDR) = 0.0;
            AV(apvr, 1, IDR) = UCAJ * g->dyci[1];
            AV(acvr, 1, IDR) = -UCAJ * (g->dyci[1] + 2.0 * g->dyfi[0]);
            UCAJ = 4.0 / (2.0 / g->dyci[M - 1] + 1.0 / g->dyfi[M - 1]);
            AV(amvr, M, IDR) = UCAJ * g->dyci[M - 1];
            AV(apvr, M, IDR) = 0.0;
            AV(acvr, M, IDR) = -UCAJ * (2.0 * g->dyfi[M - 1] + g->dyci[M - 1]);
        } else {
            for (int JC = 2; JC <= M; ++JC) {
                int JM = JC - 1;
                AV(amvr, JC, IDR) = g->dyfi[JM - 1] * g->dyci[JC - 1];
                AV(apvr, JC, IDR) = g->dyfi[JC - 1] * g->dyci[JC - 1];
                AV(acvr, JC, IDR) = -(AV(amvr, JC, IDR) + AV(apvr, JC, IDR));
            }
            AV(amvr, 1, IDR) = 0.0; AV(apvr, 1, IDR) = 0.0; AV(acvr, 1, IDR) = 1.0;
            AV(amvr, NND2, IDR) = 0.0; AV(apvr, NND2, IDR) = 0.0; AV(acvr, NND2, IDR) = 1.0;
        }
    }
#undef AV
    /* modified wavenumbers (src/possion3d_FFT99.f90:152-179) */
    int N3MH = N / 2, N3MP = N3MH + 1, N1MH = L / 2, N1MP = N1MH + 1;
    for (int K = 1; K <= N3MH; ++K) {
        double AN = (double)(K - 1) * 2.0 * PI;
        g->ak3[K - 1] = 2.0 * (1.0 - cos(AN / (double)N)) * g->dzqi;
    }
    for (int K = N3MP; K <= N; ++K) {
        double AN = -(double)(N - K + 1) * 2.0 * PI;
        g->ak3[K - 1] = 2.0 * (1.0 - cos(AN / (double)N)) * g->dzqi;
    }
    for (int I = 1; I <= N1MH; ++I) {
        double AP = (double)(I - 1) * 2.0 * PI;
        g->ak1[I - 1] = 2.0 * (1.0 - cos(AP / (double)L)) * g->dxqi;
    }
    for (int I = N1MP; I <= L; ++I) {
        double AP = -(double)(L - I + 1) * 2.0 * PI;
        g->ak1[I - 1] = 2.0 * (1.0 - cos(AP / (double)L)) * g->dxqi;
    }
}

/* ---------------------------------------------------------------------------
 * Tridiagonal line solvers.  Element (line l, row r) lives at l*ls + r*es, so
 * one routine covers the "IJJ" (march along index 2, es = ISZ, ls = 1) and
 * "IJI" (march along index 1, es = 1, ls = ISZ) forms.
 * orc_tdma_noncyc: TDMAIJJ_nonCYC (src/TDMAIJJ_nonCYC.f90:67-123) and
 *                  TDMAIJI_nonCYC (src/TDMAIJI_nonCYC.f90:67-123).
 *   bc: BCJ/BCI(line, 1:2), column-major -> bc[l], bc[nlines + l]; may be NULL.
 * orc_tdma_cyc:    TDMAIJJ_CYC (src/TDMAIJJ_CYC.f90:67-157) and
 *                  TDMAIJI_CYC (src/TDMAIJI_CYC.f90:67-157).
 * n >= 3 is required (the reference reads outside its arrays for n < 3).
 * ------------------------------------------------------------------------- */
void orc_tdma_noncyc(const double *A, const double *B, const double *C, double *R,
                     const double *bc, int n, int nlines, long es, long ls)
{
#pragma omp parallel
    {
        double *H = (double *)malloc(sizeof(double) * (size_t)n);
        double *G = (double *)malloc(sizeof(double) * (size_t)n);
#pragma omp for schedule(static)
        for (int l = 0; l < nlines; ++l) {
            const double *a = A + (long)l * ls, *b = B + (long)l * ls, *c = C + (long)l * ls;
            double *r = R + (long)l * ls;
            double blo = bc ? bc[l] : 0.0, bhi = bc ? bc[nlines + l] : 0.0;
            H[0] = c[0] / b[0];
            G[0] = (r[0] - a[0] * blo) / b[0];
            for (int j = 1; j < n - 1; ++j) {
                double PP = 1.0 / (b[j * es] - H[j - 1] * a[j * es]);
                H[j] = c[j * es] * PP;
                G[j] = (r[j * es] - a[j * es] * G[j - 1]) * PP;
            }
            {
                int j = n - 1;
                double PP = 1.0 / (b[j * es] - H[j - 1] * a[j * es]);
                G[j] = (r[j * es] - c[j * es] * bhi - a[j * es] * G[j - 1]) * PP;
                r[j * es] = G[j];
            }
            for (int j = n - 2; j >= 0; --j) r[j * es] = G[j] - H[j] * r[(j + 1) * es];
        }
        free(H);
        free(G);
    }
}

void orc_tdma_cyc(const double *A, const double *B, const double *C, double *R,
                  int n, int nlines, long es, long ls)
{
#pragma omp parallel
    {
        double *H0 = (double *)calloc((size_t)n, sizeof(double));
        double *G1 = (double *)calloc((size_t)n, sizeof(double));
        double *G2 = (double *)calloc((size_t)n, sizeof(double));
#pragma omp for schedule(static)
        for (int l = 0; l < nlines; ++l) {
            const double *a = A + (long)l * ls, *b = B + (long)l * ls, *c = C + (long)l * ls;
            double *r = R + (long)l * ls;
            H0[n - 2] = 0.0; /* never set by the reference: stays at its zero fill */
            H0[0] = c[0] / b[0];
            G1[0] = r[0] / b[0];
            G2[0] = -a[0] / b[0];
            for (int i = 1; i <= n - 3; ++i) {
                double PP = 1.0 / (b[i * es] - H0[i - 1] * a[i * es]);
                double QQ1 = G1[i - 1] * a[i * es];
                double QQ2 = G2[i - 1] * a[i * es];
                H0[i] = c[i * es] * PP;
                G1[i] = (r[i * es] - QQ1) * PP;
                G2[i] = -QQ2 * PP;
            }
            {
                int i = n - 2;
                double PP = 1.0 / (b[i * es] - H0[i - 1] * a[i * es]);
                double QQ1 = G1[i - 1] * a[i * es];
                double QQ2 = G2[i - 1] * a[i * es];
                G1[i] = (r[i * es] - QQ1) * PP;
                G2[i] = (-c[i * es] - QQ2) * PP;
            }
            for (int i = n - 3; i >= 0; --i) {
                G1[i] = -H0[i] * G1[i + 1] + G1[i];
                G2[i] = -H0[i] * G2[i + 1] + G2[i];
            }
            {
                int i = n - 1;
                double PP = r[i * es] - c[i * es] * G1[0] - a[i * es] * G1[i - 1];
                double QQ1 = b[i * es] + c[i * es] * G2[0] + a[i * es] * G2[i - 1];
                r[i * es] = PP / QQ1;
            }
            double xn = r[(n - 1) * es];
            for (int i = n - 2; i >= 0; --i) r[i * es] = G1[i] + G2[i] * xn;
        }
        free(H0);
        free(G1);
        free(G2);
    }
}

/* ---------------------------------------------------------------------------
 * Discrete Fourier transforms with the conventions of the vendored FFT99
 * package (libs/fft99/FFT99.f90:98-105, libs/fft99/CFFT99.f90:182-197):
 *   FFT99 ISIGN=-1: a_k + i b_k = (1/n) sum_j x_j exp(-2 pi i jk/n), k=0..n/2
 *   FFT99 ISIGN=+1: x_j = sum_k C_k exp(+2 pi i jk/n), C_k = a_k+ib_k,
 *                   C_{n-k} = a_k - i b_k, with b_0 = b_{n/2} taken as 0
 *   CFFT99: unnormalised, exp(-...) for ISIGN=-1, exp(+...) for ISIGN=+1.
 * Computed here by a plain mixed-radix (2,3,5, any other factor by direct
 * DFT) recursive Cooley-Tukey transform in FP64.
 * ------------------------------------------------------------------------- */
static void cfft_rec(const double *xin, double *out, int n, int stride, int sign,
                     const double *wr, const double *wi, int nfull, double *scratch)
{
    /* xin: complex interleaved with element stride 'stride' (in complex units)
     * out: contiguous n complex.  wr/wi: twiddles of the full length nfull. */
    if (n == 1) {
        out[0] = xin[0];
        out[1] = xin[1];
        return;
    }
    int p = 0;
    if (n % 4 == 0) p = 4;
    else if (n % 2 == 0) p = 2;
    else if (n % 3 == 0) p = 3;
    else if (n % 5 == 0) p = 5;
    else p = n;
    int m = n / p;
    /* sub-transforms of the p decimated sequences */
    for (int q = 0; q < p; ++q)
        cfft_rec(xin + 2 * (long)q * stride, scratch + 2 * (long)q * m, m, stride * p, sign,
                 wr, wi, nfull, scratch + 2 * (long)n);
    int tw = nfull / n; /* twiddle step for this length */
    for (int k = 0; k < m; ++k) {
        for (int s = 0; s < p; ++s) {
            /* out[k + s m] = sum_q W_n^{q (k + s m)} sub_q[k] */
            double accr = 0.0, acci = 0.0;
            for (int q = 0; q < p; ++q) {
                long e = ((long)q * (k + s * m)) % n;
                double c = wr[e * tw], sn = sign * wi[e * tw];
                double xr = scratch[2 * ((long)q * m + k)], xi = scratch[2 * ((long)q * m + k) + 1];
                accr += c * xr - sn * xi;
                acci += c * xi + sn * xr;
            }
            out[2 * (k + s * m)] = accr;
            out[2 * (k + s * m) + 1] = acci;
        }
    }
}

typedef struct {
    int n;
    double *wr, *wi; /* cos/sin(2 pi e / n), e = 0..n-1, sign applied at use */
    double *scratch;
} cplan;

static void cplan_init(cplan *p, int n)
{
    const double PI = 2.0 * asin(1.0);
    p->n = n;
    p->wr = (double *)malloc(sizeof(double) * (size_t)n);
    p->wi = (double *)malloc(sizeof(double) * (size_t)n);
    for (int e = 0; e < n; ++e) {
        p->wr[e] = cos(2.0 * PI * (double)e / (double)n);
        p->wi[e] = -sin(2.0 * PI * (double)e / (double)n); /* exp(-i theta) */
    }
    /* scratch: recursion depth <= log2 n, each level needs 2n doubles */
    p->scratch = (double *)malloc(sizeof(double) * (size_t)(4 * n * 24 + 16));
}

static void cplan_free(cplan *p)
{
    free(p->wr);
    free(p->wi);
    free(p->scratch);
}

/* in-place complex transform of one contiguous interleaved vector;
 * isign = -1 gives exp(-i..), +1 gives exp(+i..); unnormalised. */
static void cfft_run(cplan *p, double *x, int isign, double *tmp)
{
    cfft_rec(x, tmp, p->n, 1, isign < 0 ? 1 : -1, p->wr, p->wi, p->n, p->scratch);
    memcpy(x, tmp, sizeof(double) * 2 * (size_t)p->n);
}

/* single-vector API used by the tests */
void orc_cfft99(double *x, int n, int isign)
{
    cplan p;
    cplan_init(&p, n);
    double *tmp = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    cfft_run(&p, x, isign, tmp);
    free(tmp);
    cplan_free(&p);
}

/* x[n] -> ab[n+2] = a0,b0,a1,b1,...,a_{n/2},b_{n/2} (FFT99 ISIGN=-1) */
void orc_fft99_fwd(const double *x, double *ab, int n)
{
    cplan p;
    cplan_init(&p, n);
    double *z = (double *)malloc(sizeof(double) * 4 * (size_t)n);
    for (int j = 0; j < n; ++j) { z[2 * j] = x[j]; z[2 * j + 1] = 0.0; }
    cfft_run(&p, z, -1, z + 2 * n);
    for (int k = 0; k <= n / 2; ++k) {
        ab[2 * k] = z[2 * k] / (double)n;
        ab[2 * k + 1] = z[2 * k + 1] / (double)n;
    }
    ab[1] = 0.0;
    ab[n + 1] = 0.0;
    free(z);
    cplan_free(&p);
}

/* ab[n+2] -> x[n] (FFT99 ISIGN=+1) */
void orc_fft99_inv(const double *ab, double *x, int n)
{
    cplan p;
    cplan_init(&p, n);
    double *z = (double *)malloc(sizeof(double) * 4 * (size_t)n);
    z[0] = ab[0]; z[1] = 0.0;
    for (int k = 1; k < n / 2; ++k) {
        z[2 * k] = ab[2 * k]; z[2 * k + 1] = ab[2 * k + 1];
        z[2 * (n - k)] = ab[2 * k]; z[2 * (n - k) + 1] = -ab[2 * k + 1];
    }
    z[n] = ab[n]; z[n + 1] = 0.0;
    cfft_run(&p, z, +1, z + 2 * n);
    for (int j = 0; j < n; ++j) x[j] = z[2 * j];
    free(z);
    cplan_free(&p);
}

/* ---------------------------------------------------------------------------
 * Pressure Poisson solve, FFT99_POIS3D_periodicxz (src/possion3d_FFT99.f90:
 * 254-488) with the coefficients of TDMA_PHI_COEF (:185-220).
 *   f:   RHSLLPHI(L, M, N)
 *   dph: DPH(L, 0:M+1, N); interior written, ghost planes zeroed as in :447.
 *   cyclic_y: use TDMAIJJ_CYC (ICASE == IBOX3P) instead of TDMAIJJ_nonCYC.
 * The y solves use exactly the reference's Thomas recurrences (orc_tdma_*).
 * ------------------------------------------------------------------------- */
void orc_poisson_fft99(const orc_grid *g, const double *f, double *dph, int cyclic_y)
{
    const int L = g->L, M = g->M, N = g->N, LH = L / 2 + 1;
    const size_t spec = (size_t)LH * M * N;
    double *Re = (double *)calloc(spec, sizeof(double)); /* (LH, M, N) */
    double *Im = (double *)calloc(spec, sizeof(double));
#define SP(kx, j, kz) ((size_t)(kx) + (size_t)LH * ((size_t)(j) + (size_t)M * (size_t)(kz)))
    /* forward: per y-plane, real FFT in x (1/L), complex FFT in z, x 1/N */
#pragma omp parallel
    {
        cplan px, pz;
        cplan_init(&px, L);
        cplan_init(&pz, N);
        double *zx = (double *)malloc(sizeof(double) * 4 * (size_t)L);
        double *zz = (double *)malloc(sizeof(double) * 4 * (size_t)N);
        double *XA = (double *)malloc(sizeof(double) * 2 * (size_t)N * LH);
#pragma omp for schedule(static)
        for (int j = 0; j < M; ++j) {
            for (int k = 0; k < N; ++k) {
                const double *row = f + (size_t)L * ((size_t)j + (size_t)M * k);
                for (int i = 0; i < L; ++i) { zx[2 * i] = row[i]; zx[2 * i + 1] = 0.0; }
                cfft_run(&px, zx, -1, zx + 2 * L);
                for (int kx = 0; kx < LH; ++kx) {
                    XA[2 * ((size_t)kx * N + k)] = zx[2 * kx] / (double)L;
                    XA[2 * ((size_t)kx * N + k) + 1] = (kx == 0 || kx == L / 2) ? 0.0 : zx[2 * kx + 1] / (double)L;
                }
            }
            for (int kx = 0; kx < LH; ++kx) {
                memcpy(zz, XA + 2 * (size_t)kx * N, sizeof(double) * 2 * (size_t)N);
                cfft_run(&pz, zz, -1, zz + 2 * N);
                double UN3M = 1.0 / (double)N;
                for (int kz = 0; kz < N; ++kz) {
                    Re[SP(kx, j, kz)] = zz[2 * kz] * UN3M;
                    Im[SP(kx, j, kz)] = zz[2 * kz + 1] * UN3M;
                }
            }
        }
        free(zx); free(zz); free(XA);
        cplan_free(&px);
        cplan_free(&pz);
    }
    Re[SP(0, 0, 0)] = 0.0;
    Im[SP(0, 0, 0)] = 0.0;
    /* y-solves: lines (kx, kz), march along j with stride LH */
#pragma omp parallel
    {
        double *A = (double *)malloc(sizeof(double) * (size_t)LH * M);
        double *B = (double *)malloc(sizeof(double) * (size_t)LH * M);
        double *C = (double *)malloc(sizeof(double) * (size_t)LH * M);
#pragma omp for schedule(static)
        for (int kz = 0; kz < N; ++kz) {
            for (int kx = 0; kx < LH; ++kx)
                for (int j = 0; j < M; ++j) {
                    A[kx + (size_t)LH * j] = g->amph[j];
                    C[kx + (size_t)LH * j] = g->apph[j];
                    /* RCCI2 = 1 in the channel */
                    B[kx + (size_t)LH * j] = g->acph[j] + (-g->ak1[kx] / 1.0 - g->ak3[kz]);
                }
            if (kz == 0) {
                B[0] = g->acph[0] + (-g->ak1[0] / 1.0 - g->ak3[0]);
                C[0] = 0.0;
                A[0] = 0.0;
            }
            for (int pass = 0; pass < 2; ++pass) {
                double *F = (pass == 0 ? Re : Im) + (size_t)LH * M * kz;
                if (kz == 0) F[0] = 0.0;
                /* lines over kx (ls = 1), rows over j (es = LH) -- the real
                 * matrices are identical for the two passes */
                /* orc_tdma_* open their own parallel regions; call serial cores */
                if (cyclic_y) {
                    double *H0 = (double *)calloc((size_t)M, sizeof(double));
                    double *G1 = (double *)calloc((size_t)M, sizeof(double));
                    double *G2 = (double *)calloc((size_t)M, sizeof(double));
                    for (int l = 0; l < LH; ++l) {
                        const double *a = A + l, *b = B + l, *c = C + l;
                        double *r = F + l;
                        long es = LH;
                        int n = M;
                        H0[n - 2] = 0.0;
                        H0[0] = c[0] / b[0]; G1[0] = r[0] / b[0]; G2[0] = -a[0] / b[0];
                        for (int i = 1; i <= n - 3; ++i) {
                            double PP = 1.0 / (b[i * es] - H0[i - 1] * a[i * es]);
                            double QQ1 = G1[i - 1] * a[i * es], QQ2 = G2[i - 1] * a[i * es];
                            H0[i] = c[i * es] * PP; G1[i] = (r[i * es] - QQ1) * PP; G2[i] = -QQ2 * PP;
                        }
                        {
                            int i = n - 2;
                            double PP = 1.0 / (b[i * es] - H0[i - 1] * a[i * es]);
                            double QQ1 = G1[i - 1] * a[i * es], QQ2 = G2[i - 1] * a[i * es];
                            G1[i] = (r[i * es] - QQ1) * PP; G2[i] = (-c[i * es] - QQ2) * PP;
                        }
                        for (int i = n - 3; i >= 0; --i) {
                            G1[i] = -H0[i] * G1[i + 1] + G1[i];
                            G2[i] = -H0[i] * G2[i + 1] + G2[i];
                        }
                        {
                            int i = n - 1;
                            double PP = r[i * es] - c[i * es] * G1[0] - a[i * es] * G1[i - 1];
                            double QQ1 = b[i * es] + c[i * es] * G2[0] + a[i * es] * G2[i - 1];
                            r[i * es] = PP / QQ1;
                        }
                        double xn = r[(n - 1) * es];
                        for (int i = n - 2; i >= 0; --i) r[i * es] = G1[i] + G2[i] * xn;
                    }
                    free(H0); free(G1); free(G2);
                } else {
                    double *H = (double *)malloc(sizeof(double) * (size_t)M);
                    double *G = (double *)malloc(sizeof(double) * (size_t)M);
                    for (int l = 0; l < LH; ++l) {
                        const double *a = A + l, *b = B + l, *c = C + l;
                        double *r = F + l;
                        long es = LH;
                        int n = M;
                        H[0] = c[0] / b[0];
                        G[0] = (r[0] - a[0] * 0.0) / b[0];
                        for (int j = 1; j < n - 1; ++j) {
                            double PP = 1.0 / (b[j * es] - H[j - 1] * a[j * es]);
                            H[j] = c[j * es] * PP;
                            G[j] = (r[j * es] - a[j * es] * G[j - 1]) * PP;
                        }
                        {
                            int j = n - 1;
                            double PP = 1.0 / (b[j * es] - H[j - 1] * a[j * es]);
                            G[j] = (r[j * es] - c[j * es] * 0.0 - a[j * es] * G[j - 1]) * PP;
                            r[j * es] = G[j];
                        }
                        for (int j = n - 2; j >= 0; --j) r[j * es] = G[j] - H[j] * r[(j + 1) * es];
                    }
                    free(H); free(G);
                }
            }
        }
        free(A); free(B); free(C);
    }
    /* inverse: complex FFT in z (+1), real FFT in x (+1), both unnormalised */
    memset(dph, 0, sizeof(double) * (size_t)L * (M + 2) * N);
#pragma omp parallel
    {
        cplan px, pz;
        cplan_init(&px, L);
        cplan_init(&pz, N);
        double *zx = (double *)malloc(sizeof(double) * 4 * (size_t)L);
        double *zz = (double *)malloc(sizeof(double) * 4 * (size_t)N);
        double *XA = (double *)malloc(sizeof(double) * 2 * (size_t)N * LH);
#pragma omp for schedule(static)
        for (int j = 0; j < M; ++j) {
            for (int kx = 0; kx < LH; ++kx) {
                for (int kz = 0; kz < N; ++kz) {
                    zz[2 * kz] = Re[SP(kx, j, kz)];
                    zz[2 * kz + 1] = Im[SP(kx, j, kz)];
                }
                cfft_run(&pz, zz, +1, zz + 2 * N);
                memcpy(XA + 2 * (size_t)kx * N, zz, sizeof(double) * 2 * (size_t)N);
            }
            for (int k = 0; k < N; ++k) {
                /* half-complex -> full Hermitian, b_0 = b_{L/2} = 0 */
                zx[0] = XA[2 * ((size_t)0 * N + k)]; zx[1] = 0.0;
                for (int kx = 1; kx < L / 2; ++kx) {
                    double ar = XA[2 * ((size_t)kx * N + k)], ai = XA[2 * ((size_t)kx * N + k) + 1];
                    zx[2 * kx] = ar; zx[2 * kx + 1] = ai;
                    zx[2 * (L - kx)] = ar; zx[2 * (L - kx) + 1] = -ai;
                }
                zx[L] = XA[2 * ((size_t)(L / 2) * N + k)]; zx[L + 1] = 0.0;
                cfft_run(&px, zx, +1, zx + 2 * L);
                double *row = dph + (size_t)L * ((size_t)(j + 1) + (size_t)(M + 2) * k);
                for (int i = 0; i < L; ++i) row[i] = zx[2 * i];
            }
        }
        free(zx); free(zz); free(XA);
        cplan_free(&px);
        cplan_free(&pz);
    }
#undef SP
    free(Re);
    free(Im);
}

/* Residual of the discrete Poisson operator, CHECK_FFT_SOLVER
 * (src/PRESSURE_CORRECTION.f90:159-202): res = L_h(dph) - f, with DPH ghost
 * planes taken as given (the reference applies BC_WALL_DPH first).  Returns
 * max |res| and writes the field if res != NULL. */
double orc_poisson_residual(const orc_grid *g, const double *dph, const double *f, double *res)
{
    const int L = g->L, M = g->M, N = g->N;
    double mx = 0.0;
#define PH(i, j, k) dph[(size_t)((i)-1) + (size_t)L * ((size_t)(j) + (size_t)(M + 2) * ((k)-1))]
#pragma omp parallel for reduction(max : mx) schedule(static)
    for (int JC = 1; JC <= M; ++JC) {
        int JM = JC - 1, JP = JC + 1;
        for (int KC = 1; KC <= N; ++KC) {
            int KP = KC == N ? 1 : KC + 1, KM = KC == 1 ? N : KC - 1;
            for (int IC = 1; IC <= L; ++IC) {
                int IP = IC == L ? 1 : IC + 1, IM = IC == 1 ? L : IC - 1;
                double lx = (PH(IP, JC, KC) - 2.0 * PH(IC, JC, KC) + PH(IM, JC, KC)) * g->dxqi;
                double ly = PH(IC, JP, KC) * g->apph[JC - 1] + PH(IC, JC, KC) * g->acph[JC - 1] +
                            PH(IC, JM, KC) * g->amph[JC - 1];
                double lz = (PH(IC, JC, KP) - 2.0 * PH(IC, JC, KC) + PH(IC, JC, KM)) * g->dzqi;
                double r = lx + ly + lz - f[(size_t)(IC - 1) + (size_t)L * ((size_t)(JC - 1) + (size_t)M * (KC - 1))];
                if (res) res[(size_t)(IC - 1) + (size_t)L * ((size_t)(JC - 1) + (size_t)M * (KC - 1))] = r;
                if (fabs(r) > mx) mx = fabs(r);
            }
        }
    }
#undef PH
    return mx;
}

/* ---------------------------------------------------------------------------
 * One RK3 substep of the turbulence-generator channel (tg domain), one rank.
 * SOLVERRK3_MOM_tg (src/SOLVERRK3_MOM_tg.f90:18-61) and everything it calls.
 * State: q = Q_tg, pr = PR_tg, convh0 = CONVH0_tg.  Work buffers: dph = DPH_tg,
 * rhs = RHS_tg, rhsllphi = RHSLLPHI_tg, qtmp = Qtmp_tg; their contents on exit
 * are those the reference leaves (Qtmp holds Hx, RHS holds the last solve).
 * ------------------------------------------------------------------------- */
typedef struct {
    double *q, *pr, *convh0;          /* state */
    double *dph, *rhs, *rhsllphi, *qtmp; /* work, persistent across calls */
    double dpgrns;                    /* last mass-flux correction (out) */
} orc_state;

#define GL (g->L)
#define GM (g->M)
#define GN (g->N)
#define QQ(i, j, k, c) s->q[(size_t)((i)-1) + (size_t)GL * ((size_t)(j) + (size_t)(GM + 2) * ((size_t)((k)-1) + (size_t)GN * ((c)-1)))]
#define G3(arr, i, j, k) arr[(size_t)((i)-1) + (size_t)GL * ((size_t)(j) + (size_t)(GM + 2) * ((k)-1))]
#define F3(arr, i, j, k) arr[(size_t)((i)-1) + (size_t)GL * ((size_t)((j)-1) + (size_t)GM * ((k)-1))]
#define CH0(i, j, k, c) s->convh0[(size_t)((i)-1) + (size_t)GL * ((size_t)((j)-1) + (size_t)GM * ((size_t)((k)-1) + (size_t)GN * ((c)-1)))]
#define IPV(i) ((i) == GL ? 1 : (i) + 1)
#define IMV(i) ((i) == 1 ? GL : (i)-1)
#define KPV(k) ((k) == GN ? 1 : (k) + 1)
#define KMV(k) ((k) == 1 ? GN : (k)-1)
#define AMV(JC, IDR) g->amvr[((IDR)-1) * (GM + 1) + (JC)-1]
#define ACV(JC, IDR) g->acvr[((IDR)-1) * (GM + 1) + (JC)-1]
#define APV(JC, IDR) g->apvr[((IDR)-1) * (GM + 1) + (JC)-1]

static const double TGAM[4] = {1.0, 8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
static const double TROH[4] = {0.0, 0.0, -17.0 / 60.0, -5.0 / 12.0};
static double talp(int ns) { return TGAM[ns] + TROH[ns]; }

/* CONVECTION_X_tg (src/CONVECTION_X.f90:23-74) -> Qtmp */
void orc_convection_x(const orc_grid *g, orc_state *s)
{
    memset(s->qtmp, 0, sizeof(double) * (size_t)GL * GM * GN);
    double COE1 = 0.5 * 0.5 * g->dxi;
#pragma omp parallel for schedule(static)
    for (int JC = 1; JC <= GM; ++JC) {
        int JM = JC - 1, JP = JC + 1, JJ = JC, JJP = JJ + 1;
        double COE2 = g->dyfi[JJ - 1] * 0.5 * 1.0;
        double COE3 = 0.5 * 0.5 * g->dzi * 1.0;
        for (int KC = 1; KC <= GN; ++KC) {
            int KM = KMV(KC), KP = KPV(KC);
            for (int IC = 1; IC <= GL; ++IC) {
                int IP = IPV(IC), IM = IMV(IC);
                double H11 = ((QQ(IP, JC, KC, 1) + QQ(IC, JC, KC, 1)) * (QQ(IP, JC, KC, 1) + QQ(IC, JC, KC, 1)) -
                              (QQ(IM, JC, KC, 1) + QQ(IC, JC, KC, 1)) * (QQ(IM, JC, KC, 1) + QQ(IC, JC, KC, 1))) * COE1;
                double H12 = ((QQ(IC, JP, KC, 2) + QQ(IM, JP, KC, 2)) *
                                  (g->wff[JJP - 1] * QQ(IC, JP, KC, 1) + g->wfb[JJP - 1] * QQ(IC, JC, KC, 1)) -
                              (QQ(IC, JC, KC, 2) + QQ(IM, JC, KC, 2)) *
                                  (g->wff[JJ - 1] * QQ(IC, JC, KC, 1) + g->wfb[JJ - 1] * QQ(IC, JM, KC, 1))) * COE2;
                double H13 = ((QQ(IC, JC, KP, 3) + QQ(IM, JC, KP, 3)) * (QQ(IC, JC, KP, 1) + QQ(IC, JC, KC, 1)) -
                              (QQ(IC, JC, KC, 3) + QQ(IM, JC, KC, 3)) * (QQ(IC, JC, KC, 1) + QQ(IC, JC, KM, 1))) * COE3;
                F3(s->qtmp, IC, JC, KC) = -(H11 + H12 + H13);
            }
        }
    }
}

/* CONVECTION_Y_tg (src/CONVECTION_Y.f90:22-93), channel branch -> DPH interior */
void orc_convection_y(const orc_grid *g, orc_state *s)
{
    memset(s->dph, 0, sizeof(double) * (size_t)GL * (GM + 2) * GN);
    int NYI = 2;
    double COE1 = g->dxi * 0.5, COE30 = g->dzi * 0.5;
#pragma omp parallel for schedule(static)
    for (int JC = NYI; JC <= GM; ++JC) {
        int JM = JC - 1, JP = JC + 1, JJ = JC;
        double COE2 = g->dyci[JJ - 1] * 0.5 * 0.5;
        double COE31 = COE30 * 1.0;
        for (int KC = 1; KC <= GN; ++KC) {
            int KM = KMV(KC), KP = KPV(KC);
            for (int IC = 1; IC <= GL; ++IC) {
                int IP = IPV(IC), IM = IMV(IC);
                double H21 = ((g->wff[JJ - 1] * QQ(IP, JC, KC, 1) + g->wfb[JJ - 1] * QQ(IP, JM, KC, 1)) *
                                  (QQ(IP, JC, KC, 2) + QQ(IC, JC, KC, 2)) -
                              (g->wff[JJ - 1] * QQ(IC, JC, KC, 1) + g->wfb[JJ - 1] * QQ(IC, JM, KC, 1)) *
                                  (QQ(IC, JC, KC, 2) + QQ(IM, JC, KC, 2))) * COE1;
                double H22 = ((QQ(IC, JP, KC, 2) * 1.0 + QQ(IC, JC, KC, 2) * 1.0) * (QQ(IC, JP, KC, 2) + QQ(IC, JC, KC, 2)) -
                              (QQ(IC, JC, KC, 2) * 1.0 + QQ(IC, JM, KC, 2) * 1.0) * (QQ(IC, JC, KC, 2) + QQ(IC, JM, KC, 2))) * COE2;
                double H23 = ((g->wff[JJ - 1] * QQ(IC, JC, KP, 3) * 1.0 + g->wfb[JJ - 1] * QQ(IC, JM, KP, 3) * 1.0) *
                                  (QQ(IC, JC, KP, 2) + QQ(IC, JC, KC, 2)) -
                              (g->wff[JJ - 1] * QQ(IC, JC, KC, 3) * 1.0 + g->wfb[JJ - 1] * QQ(IC, JM, KC, 3) * 1.0) *
                                  (QQ(IC, JC, KC, 2) + QQ(IC, JC, KM, 2))) * COE31;
                G3(s->dph, IC, JC, KC) = -(H21 + H22 + H23);
            }
        }
    }
}

/* CONVECTION_Z_tg (src/CONVECTION_Z.f90:22-76), channel branch -> RHSLLPHI */
void orc_convection_z(const orc_grid *g, orc_state *s)
{
    memset(s->rhsllphi, 0, sizeof(double) * (size_t)GL * GM * GN);
    double COE1 = 0.5 * 0.5 * g->dxi;
#pragma omp parallel for schedule(static)
    for (int JC = 1; JC <= GM; ++JC) {
        int JM = JC - 1, JP = JC + 1, JJ = JC, JJP = JJ + 1;
        double COE2 = g->dyfi[JJ - 1] * 0.5;
        double COE3 = 0.5 * 0.5 * g->dzi * 1.0;
        for (int KC = 1; KC <= GN; ++KC) {
            int KM = KMV(KC), KP = KPV(KC);
            for (int IC = 1; IC <= GL; ++IC) {
                int IP = IPV(IC), IM = IMV(IC);
                double H31 = ((QQ(IP, JC, KC, 1) + QQ(IP, JC, KM, 1)) * (QQ(IP, JC, KC, 3) + QQ(IC, JC, KC, 3)) -
                              (QQ(IC, JC, KC, 1) + QQ(IC, JC, KM, 1)) * (QQ(IC, JC, KC, 3) + QQ(IM, JC, KC, 3))) * COE1;
                double H32 = ((QQ(IC, JP, KC, 2) + QQ(IC, JP, KM, 2)) *
                                  (g->wff[JJP - 1] * QQ(IC, JP, KC, 3) * 1.0 + g->wfb[JJP - 1] * QQ(IC, JC, KC, 3) * 1.0) -
                              (QQ(IC, JC, KC, 2) + QQ(IC, JC, KM, 2)) *
                                  (g->wff[JJ - 1] * QQ(IC, JC, KC, 3) * 1.0 + g->wfb[JJ - 1] * QQ(IC, JM, KC, 3) * 1.0)) * COE2;
                double H33 = ((QQ(IC, JC, KP, 3) + QQ(IC, JC, KC, 3)) * (QQ(IC, JC, KP, 3) + QQ(IC, JC, KC, 3)) -
                              (QQ(IC, JC, KC, 3) + QQ(IC, JC, KM, 3)) * (QQ(IC, JC, KC, 3) + QQ(IC, JC, KM, 3))) * COE3;
                F3(s->rhsllphi, IC, JC, KC) = -(H31 + H32 + H33);
            }
        }
    }
}

/* RHS_CvLpGpS_tg (src/RHS_CvLpGpS_tg.f90:29-220).  The flow-drive term of
 * the streamwise component IDR = NFLOW (:178-213): FLOWDRVTP = 1 constant
 * mass flux, DPGRNS = (sum RHS / DYFI / RCCI1) / VOLM_tg; FLOWDRVTP = 2
 * constant pressure gradient, DPGRNS = -CFGV / 2 * TALP(NS) * DT.  NFLOW is
 * 1 (x) or 3 (z): the y component's wall row J = 1 has no RHS of its own. */
void orc_rhs_tg(const orc_grid *g, orc_state *s, int NS, int IDR, double DT, int flowdrvtp, int nflow, double cfgv)
{
    int NII = (IDR == 2) ? 2 : 1;
    const double *H = IDR == 1 ? s->qtmp : IDR == 3 ? s->rhsllphi : NULL;
    double COE1 = talp(NS) * g->cvisc;
#pragma omp parallel for schedule(static)
    for (int KC = 1; KC <= GN; ++KC) {
        int KM = KMV(KC), KP = KPV(KC);
        for (int JC = NII; JC <= GM; ++JC) {
            int JM = JC - 1, JP = JC + 1, JJ = JC;
            double RMC2 = 1.0; /* RNDI2/RCCI2 = 1 in the channel */
            for (int IC = 1; IC <= GL; ++IC) {
                int IP = IPV(IC), IM = IMV(IC);
                double conv = IDR == 2 ? G3(s->dph, IC, JC, KC) : F3(H, IC, JC, KC);
                double RHSC = TGAM[NS] * conv + TROH[NS] * CH0(IC, JC, KC, IDR);
                double RHSL = COE1 * ((QQ(IP, JC, KC, IDR) - 2.0 * QQ(IC, JC, KC, IDR) + QQ(IM, JC, KC, IDR)) * g->dxqi +
                                      (QQ(IC, JC, KP, IDR) - 2.0 * QQ(IC, JC, KC, IDR) + QQ(IC, JC, KM, IDR)) * g->dzqi * RMC2 +
                                      (QQ(IC, JP, KC, IDR) * APV(JJ, IDR) + QQ(IC, JC, KC, IDR) * ACV(JJ, IDR) +
                                       QQ(IC, JM, KC, IDR) * AMV(JJ, IDR)));
                CH0(IC, JC, KC, IDR) = conv;
                F3(s->rhs, IC, JC, KC) = (RHSC + RHSL) * DT;
            }
        }
    }
    double COE2 = talp(NS) * DT;
#pragma omp parallel for schedule(static)
    for (int KC = 1; KC <= GN; ++KC) {
        for (int JC = NII; JC <= GM; ++JC) {
            for (int IC = 1; IC <= GL; ++IC) {
                double PGM;
                if (IDR == 1) PGM = (G3(s->pr, IC, JC, KC) - G3(s->pr, IMV(IC), JC, KC)) * g->dxi * COE2;
                else if (IDR == 2) PGM = (G3(s->pr, IC, JC, KC) - G3(s->pr, IC, JC - 1, KC)) * (g->dyci[JC - 1] / 1.0) * COE2;
                else PGM = (G3(s->pr, IC, JC, KC) - G3(s->pr, IC, JC, KMV(KC))) * g->dzi * COE2;
                F3(s->rhs, IC, JC, KC) = F3(s->rhs, IC, JC, KC) - PGM;
            }
        }
    }
    if (IDR == nflow) {
        double DPGRNS = 0.0;
        if (flowdrvtp == 1) {
            double INTGRHSY = 0.0;
            /* summed in a fixed order (IC, JC, KC) as in the reference */
            for (int IC = 1; IC <= GL; ++IC)
                for (int JC = 1; JC <= GM; ++JC)
                    for (int KC = 1; KC <= GN; ++KC)
                        INTGRHSY = INTGRHSY + F3(s->rhs, IC, JC, KC) / g->dyfi[JC - 1] / 1.0;
            DPGRNS = INTGRHSY / g->volm;
        } else if (flowdrvtp == 2) {
            DPGRNS = -0.50 * cfgv * COE2;
        }
        s->dpgrns = DPGRNS;
#pragma omp parallel for schedule(static)
        for (int K = 1; K <= GN; ++K)
            for (int I = 1; I <= GL; ++I)
                for (int J = NII; J <= GM; ++J) F3(s->rhs, I, J, K) = F3(s->rhs, I, J, K) - DPGRNS;
    }
}

/* MOMFA_tg (src/MOMFA_tg.f90:17-240): x-, z-, y- CN line solves, Q += RHS */
void orc_momfa_tg(const orc_grid *g, orc_state *s, int NS, int IDR, double DT)
{
    double FACOE = 0.5 * talp(NS) * DT * g->cvisc;
    int N2I = (IDR == 2) ? 2 : 1;
    int JSZ = GM - N2I + 1;
    /* MOMFA1_X_tg: per k-plane, lines along i (contiguous), rows j=N2I..M */
    {
        double COE = -FACOE * g->dxqi;
        double *A = (double *)malloc(sizeof(double) * (size_t)GL * GM);
        double *B = (double *)malloc(sizeof(double) * (size_t)GL * GM);
        double *C = (double *)malloc(sizeof(double) * (size_t)GL * GM);
        for (size_t t = 0; t < (size_t)GL * GM; ++t) {
            C[t] = COE;
            A[t] = COE;
            B[t] = 1.0 - C[t] - A[t];
        }
        for (int K = 1; K <= GN; ++K) {
            double *F = &F3(s->rhs, 1, N2I, K);
            orc_tdma_cyc(A, B, C, F, GL, JSZ, 1, GL);
        }
        free(A); free(B); free(C);
    }
    /* MOMFA2_Z_tg: per i, lines along k (stride L*M), rows j=N2I..M */
    {
        double *A = (double *)malloc(sizeof(double) * (size_t)GN * GM);
        double *B = (double *)malloc(sizeof(double) * (size_t)GN * GM);
        double *C = (double *)malloc(sizeof(double) * (size_t)GN * GM);
        double *FK = (double *)malloc(sizeof(double) * (size_t)GN * GM);
        for (int I = 1; I <= GL; ++I) {
            for (int K = 1; K <= GN; ++K)
                for (int J = N2I; J <= GM; ++J) {
                    double RMC2 = 1.0 * (-FACOE * g->dzqi);
                    size_t t = (size_t)(K - 1) + (size_t)GN * (J - N2I);
                    C[t] = RMC2;
                    A[t] = RMC2;
                    B[t] = 1.0 - C[t] - A[t];
                    FK[t] = F3(s->rhs, I, J, K);
                }
            orc_tdma_cyc(A, B, C, FK, GN, JSZ, 1, GN);
            for (int K = 1; K <= GN; ++K)
                for (int J = N2I; J <= GM; ++J) F3(s->rhs, I, J, K) = FK[(size_t)(K - 1) + (size_t)GN * (J - N2I)];
        }
        free(A); free(B); free(C); free(FK);
    }
    /* MOMFA3_Y_tg: per k-plane, lines along j (stride L) over all j=1..M */
    {
        double *A = (double *)malloc(sizeof(double) * (size_t)GL * GM);
        double *B = (double *)malloc(sizeof(double) * (size_t)GL * GM);
        double *C = (double *)malloc(sizeof(double) * (size_t)GL * GM);
        for (int I = 1; I <= GL; ++I)
            for (int J = 1; J <= GM; ++J) {
                size_t t = (size_t)(I - 1) + (size_t)GL * (J - 1);
                B[t] = 1.0 - FACOE * ACV(J, IDR);
                C[t] = -FACOE * APV(J, IDR);
                A[t] = -FACOE * AMV(J, IDR);
            }
        for (int K = 1; K <= GN; ++K) {
            double *F = &F3(s->rhs, 1, 1, K);
            orc_tdma_noncyc(A, B, C, F, NULL, GM, GL, GL, 1);
        }
        free(A); free(B); free(C);
    }
#pragma omp parallel for schedule(static)
    for (int K = 1; K <= GN; ++K)
        for (int J = N2I; J <= GM; ++J)
            for (int I = 1; I <= GL; ++I) QQ(I, J, K, IDR) = F3(s->rhs, I, J, K) + QQ(I, J, K, IDR);
}

/* BC_WALL_Q_tg (src/BC_WALL.f90:185-227) after INTFC_VARS3 on one rank */
void orc_bc_wall_q(const orc_grid *g, orc_state *s)
{
    for (int K = 1; K <= GN; ++K)
        for (int I = 1; I <= GL; ++I) {
            QQ(I, GM + 1, K, 1) = 0.0;
            QQ(I, GM + 1, K, 3) = 0.0;
            QQ(I, GM + 1, K, 2) = 0.0;
            QQ(I, 0, K, 1) = 0.0;
            QQ(I, 0, K, 3) = 0.0;
            QQ(I, 1, K, 2) = 0.0;
            QQ(I, 0, K, 2) = 0.0;
        }
}

/* BC_WALL_PR / BC_WALL_DPH (src/BC_WALL.f90:268-408): Neumann ghost copy */
void orc_bc_wall_neumann(const orc_grid *g, double *arr)
{
    for (int K = 1; K <= GN; ++K)
        for (int I = 1; I <= GL; ++I) {
            G3(arr, I, GM + 1, K) = G3(arr, I, GM, K);
            G3(arr, I, 0, K) = G3(arr, I, 1, K);
        }
}

/* DIVG_tg (src/DIVG.f90:21-57) -> RHSLLPHI */
void orc_divg_tg(const orc_grid *g, orc_state *s, int NS, double DT)
{
    double COE0 = 1.0 / (DT * talp(NS));
#pragma omp parallel for schedule(static)
    for (int JC = 1; JC <= GM; ++JC) {
        int JP = JC + 1;
        double COE1 = g->dxi / 1.0, COE2 = g->dyfi[JC - 1] / 1.0;
        for (int KC = 1; KC <= GN; ++KC) {
            int KP = KPV(KC);
            for (int IC = 1; IC <= GL; ++IC) {
                int IP = IPV(IC);
                double D = (QQ(IP, JC, KC, 1) - QQ(IC, JC, KC, 1)) * COE1 + (QQ(IC, JP, KC, 2) - QQ(IC, JC, KC, 2)) * COE2 +
                           (QQ(IC, JC, KP, 3) - QQ(IC, JC, KC, 3)) * g->dzi;
                F3(s->rhsllphi, IC, JC, KC) = D * COE0;
            }
        }
    }
}

/* VELOUPDT_tg (src/VELO_CORRECTION.f90:19-75) */
void orc_veloupdt_tg(const orc_grid *g, orc_state *s, int NS, double DT)
{
    double COE1 = DT * talp(NS) * g->dxi;
#pragma omp parallel for schedule(static)
    for (int KC = 1; KC <= GN; ++KC)
        for (int JC = 1; JC <= GM; ++JC)
            for (int IC = 1; IC <= GL; ++IC)
                QQ(IC, JC, KC, 1) = QQ(IC, JC, KC, 1) - (G3(s->dph, IC, JC, KC) - G3(s->dph, IMV(IC), JC, KC)) * COE1;
    COE1 = DT * talp(NS) * g->dzi;
#pragma omp parallel for schedule(static)
    for (int KC = 1; KC <= GN; ++KC)
        for (int JC = 1; JC <= GM; ++JC)
            for (int IC = 1; IC <= GL; ++IC)
                QQ(IC, JC, KC, 3) = QQ(IC, JC, KC, 3) - (G3(s->dph, IC, JC, KC) - G3(s->dph, IC, JC, KMV(KC))) * COE1;
#pragma omp parallel for schedule(static)
    for (int JC = 2; JC <= GM; ++JC) {
        double C1 = DT * talp(NS) * g->dyci[JC - 1] / 1.0;
        for (int KC = 1; KC <= GN; ++KC)
            for (int IC = 1; IC <= GL; ++IC)
                QQ(IC, JC, KC, 2) = QQ(IC, JC, KC, 2) - (G3(s->dph, IC, JC, KC) - G3(s->dph, IC, JC - 1, KC)) * C1;
    }
}

/* PRCALC_tg (src/PRESSURE_CORRECTION.f90:1-20) */
void orc_prcalc_tg(const orc_grid *g, orc_state *s, int NS, double DT)
{
    double BE = 0.5 * talp(NS) * DT * g->cvisc;
#pragma omp parallel for schedule(static)
    for (int KC = 1; KC <= GN; ++KC)
        for (int JC = 1; JC <= GM; ++JC)
            for (int IC = 1; IC <= GL; ++IC)
                G3(s->pr, IC, JC, KC) = G3(s->pr, IC, JC, KC) + G3(s->dph, IC, JC, KC) - BE * F3(s->rhsllphi, IC, JC, KC);
}

/* DIVGCK_tg (src/DIVGCK.f90:16-58): max |div u| */
double orc_divgck_tg(const orc_grid *g, const orc_state *s)
{
    double DIVX = 0.0;
#pragma omp parallel for reduction(max : DIVX) schedule(static)
    for (int K = 1; K <= GN; ++K)
        for (int J = 1; J <= GM; ++J)
            for (int I = 1; I <= GL; ++I) {
                double D = (QQ(IPV(I), J, K, 1) - QQ(I, J, K, 1)) * g->dxi +
                           (QQ(I, J + 1, K, 2) - QQ(I, J, K, 2)) * g->dyfi[J - 1] * 1.0 +
                           (QQ(I, J, KPV(K), 3) - QQ(I, J, K, 3)) * g->dzi * 1.0;
                if (fabs(D) > DIVX) DIVX = fabs(D);
            }
    return DIVX;
}

/* CFL_tg (src/CFL.f90:21-59): max over cells of the convective CFL rate */
double orc_cfl_tg(const orc_grid *g, const orc_state *s)
{
    double CFLMA = 0.0;
#pragma omp parallel for reduction(max : CFLMA) schedule(static)
    for (int K = 1; K <= GN; ++K)
        for (int J = 1; J <= GM; ++J)
            for (int I = 1; I <= GL; ++I) {
                double QCF = (fabs((QQ(I, J, K, 1) + QQ(IPV(I), J, K, 1)) * g->dxi) +
                              fabs((QQ(I, J, K, 2) + QQ(I, J + 1, K, 2)) * g->dyfi[J - 1] * 1.0) +
                              fabs((QQ(I, J, K, 3) + QQ(I, J, KPV(K), 3)) * g->dzi * 1.0)) * 0.5;
                if (QCF > CFLMA) CFLMA = QCF;
            }
    return CFLMA;
}

/* xz-mean statistics of PP_MEAN_ZX_FLOW_TG (src/processing/PP_MEAN_ZX.f90:
 * 104-237), channel (RCCI1 = 1), first and second moments only: per row J
 * (out[13*(J-1) + m]):
 *   m = 0..2  U1xz(J,1:3) = VL1313/2 sum (cell-centred u, v, w sums)
 *   m = 3     U1xz(J,4)   = VL1313 sum p
 *   m = 4..9  U2xz(J,1:6) = VL1313/4 sum Ucct(M) Ucct(N), M <= N
 *   m = 10..12 UPxz(J,1:3) = VL1313/2 sum Ucct(M) p
 * with Ucct(1) = u(I)+u(I+1), Ucct(2) = v(J)+v(J+1), Ucct(3) = w(K)+w(K+1). */
void orc_mean_xz(const orc_grid *g, const orc_state *s, double *out)
{
    const double VL = 1.0 / (double)(GL * GN);
    const double COE0 = VL, COE1 = VL * 0.5, COE2 = VL * (0.5 * 0.5);
#pragma omp parallel for schedule(static)
    for (int J = 1; J <= GM; ++J) {
        double u1[4] = {0, 0, 0, 0}, u2[6] = {0, 0, 0, 0, 0, 0}, up[3] = {0, 0, 0};
        for (int I = 1; I <= GL; ++I) {
            const int IP = IPV(I);
            for (int K = 1; K <= GN; ++K) {
                const int KP = KPV(K);
                double U[3];
                U[0] = QQ(I, J, K, 1) + QQ(IP, J, K, 1);
                U[2] = (QQ(I, J, K, 3) + QQ(I, J, KP, 3)) * 1.0;
                U[1] = (QQ(I, J, K, 2) + QQ(I, J + 1, K, 2)) * 1.0;
                const double P = G3(s->pr, I, J, K);
                for (int M = 0; M < 3; ++M) {
                    u1[M] += U[M];
                    up[M] += U[M] * P;
                }
                u1[3] += P;
                int L = 0;
                for (int M = 0; M < 3; ++M)
                    for (int N = M; N < 3; ++N) u2[L++] += U[M] * U[N];
            }
        }
        double *o = out + 13 * (J - 1);
        o[3] = u1[3] * COE0;
        for (int M = 0; M < 3; ++M) o[M] = u1[M] * COE1;
        for (int L = 0; L < 6; ++L) o[4 + L] = u2[L] * COE2;
        for (int M = 0; M < 3; ++M) o[10 + M] = up[M] * COE1;
    }
}

/* SOLVERRK3_MOM_tg(NS) */
void orc_substep_tg(const orc_grid *g, orc_state *s, int NS, double DT, int flowdrvtp, int nflow, double cfgv)
{
    orc_convection_x(g, s);
    orc_convection_y(g, s);
    orc_convection_z(g, s);
    for (int IDR = 1; IDR <= 3; ++IDR) {
        orc_rhs_tg(g, s, NS, IDR, DT, flowdrvtp, nflow, cfgv);
        orc_momfa_tg(g, s, NS, IDR, DT);
    }
    orc_bc_wall_q(g, s);
    orc_divg_tg(g, s, NS, DT);
    orc_poisson_fft99(g, s->rhsllphi, s->dph, g->icase == IBOX3P);
    orc_bc_wall_neumann(g, s->dph);
    orc_veloupdt_tg(g, s, NS, DT);
    orc_bc_wall_q(g, s);
    orc_prcalc_tg(g, s, NS, DT);
    orc_bc_wall_neumann(g, s->pr);
}

#undef GL
#undef GM
#undef GN

/* ---------------------------------------------------------------------------
 * Thermophysical property tables.
 * CUBIC_SPLINE (src/Flow_State.f90:1693-1788): Forsythe-Malcolm-Moler spline
 *   with end conditions from third divided differences.
 * orc_sleval: NIST_SLEVAL_* (src/Flow_State.f90:1015-1070, 1420-1530):
 *   binary search for x(i) <= u < x(i+1), then the cubic.  Out of range
 *   returns status 1 (the reference STOPs).  'hack' selects the post-fix:
 *   0 none, 1: |u| < 1e-10 -> 1 (HT), 2: |u-1| < 1e-10 -> 1 (TD/TM/TK),
 *   3: |u-1| < 1e-10 -> 0 (TH).
 * ------------------------------------------------------------------------- */
void orc_cubic_spline(int N, const double *X, const double *Y, double *B, double *C, double *D)
{
    int NM1 = N - 1;
    if (N < 2) return;
    if (N < 3) {
        B[0] = (Y[1] - Y[0]) / (X[1] - X[0]);
        C[0] = 0.0; D[0] = 0.0;
        B[1] = B[0]; C[1] = 0.0; D[1] = 0.0;
        return;
    }
    /* Fortran 1-based -> macros */
#define X_(i) X[(i)-1]
#define Y_(i) Y[(i)-1]
#define B_(i) B[(i)-1]
#define C_(i) C[(i)-1]
#define D_(i) D[(i)-1]
    D_(1) = X_(2) - X_(1);
    C_(2) = (Y_(2) - Y_(1)) / D_(1);
    for (int I = 2; I <= NM1; ++I) {
        D_(I) = X_(I + 1) - X_(I);
        B_(I) = 2.0 * (D_(I - 1) + D_(I));
        C_(I + 1) = (Y_(I + 1) - Y_(I)) / D_(I);
        C_(I) = C_(I + 1) - C_(I);
    }
    B_(1) = -D_(1);
    B_(N) = -D_(N - 1);
    C_(1) = 0.0;
    C_(N) = 0.0;
    if (N != 3) {
        C_(1) = C_(3) / (X_(4) - X_(2)) - C_(2) / (X_(3) - X_(1));
        C_(N) = C_(N - 1) / (X_(N) - X_(N - 2)) - C_(N - 2) / (X_(N - 1) - X_(N - 3));
        C_(1) = C_(1) * D_(1) * D_(1) / (X_(4) - X_(1));
        C_(N) = -C_(N) * D_(N - 1) * D_(N - 1) / (X_(N) - X_(N - 3));
    }
    for (int I = 2; I <= N; ++I) {
        double T = D_(I - 1) / B_(I - 1);
        B_(I) = B_(I) - T * D_(I - 1);
        C_(I) = C_(I) - T * C_(I - 1);
    }
    C_(N) = C_(N) / B_(N);
    for (int L = 1; L <= NM1; ++L) {
        int I = N - L;
        C_(I) = (C_(I) - D_(I) * C_(I + 1)) / B_(I);
    }
    B_(N) = (Y_(N) - Y_(NM1)) / D_(NM1) + D_(NM1) * (C_(NM1) + 2.0 * C_(N));
    for (int I = 1; I <= NM1; ++I) {
        B_(I) = (Y_(I + 1) - Y_(I)) / D_(I) - D_(I) * (C_(I + 1) + 2.0 * C_(I));
        D_(I) = (C_(I + 1) - C_(I)) / D_(I);
        C_(I) = 3.0 * C_(I);
    }
    C_(N) = 3.0 * C_(N);
    D_(N) = D_(NM1);
#undef X_
#undef Y_
#undef B_
#undef C_
#undef D_
}

int orc_sleval(int n, const double *x, const double *y, const double *b, const double *c, const double *d,
               double u, int hack, double *eval)
{
    if (u <= x[0] || u >= x[n - 1]) return 1;
    int i = 1, j = n + 1;
    while (j > i + 1) {
        int k = (i + j) / 2;
        if (u < x[k - 1]) j = k;
        else i = k;
    }
    double dx = u - x[i - 1];
    double e = y[i - 1] + dx * (b[i - 1] + dx * (c[i - 1] + dx * d[i - 1]));
    if (hack == 1 && fabs(u) < 1.0e-10) e = 1.0;
    if (hack == 2 && fabs(u - 1.0) < 1.0e-10) e = 1.0;
    if (hack == 3 && fabs(u - 1.0) < 1.0e-10) e = 0.0;
    *eval = e;
    return 0;
}

/* THERM_PROP_UPDATE (src/THERM_PROP_UPDATE.f90:129-205) over a flat list of
 * cells, plus c_p = S_TCp(T) (NIST_SLEVAL_TCp) for the surrogate's fifth
 * output.  Table arrays (each length n): H, T, D, M, K, CP and the spline
 * coefficient triples of H->T and T->{D, M, K, CP}:
 *   tab = [H T D M K CP | HTb HTc HTd | TDb TDc TDd | TMb TMc TMd |
 *          TKb TKc TKd | TCPb TCPc TCPd]  (21 n doubles)
 * Cells with |h| < 1e-10 leave their outputs untouched, as the reference
 * does.  Returns the number of out-of-range cells (the reference STOPs);
 * those cells are left untouched too. */
long orc_therm_prop_update(int n, const double *tab, long ncell, const double *h,
                           double *T, double *rho, double *mu, double *kap, double *cp)
{
    const double *H = tab, *TT = tab + n, *DD = tab + 2 * n, *MM = tab + 3 * n, *KK = tab + 4 * n,
                 *CP = tab + 5 * n;
    const double *ht = tab + 6 * n, *td = tab + 9 * n, *tm = tab + 12 * n, *tk = tab + 15 * n,
                 *tc = tab + 18 * n;
    long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (long q = 0; q < ncell; ++q) {
        double hv = h[q];
        if (fabs(hv) < 1.0e-10) continue;
        double t, dv, mv, kv, cv;
        if (orc_sleval(n, H, TT, ht, ht + n, ht + 2 * n, hv, 1, &t)) { bad++; continue; }
        int e1 = orc_sleval(n, TT, MM, tm, tm + n, tm + 2 * n, t, 2, &mv);
        int e2 = orc_sleval(n, TT, KK, tk, tk + n, tk + 2 * n, t, 2, &kv);
        int e3 = orc_sleval(n, TT, DD, td, td + n, td + 2 * n, t, 2, &dv);
        int e4 = orc_sleval(n, TT, CP, tc, tc + n, tc + 2 * n, t, 2, &cv);   /* NIST_SLEVAL_TCp, src/Flow_State.f90:1686 */
        if (e1 | e2 | e3 | e4) { bad++; continue; }
        T[q] = t; mu[q] = mv; kap[q] = kv; rho[q] = dv;
        if (cp) cp[q] = cv;
    }
    return bad;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* threads of the following parallel regions (the timed CPU baseline's 1-thread figure) */
void orc_set_num_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

/* ---------------------------------------------------------------------------
 * Energy substep of the heated channel, one rank, periodic x and z, iso-flux
 * walls (BCWALLHEAT = isoFluxWall at both walls), property table lookup:
 *   SOLVERRK3_ENG_IO (src/SOLVERRK3_ENG_IO.f90:1-77):
 *     VISCOUSITY0 = VISCOUSITY, DENSITY1 = DENSITY0, DENSITY0 = DENSITY
 *     RHS_ENERGY_EXPLICIT (src/RHS_ENERGY_EXPLICIT.f90:1-150)
 *     RHOH += TGAM*DT*RHS + TROH*DT*RHS0, RHOH = min(RHOHmax + REALMIN, RHOH);
 *     RHS0 = RHS
 *     ENTHALPY_UPDATE, DENSITY_UPDATE, ENTHALPY_UPDATE, THERM_PROP_UPDATE
 *     (src/THERM_PROP_UPDATE.f90:1-33, 70-128, 129-205)
 *     INTFC_MFD_THERMAL_io (src/INTFC_THERMAL_io.f90): on one rank every
 *     field it exchanges is then overwritten at the walls, so it is omitted
 *     BC_WALL_THERMAL -> BC_WALL_ISOFLUX top, then bottom
 *     (src/BC_WALL.f90:522-650)
 * Channel metrics: RCCI1 = RNDI1 = RCCI2 = 1, XND2CL = ZND2CL = 1/2.
 * Arrays (L, 0:M+1, N): rhoh h T rho rho0 rho1 mu mu0 kap; rhs0 (L, M, N);
 * g, g0 (L, 0:M+1, N, 3); wallflux (L, 2).  tab: as orc_therm_prop_update.
 * Returns the number of table lookups out of range (the reference STOPs).
 * ------------------------------------------------------------------------- */
typedef struct {
    double *rhoh, *h, *T, *rho, *rho0, *rho1, *mu, *mu0, *kap, *rhs0;
    const double *g, *g0, *wallflux;
} orc_thermal;

#define TL (g->L)
#define TM (g->M)
#define TN (g->N)
#define T3(arr, i, j, k) (arr)[(size_t)((i)-1) + (size_t)TL * ((size_t)(j) + (size_t)(TM + 2) * ((k)-1))]
#define TF3(arr, i, j, k) (arr)[(size_t)((i)-1) + (size_t)TL * ((size_t)((j)-1) + (size_t)TM * ((k)-1))]
#define TG(arr, i, j, k, c) (arr)[(size_t)((i)-1) + (size_t)TL * ((size_t)(j) + (size_t)(TM + 2) * ((size_t)((k)-1) + (size_t)TN * ((c)-1)))]
#define TIPV(i) ((i) == TL ? 1 : (i) + 1)
#define TIMV(i) ((i) == 1 ? TL : (i)-1)
#define TKPV(k) ((k) == TN ? 1 : (k) + 1)
#define TKMV(k) ((k) == 1 ? TN : (k)-1)

static int sl_ht(int n, const double *tab, double h, double *T)
{
    return orc_sleval(n, tab, tab + n, tab + 6 * n, tab + 7 * n, tab + 8 * n, h, 1, T);
}
/* T -> D, M, K with the NIST_SLEVAL_TD/TM/TK post-fix */
static int sl_t(int n, const double *tab, double T, double *D, double *Mu, double *K)
{
    int e = 0;
    if (Mu) e |= orc_sleval(n, tab + n, tab + 3 * n, tab + 12 * n, tab + 13 * n, tab + 14 * n, T, 2, Mu);
    if (K) e |= orc_sleval(n, tab + n, tab + 4 * n, tab + 15 * n, tab + 16 * n, tab + 17 * n, T, 2, K);
    if (D) e |= orc_sleval(n, tab + n, tab + 2 * n, tab + 9 * n, tab + 10 * n, tab + 11 * n, T, 2, D);
    return e;
}

long orc_energy_substep(const orc_grid *g, orc_thermal *t, int n, const double *tab, int NS, double DT,
                        double CTHECD, double RHOHMAX)
{
    const size_t nall = (size_t)TL * (TM + 2) * TN;
    long bad = 0;
    memcpy(t->mu0, t->mu, nall * sizeof(double));
    memcpy(t->rho1, t->rho0, nall * sizeof(double));
    memcpy(t->rho0, t->rho, nall * sizeof(double));
    /* RHS_ENERGY_EXPLICIT + the RK3 update of RHOH */
    const double COE1 = g->dxi * 0.5 * 0.5;
    const double COE11 = g->dxqi * 0.5 * CTHECD;
    const double RK1 = TGAM[NS] * DT, RK0 = TROH[NS] * DT;
#pragma omp parallel for schedule(static)
    for (int KC = 1; KC <= TN; ++KC) {
        const int KP = TKPV(KC), KM = TKMV(KC);
        for (int JC = 1; JC <= TM; ++JC) {
            const int JP = JC + 1, JM = JC - 1, JJ = JC, JJP = JJ + 1;
            const double COE2 = g->dyfi[JJ - 1] * 1.0 * 0.5;
            const double COE3 = g->dzi * 0.5 * 1.0 * 0.5;
            const double COE33 = g->dzqi * 0.5 * 1.0 * CTHECD;
            const double COE22 = g->dyfi[JJ - 1] * 1.0 * CTHECD;
            double COE221, COE222;
            if (JJ == 1) {
                COE221 = g->dyci[JJP - 1] / 1.0;
                COE222 = 1.0 / 1.0;                       /* iso-flux bottom wall */
            } else if (JJ == TM) {
                COE222 = g->dyci[JJ - 1] / 1.0;
                COE221 = 1.0 / 1.0;                       /* iso-flux top wall */
            } else {
                COE221 = g->dyci[JJP - 1] / 1.0;
                COE222 = g->dyci[JJ - 1] / 1.0;
            }
            const double wffp = g->wff[JJP - 1], wfbp = g->wfb[JJP - 1], wff = g->wff[JJ - 1], wfb = g->wfb[JJ - 1];
            for (int IC = 1; IC <= TL; ++IC) {
                const int IP = TIPV(IC), IM = TIMV(IC);
                const double *H = t->h, *TT = t->T, *KK = t->kap;
                double DX_RHOHU = (T3(H, IP, JC, KC) + T3(H, IC, JC, KC)) * (TG(t->g0, IP, JC, KC, 1) + TG(t->g, IP, JC, KC, 1)) -
                                  (T3(H, IC, JC, KC) + T3(H, IM, JC, KC)) * (TG(t->g0, IC, JC, KC, 1) + TG(t->g, IC, JC, KC, 1));
                DX_RHOHU = DX_RHOHU * COE1;
                double DY_RHOHV = (wffp * T3(H, IC, JP, KC) + wfbp * T3(H, IC, JC, KC)) *
                                      (TG(t->g0, IC, JP, KC, 2) + TG(t->g, IC, JP, KC, 2)) -
                                  (wff * T3(H, IC, JC, KC) + wfb * T3(H, IC, JM, KC)) *
                                      (TG(t->g0, IC, JC, KC, 2) + TG(t->g, IC, JC, KC, 2));
                DY_RHOHV = DY_RHOHV * COE2;
                double DZ_RHOHW = (T3(H, IC, JC, KP) + T3(H, IC, JC, KC)) * (TG(t->g0, IC, JC, KP, 3) + TG(t->g, IC, JC, KP, 3)) -
                                  (T3(H, IC, JC, KC) + T3(H, IC, JC, KM)) * (TG(t->g0, IC, JC, KC, 3) + TG(t->g, IC, JC, KC, 3));
                DZ_RHOHW = DZ_RHOHW * COE3;
                double DX_KDT = (T3(KK, IP, JC, KC) + T3(KK, IC, JC, KC)) * (T3(TT, IP, JC, KC) - T3(TT, IC, JC, KC)) -
                                (T3(KK, IC, JC, KC) + T3(KK, IM, JC, KC)) * (T3(TT, IC, JC, KC) - T3(TT, IM, JC, KC));
                DX_KDT = DX_KDT * COE11;
                double KDTy1 = (wffp * T3(KK, IC, JP, KC) + wfbp * T3(KK, IC, JC, KC)) *
                               (T3(TT, IC, JP, KC) - T3(TT, IC, JC, KC)) * COE221;
                double KDTy2 = (wff * T3(KK, IC, JC, KC) + wfb * T3(KK, IC, JM, KC)) *
                               (T3(TT, IC, JC, KC) - T3(TT, IC, JM, KC)) * COE222;
                if (JJ == 1)
                    KDTy2 = -t->wallflux[(IC - 1)] * COE222;
                else if (JJ == TM)
                    KDTy1 = -t->wallflux[(IC - 1) + TL] * COE221;
                const double DY_KDT = (KDTy1 - KDTy2) * COE22;
                double DZ_KDT = (T3(KK, IC, JC, KP) + T3(KK, IC, JC, KC)) * (T3(TT, IC, JC, KP) - T3(TT, IC, JC, KC)) -
                                (T3(KK, IC, JC, KM) + T3(KK, IC, JC, KC)) * (T3(TT, IC, JC, KC) - T3(TT, IC, JC, KM));
                DZ_KDT = DZ_KDT * COE33;
                const double RHS = -DX_RHOHU - DY_RHOHV - DZ_RHOHW + DX_KDT + DY_KDT + DZ_KDT;
                double r = T3(t->rhoh, IC, JC, KC) + RK1 * RHS + RK0 * TF3(t->rhs0, IC, JC, KC);
                r = fmin(RHOHMAX + 1.0e-14, r);
                T3(t->rhoh, IC, JC, KC) = r;
                TF3(t->rhs0, IC, JC, KC) = RHS;
            }
        }
    }
    /* ENTHALPY_UPDATE, DENSITY_UPDATE, ENTHALPY_UPDATE, THERM_PROP_UPDATE (interior) */
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (int KC = 1; KC <= TN; ++KC)
        for (int JC = 1; JC <= TM; ++JC)
            for (int IC = 1; IC <= TL; ++IC) {
                double hv = T3(t->rhoh, IC, JC, KC) / T3(t->rho, IC, JC, KC);
                double Tt, D;
                if (sl_ht(n, tab, hv, &Tt) || sl_t(n, tab, Tt, &D, NULL, NULL)) { bad++; continue; }
                T3(t->rho, IC, JC, KC) = D;
                hv = T3(t->rhoh, IC, JC, KC) / D;
                T3(t->h, IC, JC, KC) = hv;
                if (fabs(hv) < 1.0e-10) continue;
                double Mu, K;
                if (sl_ht(n, tab, hv, &Tt) || sl_t(n, tab, Tt, &D, &Mu, &K)) { bad++; continue; }
                T3(t->T, IC, JC, KC) = Tt;
                T3(t->mu, IC, JC, KC) = Mu;
                T3(t->kap, IC, JC, KC) = K;
                T3(t->rho, IC, JC, KC) = D;
            }
    /* BC_WALL_ISOFLUX: top wall (J = M+1), then bottom wall (J = 0) */
    for (int wall = 0; wall < 2; ++wall) {
        const int J = wall == 0 ? TM + 1 : 0;
        for (int KC = 1; KC <= TN; ++KC)
            for (int IC = 1; IC <= TL; ++IC) {
                double hv;
                if (J == 0)
                    hv = 2.0 * T3(t->h, IC, 1, KC) - g->wff[2 - 1] * T3(t->h, IC, 2, KC) - g->wfb[2 - 1] * T3(t->h, IC, 1, KC);
                else
                    hv = 2.0 * T3(t->h, IC, J - 1, KC) - g->wff[J - 1 - 1] * T3(t->h, IC, J - 1, KC) -
                         g->wfb[J - 1 - 1] * T3(t->h, IC, J - 2, KC);
                T3(t->h, IC, J, KC) = hv;
                double Tt, Mu, K, D;
                if (sl_ht(n, tab, hv, &Tt) || sl_t(n, tab, Tt, &D, &Mu, &K)) { bad++; continue; }
                T3(t->rho, IC, J, KC) = D;
                T3(t->T, IC, J, KC) = Tt;
                T3(t->mu, IC, J, KC) = Mu;
                T3(t->kap, IC, J, KC) = K;
                T3(t->rhoh, IC, J, KC) = hv * D;
            }
    }
    return bad;
}
#undef TL
#undef TM
#undef TN
#undef T3
#undef TF3
#undef TG
#undef TIPV
#undef TIMV
#undef TKPV
#undef TKMV

/* ---------------------------------------------------------------------------
 * The io (main-domain) momentum substep of the heated periodic channel:
 * variable properties, mass fluxes G = rho u, explicit viscous terms
 * (visthemflg = 0, the reference's default), no weighted pressure.
 *
 * orc_io_prepare(NS): what SOLVE runs between SOLVERRK3_ENG_IO(NS) and
 *   SOLVERRK3_MOM_IO(NS) (src/SOLVE.f90:121-133):
 *     DENSITY_TIME_DERIVATION(NS)   src/THERM_PROP_UPDATE.f90:225-291
 *     DENSITY_Staggered             src/MASSFLUX_CALC_IO.f90:242-311
 *     MU_Staggered                  src/MASSFLUX_CALC_IO.f90:147-237
 *     VELOCITY_CALC_io              src/MASSFLUX_CALC_IO.f90:82-142
 *   (the first three only with thermal = 1; otherwise RHO_STG = MU_STG = 1
 *   and DrhoDtP = 0 as allocated, src/MEMALLOCT.f90:197-201).
 * orc_io_substep(NS): SOLVERRK3_MOM_IO (src/SOLVERRK3_MOM_IO.f90:1-168):
 *     G0 = G
 *     CONVECTION_X_io / _Y_io / _Z_io  src/CONVECTION_X.f90:77-176,
 *                                      src/CONVECTION_Y.f90:128-287,
 *                                      src/CONVECTION_Z.f90:115-314
 *     DIVG_U_io                        src/DIVG.f90:153-259
 *     VISCOUS_ALL_EXPLT_X/Y/Z_io       src/VISCOUS_X_EXPLICIT_io.f90:1-113,
 *                                      src/VISCOUS_Y_EXPLICIT_io.f90:1-99,
 *                                      src/VISCOUS_Z_EXPLICIT_io.f90:1-100
 *     RHS_MOM_EXPLICIT_io(NS, IDR)     src/RHS_MOM_EXPLICIT_io.f90:1-297
 *     MASSFLUX_CALC_IO                 src/MASSFLUX_CALC_IO.f90:1-79
 *     BC_WALL_G_io                     src/BC_WALL.f90:412-458, 682-694
 *     DIVG_io(NS)                      src/DIVG.f90:60-150
 *     the FFT99 Poisson solve, the reference's "Method One" for the io
 *     domain (src/SOLVERRK3_MOM_IO.f90:123), then BC_WALL_DPH_io
 *     PRCALC_io (explicit branch)      src/PRESSURE_CORRECTION.f90:88-156
 *     MASSFLUX_UPDATE_IO               src/VELO_CORRECTION.f90:109-185
 * INTFC_VARS1/3 on one GPU is the periodic y wrap that the reference's
 * exchange produces across its ranks (src/INTFC_VARS.f90:41-80): row 0 <-
 * row M, row M+1 <- row 1; wall routines then overwrite what they set.
 * Channel metrics: RCCI1 = RNDI1 = RCCI2 = 1, XND2CL = YND2CL = ZND2CL = 1/2.
 * Arrays: q g g0 rho_stg mu_stg (L, 0:M+1, N, 3); pr rho rho0 mu mu0
 * drhodtp divu dph (L, 0:M+1, N); explt0 (L, M, N, 3); hx hy hz rhsllphi
 * (L, M, N).
 * ------------------------------------------------------------------------- */
typedef struct {
    double *q, *g, *g0, *pr;
    const double *rho, *rho0, *mu, *mu0;
    double *drhodtp, *divu, *rho_stg, *mu_stg, *explt0;
    double *dph, *hx, *hy, *hz, *rhsllphi;
    double dpgrns;
    double *drhoi_stg;   /* DRHOI_STG (L, 0:M+1, N, 3), visimplicit only */
    double *im;          /* implicit viscous parts (RHS_io of VISCOUS_PAR_*), (L, M, N, 3) */
    double *ad;          /* their density corrections (DIVU_io of VISCOUS_PAR_*), (L, M, N, 3) */
} orc_io;

typedef struct {
    int thermal;      /* thermlflg */
    int nflow;        /* NFLOW: streamwise direction 1..3 */
    int flowdrvtp;    /* FLOWDRVTP: 1 constant mass flux, 2 constant pressure gradient */
    int gravdir;      /* GRAVDIR: 0 none, 1..3 */
    double cfgv;      /* CFGV (FLOWDRVTP = 2) */
    double f_a;       /* F_A (src/Flow_State.f90:97-110) */
    int visimplicit;  /* visthemflg: 0 explicit viscous terms, 1 RK-CN (MOMFA_io) */
} orc_io_params;

#define WL (g->L)
#define WM (g->M)
#define WN (g->N)
#define W3(arr, i, j, k) (arr)[(size_t)((i)-1) + (size_t)WL * ((size_t)(j) + (size_t)(WM + 2) * ((k)-1))]
#define WF3(arr, i, j, k) (arr)[(size_t)((i)-1) + (size_t)WL * ((size_t)((j)-1) + (size_t)WM * ((k)-1))]
#define WG(arr, i, j, k, c) (arr)[(size_t)((i)-1) + (size_t)WL * ((size_t)(j) + (size_t)(WM + 2) * ((size_t)((k)-1) + (size_t)WN * ((c)-1)))]
#define WE(arr, i, j, k, c) (arr)[(size_t)((i)-1) + (size_t)WL * ((size_t)((j)-1) + (size_t)WM * ((size_t)((k)-1) + (size_t)WN * ((c)-1)))]
#define WIPV(i) ((i) == WL ? 1 : (i) + 1)
#define WIMV(i) ((i) == 1 ? WL : (i)-1)
#define WKPV(k) ((k) == WN ? 1 : (k) + 1)
#define WKMV(k) ((k) == 1 ? WN : (k)-1)

/* INTFC_VARS on one GPU: the periodic y wrap of a (L, 0:M+1, N) array */
static void io_wrap(const orc_grid *g, double *a)
{
    for (int K = 1; K <= WN; ++K)
        for (int I = 1; I <= WL; ++I) {
            W3(a, I, 0, K) = W3(a, I, WM, K);
            W3(a, I, WM + 1, K) = W3(a, I, 1, K);
        }
}

static void io_wrap3(const orc_grid *g, double *a)
{
    const size_t ngh = (size_t)WL * (WM + 2) * WN;
    for (int c = 0; c < 3; ++c) io_wrap(g, a + c * ngh);
}

/* BC_WALL_Q / BC_WALL_G for the io domain (src/BC_WALL.f90:185-263, 412-458) */
static void io_bc_wall_vec(const orc_grid *g, double *a)
{
    for (int K = 1; K <= WN; ++K)
        for (int I = 1; I <= WL; ++I) {
            WG(a, I, WM + 1, K, 1) = 0.0;
            WG(a, I, WM + 1, K, 3) = 0.0;
            WG(a, I, WM + 1, K, 2) = 0.0;
            WG(a, I, 0, K, 1) = 0.0;
            WG(a, I, 0, K, 3) = 0.0;
            WG(a, I, 1, K, 2) = 0.0;
            WG(a, I, 0, K, 2) = 0.0;
        }
}

void orc_io_prepare(const orc_grid *g, orc_io *s, const orc_io_params *p, int NS)
{
    if (p->thermal) {
        /* DENSITY_TIME_DERIVATION(NS) */
#pragma omp parallel for schedule(static)
        for (int JC = 1; JC <= WM; ++JC) {
            const int JP = JC + 1;
            const double COE2 = g->dyfi[JC - 1] * 1.0, COE3 = g->dzi * 1.0;
            for (int IC = 1; IC <= WL; ++IC) {
                const int IP = WIPV(IC);
                for (int KC = 1; KC <= WN; ++KC) {
                    const int KP = WKPV(KC);
                    double DIVX = (WG(s->g, IP, JC, KC, 1) - WG(s->g, IC, JC, KC, 1)) * g->dxi;
                    double DIVX0 = (WG(s->g0, IP, JC, KC, 1) - WG(s->g0, IC, JC, KC, 1)) * g->dxi;
                    double DIVY = (WG(s->g, IC, JP, KC, 2) - WG(s->g, IC, JC, KC, 2)) * COE2;
                    double DIVY0 = (WG(s->g0, IC, JP, KC, 2) - WG(s->g0, IC, JC, KC, 2)) * COE2;
                    double DIVZ = (WG(s->g, IC, JC, KP, 3) - WG(s->g, IC, JC, KC, 3)) * COE3;
                    double DIVZ0 = (WG(s->g0, IC, JC, KP, 3) - WG(s->g0, IC, JC, KC, 3)) * COE3;
                    W3(s->drhodtp, IC, JC, KC) =
                        ((TGAM[NS] + TROH[NS]) * (DIVX + DIVY + DIVZ) + TROH[NS] * (DIVX0 + DIVY0 + DIVZ0)) /
                        (TGAM[NS] - 2.0);
                }
            }
        }
        /* DENSITY_Staggered */
        const size_t ngh = (size_t)WL * (WM + 2) * WN;
        const double *D = s->rho, *D0 = s->rho0;
        double *RS = s->rho_stg;
#pragma omp parallel for schedule(static)
        for (int JC = 1; JC <= WM; ++JC)
            for (int KC = 1; KC <= WN; ++KC)
                for (int IC = 1; IC <= WL; ++IC) {
                    const int IM = WIMV(IC), KM = WKMV(KC);
                    WG(RS, IC, JC, KC, 1) =
                        (W3(D, IC, JC, KC) + W3(D, IM, JC, KC) + W3(D0, IC, JC, KC) + W3(D0, IM, JC, KC)) * 0.5 * 0.5;
                    WG(RS, IC, JC, KC, 3) =
                        (W3(D, IC, JC, KC) + W3(D, IC, JC, KM) + W3(D0, IC, JC, KC) + W3(D0, IC, JC, KM)) * 0.5 * 0.5;
                    if (JC >= 2) {
                        const int JM = JC - 1;
                        WG(RS, IC, JC, KC, 2) = (g->wff[JC - 1] * (W3(D, IC, JC, KC) + W3(D0, IC, JC, KC)) +
                                                 g->wfb[JC - 1] * (W3(D, IC, JM, KC) + W3(D0, IC, JM, KC))) *
                                                0.5;
                    }
                }
        io_wrap3(g, RS);
        for (int KC = 1; KC <= WN; ++KC)
            for (int IC = 1; IC <= WL; ++IC) {
                WG(RS, IC, 1, KC, 2) = 0.5 * (W3(D0, IC, 0, KC) + W3(D, IC, 0, KC));
                WG(RS, IC, WM + 1, KC, 2) = 0.5 * (W3(D0, IC, WM + 1, KC) + W3(D, IC, WM + 1, KC));
            }
        /* MU_Staggered */
        const double *V = s->mu, *V0 = s->mu0;
        double *MS = s->mu_stg;
#pragma omp parallel for schedule(static)
        for (int JC = 1; JC <= WM; ++JC)
            for (int KC = 1; KC <= WN; ++KC)
                for (int IC = 1; IC <= WL; ++IC) {
                    const int IM = WIMV(IC), KM = WKMV(KC), JM = JC - 1;
                    if (JC >= 2) {
                        WG(MS, IC, JC, KC, 1) =
                            ((W3(V, IC, JC, KC) + W3(V, IM, JC, KC) + W3(V0, IC, JC, KC) + W3(V0, IM, JC, KC)) *
                                 g->wff[JC - 1] +
                             (W3(V, IC, JM, KC) + W3(V, IM, JM, KC) + W3(V0, IC, JM, KC) + W3(V0, IM, JM, KC)) *
                                 g->wfb[JC - 1]) *
                            0.5 * 0.5;
                        WG(MS, IC, JC, KC, 3) =
                            ((W3(V, IC, JC, KM) + W3(V, IC, JC, KC) + W3(V0, IC, JC, KM) + W3(V0, IC, JC, KC)) *
                                 g->wff[JC - 1] +
                             (W3(V, IC, JM, KM) + W3(V, IC, JM, KC) + W3(V0, IC, JM, KM) + W3(V0, IC, JM, KC)) *
                                 g->wfb[JC - 1]) *
                            0.5 * 0.5;
                    }
                    WG(MS, IC, JC, KC, 2) = (W3(V, IC, JC, KC) + W3(V, IM, JC, KC) + W3(V0, IC, JC, KC) +
                                             W3(V0, IM, JC, KC) + W3(V, IC, JC, KM) + W3(V, IM, JC, KM) +
                                             W3(V0, IC, JC, KM) + W3(V0, IM, JC, KM)) *
                                            0.5 * 0.5 * 0.5;
                }
        io_wrap3(g, MS);
        for (int KC = 1; KC <= WN; ++KC)
            for (int IC = 1; IC <= WL; ++IC) {
                const double wb = 0.5 * (W3(V0, IC, 0, KC) + W3(V, IC, 0, KC));
                const double wt = 0.5 * (W3(V0, IC, WM + 1, KC) + W3(V, IC, WM + 1, KC));
                WG(MS, IC, 1, KC, 1) = wb;
                WG(MS, IC, 1, KC, 3) = wb;
                WG(MS, IC, WM + 1, KC, 1) = wt;
                WG(MS, IC, WM + 1, KC, 3) = wt;
            }
        (void)ngh;
        if (p->visimplicit) {
            /* DENSITY_implicit_add (src/MASSFLUX_CALC_IO.f90:316-400) */
            double *DR = s->drhoi_stg;
#pragma omp parallel for schedule(static)
            for (int JC = 1; JC <= WM; ++JC)
                for (int KC = 1; KC <= WN; ++KC)
                    for (int IC = 1; IC <= WL; ++IC) {
                        const int IM = WIMV(IC), KM = WKMV(KC);
                        double DEN_STG = (W3(D0, IC, JC, KC) + W3(D0, IM, JC, KC)) * 0.5;
                        WG(DR, IC, JC, KC, 1) = 1.0 / WG(RS, IC, JC, KC, 1) - 1.0 / DEN_STG;
                        DEN_STG = (W3(D0, IC, JC, KC) + W3(D0, IC, JC, KM)) * 0.5;
                        WG(DR, IC, JC, KC, 3) = 1.0 / WG(RS, IC, JC, KC, 3) - 1.0 / DEN_STG;
                        if (JC >= 2) {
                            DEN_STG = g->wff[JC - 1] * W3(D0, IC, JC, KC) + g->wfb[JC - 1] * W3(D0, IC, JC - 1, KC);
                            WG(DR, IC, JC, KC, 2) = 1.0 / WG(RS, IC, JC, KC, 2) - 1.0 / DEN_STG;
                        }
                    }
            io_wrap3(g, DR);
            for (int KC = 1; KC <= WN; ++KC)
                for (int IC = 1; IC <= WL; ++IC) {
                    WG(DR, IC, 1, KC, 2) = 0.0;
                    WG(DR, IC, WM + 1, KC, 2) = 0.0;
                }
        }
    }
    /* VELOCITY_CALC_io */
#pragma omp parallel for schedule(static)
    for (int JC = 1; JC <= WM; ++JC)
        for (int KC = 1; KC <= WN; ++KC)
            for (int IC = 1; IC <= WL; ++IC) {
                WG(s->q, IC, JC, KC, 1) = WG(s->g, IC, JC, KC, 1) / WG(s->rho_stg, IC, JC, KC, 1);
                WG(s->q, IC, JC, KC, 3) = WG(s->g, IC, JC, KC, 3) / WG(s->rho_stg, IC, JC, KC, 3);
                if (JC >= 2) WG(s->q, IC, JC, KC, 2) = WG(s->g, IC, JC, KC, 2) / WG(s->rho_stg, IC, JC, KC, 2);
            }
    io_wrap3(g, s->q);
    io_bc_wall_vec(g, s->q);
}

/* RHS_MOM_EXPLICIT_io(NS, IDR) on the stage array h (Qtmp, DPH or RHSLLPHI
 * of the reference), explicit branch: h <- RHS. */
static void io_rhs_mom(const orc_grid *g, orc_io *s, const orc_io_params *p, int NS, int IDR, double DT, double *h)
{
    const int NYI = IDR == 2 ? 2 : 1;
    const double TA = talp(NS);
    double *RHS = (double *)calloc((size_t)WL * WM * WN, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int JC = NYI; JC <= WM; ++JC) {
        const int JM = JC - 1;
        for (int KC = 1; KC <= WN; ++KC) {
            const int KM = WKMV(KC);
            for (int IC = 1; IC <= WL; ++IC) {
                const int IM = WIMV(IC);
                const double RHSC = WF3(h, IC, JC, KC), RHSL = WE(s->explt0, IC, JC, KC, IDR);
                double r = (TGAM[NS] * RHSC + TROH[NS] * RHSL) * DT;
                if (p->visimplicit) {
                    const double COE2 = TA * DT;
                    r = r + COE2 * WE(s->im, IC, JC, KC, IDR) + 0.5 * COE2 * WE(s->ad, IC, JC, KC, IDR);
                }
                WE(s->explt0, IC, JC, KC, IDR) = RHSC;
                double PGM;
                if (IDR == 1)
                    PGM = (W3(s->pr, IC, JC, KC) - W3(s->pr, IM, JC, KC)) * (TA * DT * g->dxi);
                else if (IDR == 2)
                    PGM = (W3(s->pr, IC, JC, KC) - W3(s->pr, IC, JM, KC)) * (TA * DT * g->dyci[JC - 1] / 1.0);
                else
                    PGM = (W3(s->pr, IC, JC, KC) - W3(s->pr, IC, JC, KM)) * (TA * DT * g->dzi);
                r = r - PGM;
                if (p->gravdir == IDR) r = r + WG(s->rho_stg, IC, JC, KC, IDR) * (TA * DT * p->f_a / 1.0);
                WF3(RHS, IC, JC, KC) = r;
            }
        }
    }
    if (IDR == p->nflow) {
        double DPGRNS = 0.0;
        if (p->flowdrvtp == 1) {
            double INTGRHSY = 0.0;
            for (int JC = 1; JC <= WM; ++JC)
                for (int IC = 1; IC <= WL; ++IC)
                    for (int KC = 1; KC <= WN; ++KC) INTGRHSY = INTGRHSY + WF3(RHS, IC, JC, KC) / g->dyfi[JC - 1] / 1.0;
            DPGRNS = INTGRHSY / g->volm;
        } else if (p->flowdrvtp == 2) {
            DPGRNS = -0.5 * p->cfgv * (TA * DT);
        }
        s->dpgrns = DPGRNS;
        for (int JC = NYI; JC <= WM; ++JC)
            for (int KC = 1; KC <= WN; ++KC)
                for (int IC = 1; IC <= WL; ++IC) WF3(RHS, IC, JC, KC) = WF3(RHS, IC, JC, KC) - DPGRNS;
    }
    for (int JC = NYI; JC <= WM; ++JC)
        for (int KC = 1; KC <= WN; ++KC)
            for (int IC = 1; IC <= WL; ++IC) WF3(h, IC, JC, KC) = WF3(RHS, IC, JC, KC);
    free(RHS);
}

/* CONVECTION_X/Y/Z_io -> hx, hy (row 1 zero), hz */
static void io_convection(const orc_grid *g, orc_io *s)
{
    const size_t ncell = (size_t)WL * WM * WN;
    memset(s->hy, 0, ncell * sizeof(double));
#pragma omp parallel for schedule(static)
    for (int JC = 1; JC <= WM; ++JC) {
        const int JM = JC - 1, JP = JC + 1, JJ = JC, JJP = JC + 1;
        const double wff = g->wff[JJ - 1], wfb = g->wfb[JJ - 1], wffp = g->wff[JJP - 1], wfbp = g->wfb[JJP - 1];
        for (int KC = 1; KC <= WN; ++KC) {
            const int KM = WKMV(KC), KP = WKPV(KC);
            for (int IC = 1; IC <= WL; ++IC) {
                const int IP = WIPV(IC), IM = WIMV(IC);
                const double *G = s->g, *Q = s->q;
                /* CONVECTION_X_io */
                {
                    const double COE1 = g->dxi * 0.5 * 0.5, COE2 = g->dyfi[JJ - 1] * 0.5 * 1.0;
                    const double COE32 = g->dzi * 0.5 * 0.5 * 1.0;
                    double H11F = (WG(G, IP, JC, KC, 1) + WG(G, IC, JC, KC, 1)) * (WG(Q, IP, JC, KC, 1) + WG(Q, IC, JC, KC, 1));
                    double H11B = (WG(G, IC, JC, KC, 1) + WG(G, IM, JC, KC, 1)) * (WG(Q, IC, JC, KC, 1) + WG(Q, IM, JC, KC, 1));
                    double H11 = (H11F - H11B) * COE1;
                    double H12F = (WG(G, IC, JP, KC, 2) + WG(G, IM, JP, KC, 2)) *
                                  (wffp * WG(Q, IC, JP, KC, 1) + wfbp * WG(Q, IC, JC, KC, 1));
                    double H12B = (WG(G, IC, JC, KC, 2) + WG(G, IM, JC, KC, 2)) *
                                  (wff * WG(Q, IC, JC, KC, 1) + wfb * WG(Q, IC, JM, KC, 1));
                    double H12 = (H12F - H12B) * COE2;
                    double H13F = (WG(G, IC, JC, KP, 3) + WG(G, IM, JC, KP, 3)) * (WG(Q, IC, JC, KP, 1) + WG(Q, IC, JC, KC, 1));
                    double H13B = (WG(G, IC, JC, KC, 3) + WG(G, IM, JC, KC, 3)) * (WG(Q, IC, JC, KC, 1) + WG(Q, IC, JC, KM, 1));
                    double H13 = (H13F - H13B) * COE32;
                    WF3(s->hx, IC, JC, KC) = -(H11 + H12 + H13);
                }
                /* CONVECTION_Y_io (rows 2..M; row 1 stays zero) */
                if (JC >= 2) {
                    const double COE1 = g->dxi * 0.5, COE2 = g->dyci[JJ - 1] * 0.5 * 0.5, COE31 = g->dzi * 0.5 * 1.0;
                    double H21F = (wff * WG(G, IP, JC, KC, 1) + wfb * WG(G, IP, JM, KC, 1)) *
                                  (WG(Q, IP, JC, KC, 2) + WG(Q, IC, JC, KC, 2));
                    double H21B = (wff * WG(G, IC, JC, KC, 1) + wfb * WG(G, IC, JM, KC, 1)) *
                                  (WG(Q, IC, JC, KC, 2) + WG(Q, IM, JC, KC, 2));
                    double H21 = (H21F - H21B) * COE1;
                    double H22F = (WG(G, IC, JP, KC, 2) + WG(G, IC, JC, KC, 2)) * (WG(Q, IC, JP, KC, 2) + WG(Q, IC, JC, KC, 2));
                    double H22B = (WG(G, IC, JC, KC, 2) + WG(G, IC, JM, KC, 2)) * (WG(Q, IC, JC, KC, 2) + WG(Q, IC, JM, KC, 2));
                    double H22 = (H22F - H22B) * COE2;
                    double H23F = (wff * WG(G, IC, JC, KP, 3) + wfb * WG(G, IC, JM, KP, 3)) *
                                  (WG(Q, IC, JC, KP, 2) + WG(Q, IC, JC, KC, 2));
                    double H23B = (wff * WG(G, IC, JC, KC, 3) + wfb * WG(G, IC, JM, KC, 3)) *
                                  (WG(Q, IC, JC, KC, 2) + WG(Q, IC, JC, KM, 2));
                    double H23 = (H23F - H23B) * COE31;
                    WF3(s->hy, IC, JC, KC) = -(H21 + H22 + H23);
                }
                /* CONVECTION_Z_io */
                {
                    const double COE1 = g->dxi * 0.5 * 0.5, COE2 = g->dyfi[JJ - 1] * 0.5;
                    const double COE31 = g->dzi * 0.5 * 0.5 * 1.0;
                    double H31F = (WG(G, IP, JC, KC, 1) + WG(G, IP, JC, KM, 1)) * (WG(Q, IP, JC, KC, 3) + WG(Q, IC, JC, KC, 3));
                    double H31B = (WG(G, IC, JC, KC, 1) + WG(G, IC, JC, KM, 1)) * (WG(Q, IC, JC, KC, 3) + WG(Q, IM, JC, KC, 3));
                    double H31 = (H31F - H31B) * COE1;
                    double H32F = (WG(G, IC, JP, KC, 2) + WG(G, IC, JP, KM, 2)) *
                                  (wffp * WG(Q, IC, JP, KC, 3) + wfbp * WG(Q, IC, JC, KC, 3));
                    double H32B = (WG(G, IC, JC, KC, 2) + WG(G, IC, JC, KM, 2)) *
                                  (wff * WG(Q, IC, JC, KC, 3) + wfb * WG(Q, IC, JM, KC, 3));
                    double H32 = (H32F - H32B) * COE2;
                    double H33F = (WG(G, IC, JC, KP, 3) + WG(G, IC, JC, KC, 3)) * (WG(Q, IC, JC, KP, 3) + WG(Q, IC, JC, KC, 3));
                    double H33B = (WG(G, IC, JC, KC, 3) + WG(G, IC, JC, KM, 3)) * (WG(Q, IC, JC, KC, 3) + WG(Q, IC, JC, KM, 3));
                    double H33 = (H33F - H33B) * COE31;
                    WF3(s->hz, IC, JC, KC) = -(H31 + H32 + H33);
                }
            }
        }
    }
}

/* The explicit terms: CONVECTION_X/Y/Z_io, DIVG_U_io and
 * VISCOUS_ALL_EXPLT_X/Y/Z_io -> hx, hy (row 1 zero), hz. */
void orc_io_terms(const orc_grid *g, orc_io *s)
{
    const double CV = g->cvisc;
    io_convection(g, s);
    /* DIVG_U_io */
#pragma omp parallel for schedule(static)
    for (int JC = 1; JC <= WM; ++JC) {
        const int JP = JC + 1, JJ = JC;
        for (int KC = 1; KC <= WN; ++KC) {
            const int KP = WKPV(KC);
            for (int IC = 1; IC <= WL; ++IC) {
                const int IP = WIPV(IC);
                const double *Q = s->q;
                double DIVX = (WG(Q, IP, JC, KC, 1) - WG(Q, IC, JC, KC, 1)) * g->dxi;
                double DIVY = (WG(Q, IC, JP, KC, 2) - WG(Q, IC, JC, KC, 2)) * (g->dyfi[JJ - 1] * 1.0);
                double DIVZ = (WG(Q, IC, JC, KP, 3) - WG(Q, IC, JC, KC, 3)) * (g->dzi * 1.0);
                W3(s->divu, IC, JC, KC) = (DIVX + DIVY + DIVZ) / 3.0;
            }
        }
    }
    io_wrap(g, s->divu);
    /* VISCOUS_ALL_EXPLT_X_io, _Y_io, _Z_io */
#pragma omp parallel for schedule(static)
    for (int JC = 1; JC <= WM; ++JC) {
        const int JM = JC - 1, JP = JC + 1, JJ = JC, JJP = JC + 1, JJM = JC - 1;
        for (int KC = 1; KC <= WN; ++KC) {
            const int KM = WKMV(KC), KP = WKPV(KC);
            for (int IC = 1; IC <= WL; ++IC) {
                const int IP = WIPV(IC), IM = WIMV(IC);
                const double *Q = s->q, *MU = s->mu, *MU0 = s->mu0, *DU = s->divu, *MS = s->mu_stg;
                {
                    const double COE1 = 2.0 * g->dxi * CV, COE2 = g->dyfi[JJ - 1] * CV * 1.0, COE32 = g->dzi * CV * 1.0;
                    double DUDX = (WG(Q, IP, JC, KC, 1) - WG(Q, IC, JC, KC, 1)) * g->dxi;
                    double TAU11F = 0.5 * (W3(MU0, IC, JC, KC) + W3(MU, IC, JC, KC)) * (DUDX - W3(DU, IC, JC, KC));
                    DUDX = (WG(Q, IC, JC, KC, 1) - WG(Q, IM, JC, KC, 1)) * g->dxi;
                    double TAU11B = 0.5 * (W3(MU0, IM, JC, KC) + W3(MU, IM, JC, KC)) * (DUDX - W3(DU, IM, JC, KC));
                    double DTAU11DX = (TAU11F - TAU11B) * COE1;
                    double DUDY = (WG(Q, IC, JP, KC, 1) - WG(Q, IC, JC, KC, 1)) * g->dyci[JJP - 1] / 1.0;
                    double DVDX = (WG(Q, IC, JP, KC, 2) - WG(Q, IM, JP, KC, 2)) * g->dxi;
                    double TAU12F = WG(MS, IC, JP, KC, 1) * (DUDY + DVDX);
                    DUDY = (WG(Q, IC, JC, KC, 1) - WG(Q, IC, JM, KC, 1)) * g->dyci[JJ - 1] / 1.0;
                    DVDX = (WG(Q, IC, JC, KC, 2) - WG(Q, IM, JC, KC, 2)) * g->dxi;
                    double TAU12B = WG(MS, IC, JC, KC, 1) * (DUDY + DVDX);
                    double DTAU12DY = (TAU12F - TAU12B) * COE2;
                    double DUDZ = (WG(Q, IC, JC, KP, 1) - WG(Q, IC, JC, KC, 1)) * g->dzi;
                    double DWDX = (WG(Q, IC, JC, KP, 3) - WG(Q, IM, JC, KP, 3)) * g->dxi;
                    double TAU13F = WG(MS, IC, JC, KP, 2) * (DUDZ + DWDX);
                    DUDZ = (WG(Q, IC, JC, KC, 1) - WG(Q, IC, JC, KM, 1)) * g->dzi;
                    DWDX = (WG(Q, IC, JC, KC, 3) - WG(Q, IM, JC, KC, 3)) * g->dxi;
                    double TAU13B = WG(MS, IC, JC, KC, 2) * (DUDZ + DWDX);
                    double DTAU13DZ = (TAU13F - TAU13B) * COE32;
                    WF3(s->hx, IC, JC, KC) = WF3(s->hx, IC, JC, KC) + DTAU11DX + DTAU12DY + DTAU13DZ;
                }
                if (JC >= 2) {
                    const double COE1 = g->dxi * CV, COE3 = g->dzi * CV, COE2 = 2.0 * g->dyci[JJ - 1] * CV;
                    double DVDY = (WG(Q, IC, JP, KC, 2) - WG(Q, IC, JC, KC, 2)) * g->dyfi[JJ - 1];
                    double TAU22F = 0.5 * (W3(MU0, IC, JC, KC) + W3(MU, IC, JC, KC)) * (DVDY - W3(DU, IC, JC, KC));
                    DVDY = (WG(Q, IC, JC, KC, 2) - WG(Q, IC, JM, KC, 2)) * g->dyfi[JJM - 1];
                    double TAU22B = 0.5 * (W3(MU0, IC, JM, KC) + W3(MU, IC, JM, KC)) * (DVDY - W3(DU, IC, JM, KC));
                    double DTAU22DY = (TAU22F - TAU22B) * COE2;
                    double DVDX = (WG(Q, IP, JC, KC, 2) - WG(Q, IC, JC, KC, 2)) * g->dxi;
                    double DUDY = (WG(Q, IP, JC, KC, 1) - WG(Q, IP, JM, KC, 1)) * g->dyci[JJ - 1];
                    double TAU21F = WG(MS, IP, JC, KC, 1) * (DVDX + DUDY);
                    DVDX = (WG(Q, IC, JC, KC, 2) - WG(Q, IM, JC, KC, 2)) * g->dxi;
                    DUDY = (WG(Q, IC, JC, KC, 1) - WG(Q, IC, JM, KC, 1)) * g->dyci[JJ - 1];
                    double TAU21B = WG(MS, IC, JC, KC, 1) * (DVDX + DUDY);
                    double DTAU21DX = (TAU21F - TAU21B) * COE1;
                    double DVDZ = (WG(Q, IC, JC, KP, 2) - WG(Q, IC, JC, KC, 2)) * g->dzi;
                    double DWDY = (WG(Q, IC, JC, KP, 3) - WG(Q, IC, JM, KP, 3)) * g->dyci[JJ - 1];
                    double TAU23F = WG(MS, IC, JC, KP, 3) * (DVDZ + DWDY);
                    DVDZ = (WG(Q, IC, JC, KC, 2) - WG(Q, IC, JC, KM, 2)) * g->dzi;
                    DWDY = (WG(Q, IC, JC, KC, 3) - WG(Q, IC, JM, KC, 3)) * g->dyci[JJ - 1];
                    double TAU23B = WG(MS, IC, JC, KC, 3) * (DVDZ + DWDY);
                    double DTAU23DZ = (TAU23F - TAU23B) * COE3;
                    WF3(s->hy, IC, JC, KC) = WF3(s->hy, IC, JC, KC) + DTAU21DX + DTAU22DY + DTAU23DZ;
                }
                {
                    const double COE1 = g->dxi * CV, COE3 = 2.0 * g->dzi * CV, COE2 = g->dyfi[JJ - 1] * CV;
                    double DWDZ = (WG(Q, IC, JC, KP, 3) - WG(Q, IC, JC, KC, 3)) * g->dzi;
                    double TAU33F = 0.5 * (W3(MU0, IC, JC, KC) + W3(MU, IC, JC, KC)) * (DWDZ - W3(DU, IC, JC, KC));
                    DWDZ = (WG(Q, IC, JC, KC, 3) - WG(Q, IC, JC, KM, 3)) * g->dzi;
                    double TAU33B = 0.5 * (W3(MU0, IC, JC, KM) + W3(MU, IC, JC, KM)) * (DWDZ - W3(DU, IC, JC, KM));
                    double DTAU33DZ = (TAU33F - TAU33B) * COE3;
                    double DWDX = (WG(Q, IP, JC, KC, 3) - WG(Q, IC, JC, KC, 3)) * g->dxi;
                    double DUDZ = (WG(Q, IP, JC, KC, 1) - WG(Q, IP, JC, KM, 1)) * g->dzi;
                    double TAU31F = WG(MS, IP, JC, KC, 2) * (DWDX + DUDZ);
                    DWDX = (WG(Q, IC, JC, KC, 3) - WG(Q, IM, JC, KC, 3)) * g->dxi;
                    DUDZ = (WG(Q, IC, JC, KC, 1) - WG(Q, IC, JC, KM, 1)) * g->dzi;
                    double TAU31B = WG(MS, IC, JC, KC, 2) * (DWDX + DUDZ);
                    double DTAU31DX = (TAU31F - TAU31B) * COE1;
                    double DWDY = (WG(Q, IC, JP, KC, 3) - WG(Q, IC, JC, KC, 3)) * g->dyci[JJP - 1];
                    double DVDZ = (WG(Q, IC, JP, KC, 2) - WG(Q, IC, JP, KM, 2)) * g->dzi;
                    double TAU32F = WG(MS, IC, JP, KC, 3) * (DWDY + DVDZ);
                    DWDY = (WG(Q, IC, JC, KC, 3) - WG(Q, IC, JM, KC, 3)) * g->dyci[JJ - 1];
                    DVDZ = (WG(Q, IC, JC, KC, 2) - WG(Q, IC, JC, KM, 2)) * g->dzi;
                    double TAU32B = WG(MS, IC, JC, KC, 3) * (DWDY + DVDZ);
                    double DTAU32DY = (TAU32F - TAU32B) * COE2;
                    WF3(s->hz, IC, JC, KC) = WF3(s->hz, IC, JC, KC) + DTAU31DX + DTAU32DY + DTAU33DZ;
                }
            }
        }
    }
}

/* VISCOUS_PAR_EXPLT_X/Y/Z_io (src/VISCOUS_X_EXPLICIT_io.f90:116-289,
 * src/VISCOUS_Y_EXPLICIT_io.f90:264-423, src/VISCOUS_Z_EXPLICIT_io.f90:289-460,
 * channel branches): the viscous stress split for RK-CN.  The explicit parts
 * are added to the convection terms (EX + h); the implicit parts (evaluated
 * explicitly) go to im, the density corrections (with DRHOI_STG) to ad.
 * Rows 2..M for y (row 1 of im, ad zero).  DIVU_io ends as the z correction
 * on rows 1..M, as the reference leaves it. */
static void io_viscous_par(const orc_grid *g, orc_io *s)
{
    const double CV = g->cvisc;
    const size_t ncell = (size_t)WL * WM * WN;
    memset(s->im, 0, 3 * ncell * sizeof(double));
    memset(s->ad, 0, 3 * ncell * sizeof(double));
#pragma omp parallel for schedule(static)
    for (int JC = 1; JC <= WM; ++JC) {
        const int JM = JC - 1, JP = JC + 1, JJ = JC, JJP = JC + 1, JJM = JC - 1;
        for (int KC = 1; KC <= WN; ++KC) {
            const int KM = WKMV(KC), KP = WKPV(KC);
            for (int IC = 1; IC <= WL; ++IC) {
                const int IP = WIPV(IC), IM = WIMV(IC);
                const double *Q = s->q, *G = s->g, *MU = s->mu, *MS = s->mu_stg, *DR = s->drhoi_stg;
                {   /* x */
                    const double COE1 = -2.0 / 3.0 * g->dxi * CV, COE2 = g->dyfi[JJ - 1] * CV * 1.0;
                    const double COE32 = g->dzi * CV * 1.0;
                    double DUDX = (WG(Q, IP, JC, KC, 1) - WG(Q, IC, JC, KC, 1)) * g->dxi;
                    double DVDY = (WG(Q, IC, JP, KC, 2) - WG(Q, IC, JC, KC, 2)) * g->dyfi[JJ - 1] * 1.0;
                    double DWDZ = (WG(Q, IC, JC, KP, 3) - WG(Q, IC, JC, KC, 3)) * g->dzi * 1.0;
                    const double F_EX = W3(MU, IC, JC, KC) * (DVDY + DWDZ), F_IM = W3(MU, IC, JC, KC) * DUDX;
                    DUDX = (WG(G, IP, JC, KC, 1) * WG(DR, IP, JC, KC, 1) - WG(G, IC, JC, KC, 1) * WG(DR, IC, JC, KC, 1)) * g->dxi;
                    const double F_AD = W3(MU, IC, JC, KC) * DUDX;
                    DUDX = (WG(Q, IC, JC, KC, 1) - WG(Q, IM, JC, KC, 1)) * g->dxi;
                    DVDY = (WG(Q, IM, JP, KC, 2) - WG(Q, IM, JC, KC, 2)) * g->dyfi[JJ - 1] * 1.0;
                    DWDZ = (WG(Q, IM, JC, KP, 3) - WG(Q, IM, JC, KC, 3)) * g->dzi * 1.0;
                    const double B_EX = W3(MU, IM, JC, KC) * (DVDY + DWDZ), B_IM = W3(MU, IM, JC, KC) * DUDX;
                    DUDX = (WG(G, IC, JC, KC, 1) * WG(DR, IC, JC, KC, 1) - WG(G, IM, JC, KC, 1) * WG(DR, IM, JC, KC, 1)) * g->dxi;
                    const double B_AD = W3(MU, IM, JC, KC) * DUDX;
                    const double X_EX = (F_EX - B_EX) * COE1, X_IM = (F_IM - B_IM) * COE1 * (-2.0);
                    const double X_AD = (F_AD - B_AD) * COE1 * (-2.0);
                    double VP = WG(MS, IC, JP, KC, 1);
                    double DVDX = (WG(Q, IC, JP, KC, 2) - WG(Q, IM, JP, KC, 2)) * g->dxi;
                    double DUDY = (WG(Q, IC, JP, KC, 1) - WG(Q, IC, JC, KC, 1)) * g->dyci[JJP - 1] / 1.0;
                    const double F2_EX = VP * DVDX, F2_IM = VP * DUDY;
                    DUDY = (WG(G, IC, JP, KC, 1) * WG(DR, IC, JP, KC, 1) - WG(G, IC, JC, KC, 1) * WG(DR, IC, JC, KC, 1)) *
                           g->dyci[JJP - 1] / 1.0;
                    const double F2_AD = VP * DUDY;
                    double VC = WG(MS, IC, JC, KC, 1);
                    DUDY = (WG(Q, IC, JC, KC, 1) - WG(Q, IC, JM, KC, 1)) * g->dyci[JJ - 1] / 1.0;
                    DVDX = (WG(Q, IC, JC, KC, 2) - WG(Q, IM, JC, KC, 2)) * g->dxi;
                    const double B2_EX = VC * DVDX, B2_IM = VC * DUDY;
                    DUDY = (WG(G, IC, JC, KC, 1) * WG(DR, IC, JC, KC, 1) - WG(G, IC, JM, KC, 1) * WG(DR, IC, JM, KC, 1)) *
                           g->dyci[JJ - 1] / 1.0;
                    const double B2_AD = VC * DUDY;
                    const double Y_EX = (F2_EX - B2_EX) * COE2, Y_IM = (F2_IM - B2_IM) * COE2, Y_AD = (F2_AD - B2_AD) * COE2;
                    VP = WG(MS, IC, JC, KP, 2);
                    double DUDZ = (WG(Q, IC, JC, KP, 1) - WG(Q, IC, JC, KC, 1)) * g->dzi;
                    double DWDX = (WG(Q, IC, JC, KP, 3) - WG(Q, IM, JC, KP, 3)) * g->dxi;
                    const double F3_EX = VP * DWDX, F3_IM = VP * DUDZ;
                    DUDZ = (WG(Q, IC, JC, KP, 1) * WG(DR, IC, JC, KP, 1) - WG(Q, IC, JC, KC, 1) * WG(DR, IC, JC, KC, 1)) * g->dzi;
                    const double F3_AD = VP * DUDZ;
                    VC = WG(MS, IC, JC, KC, 2);
                    DUDZ = (WG(Q, IC, JC, KC, 1) - WG(Q, IC, JC, KM, 1)) * g->dzi;
                    DWDX = (WG(Q, IC, JC, KC, 3) - WG(Q, IM, JC, KC, 3)) * g->dxi;
                    const double B3_EX = VC * DWDX, B3_IM = VC * DUDZ;
                    DUDZ = (WG(Q, IC, JC, KC, 1) * WG(DR, IC, JC, KC, 1) - WG(Q, IC, JC, KM, 1) * WG(DR, IC, JC, KM, 1)) * g->dzi;
                    const double B3_AD = VC * DUDZ;
                    const double Z_EX = (F3_EX - B3_EX) * COE32, Z_IM = (F3_IM - B3_IM) * COE32, Z_AD = (F3_AD - B3_AD) * COE32;
                    WF3(s->hx, IC, JC, KC) = X_EX + Y_EX + Z_EX + WF3(s->hx, IC, JC, KC);
                    WE(s->im, IC, JC, KC, 1) = X_IM + Y_IM + Z_IM;
                    WE(s->ad, IC, JC, KC, 1) = X_AD + Y_AD + Z_AD;
                }
                if (JC >= 2) {   /* y */
                    const double COE1 = g->dxi * CV, COE3 = g->dzi * CV, COE2 = -2.0 / 3.0 * g->dyci[JJ - 1] * CV;
                    double DUDX = (WG(Q, IP, JC, KC, 1) - WG(Q, IC, JC, KC, 1)) * g->dxi;
                    double DVDY = (WG(Q, IC, JP, KC, 2) - WG(Q, IC, JC, KC, 2)) * g->dyfi[JJ - 1];
                    double DWDZ = (WG(Q, IC, JC, KP, 3) - WG(Q, IC, JC, KC, 3)) * g->dzi;
                    const double F_EX = W3(MU, IC, JC, KC) * (DUDX + DWDZ), F_IM = W3(MU, IC, JC, KC) * DVDY;
                    DVDY = (WG(G, IC, JP, KC, 2) * WG(DR, IC, JP, KC, 2) - WG(G, IC, JC, KC, 2) * WG(DR, IC, JC, KC, 2)) *
                           g->dyfi[JJ - 1];
                    const double F_AD = W3(MU, IC, JC, KC) * DVDY;
                    DUDX = (WG(Q, IP, JM, KC, 1) - WG(Q, IC, JM, KC, 1)) * g->dxi;
                    DVDY = (WG(Q, IC, JC, KC, 2) - WG(Q, IC, JM, KC, 2)) * g->dyfi[JJM - 1];
                    DWDZ = (WG(Q, IC, JM, KP, 3) - WG(Q, IC, JM, KC, 3)) * g->dzi;
                    const double B_EX = W3(MU, IC, JM, KC) * (DUDX + DWDZ), B_IM = W3(MU, IC, JM, KC) * DVDY;
                    DVDY = (WG(G, IC, JC, KC, 2) * WG(DR, IC, JC, KC, 2) - WG(G, IC, JM, KC, 2) * WG(DR, IC, JM, KC, 2)) *
                           g->dyfi[JJM - 1];
                    const double B_AD = W3(MU, IC, JM, KC) * DVDY;
                    const double Y_EX = (F_EX - B_EX) * COE2, Y_IM = (F_IM - B_IM) * COE2 * (-2.0);
                    const double Y_AD = (F_AD - B_AD) * COE2 * (-2.0);
                    double VP = WG(MS, IP, JC, KC, 1);
                    double DVDX = (WG(Q, IP, JC, KC, 2) - WG(Q, IC, JC, KC, 2)) * g->dxi;
                    double DUDY = (WG(Q, IP, JC, KC, 1) - WG(Q, IP, JM, KC, 1)) * g->dyci[JJ - 1];
                    const double F1_EX = VP * DUDY, F1_IM = VP * DVDX;
                    DVDX = (WG(G, IP, JC, KC, 2) * WG(DR, IP, JC, KC, 2) - WG(G, IC, JC, KC, 2) * WG(DR, IC, JC, KC, 2)) * g->dxi;
                    const double F1_AD = VP * DVDX;
                    double VC = WG(MS, IC, JC, KC, 1);
                    DVDX = (WG(Q, IC, JC, KC, 2) - WG(Q, IM, JC, KC, 2)) * g->dxi;
                    DUDY = (WG(Q, IC, JC, KC, 1) - WG(Q, IC, JM, KC, 1)) * g->dyci[JJ - 1];
                    const double B1_EX = VC * DUDY, B1_IM = VC * DVDX;
                    DVDX = (WG(G, IC, JC, KC, 2) * WG(DR, IC, JC, KC, 2) - WG(G, IM, JC, KC, 2) * WG(DR, IM, JC, KC, 2)) * g->dxi;
                    const double B1_AD = VC * DVDX;
                    const double X_EX = (F1_EX - B1_EX) * COE1, X_IM = (F1_IM - B1_IM) * COE1, X_AD = (F1_AD - B1_AD) * COE1;
                    VP = WG(MS, IC, JC, KP, 3);
                    double DVDZ = (WG(Q, IC, JC, KP, 2) - WG(Q, IC, JC, KC, 2)) * g->dzi;
                    double DWDY = (WG(Q, IC, JC, KP, 3) - WG(Q, IC, JM, KP, 3)) * g->dyci[JJ - 1];
                    const double F3_EX = VP * DWDY, F3_IM = VP * DVDZ;
                    DVDZ = (WG(G, IC, JC, KP, 2) * WG(DR, IC, JC, KP, 2) - WG(G, IC, JC, KC, 2) * WG(DR, IC, JC, KC, 2)) * g->dzi;
                    const double F3_AD = VP * DVDZ;
                    VC = WG(MS, IC, JC, KC, 3);
                    DVDZ = (WG(Q, IC, JC, KC, 2) - WG(Q, IC, JC, KM, 2)) * g->dzi;
                    DWDY = (WG(Q, IC, JC, KC, 3) - WG(Q, IC, JM, KC, 3)) * g->dyci[JJ - 1];
                    const double B3_EX = VC * DWDY, B3_IM = VC * DVDZ;
                    DVDZ = (WG(G, IC, JC, KC, 2) * WG(DR, IC, JC, KC, 2) - WG(G, IC, JC, KM, 2) * WG(DR, IC, JC, KM, 2)) * g->dzi;
                    const double B3_AD = VC * DVDZ;
                    const double Z_EX = (F3_EX - B3_EX) * COE3, Z_IM = (F3_IM - B3_IM) * COE3, Z_AD = (F3_AD - B3_AD) * COE3;
                    WF3(s->hy, IC, JC, KC) = X_EX + Y_EX + Z_EX + WF3(s->hy, IC, JC, KC);
                    WE(s->im, IC, JC, KC, 2) = X_IM + Y_IM + Z_IM;
                    WE(s->ad, IC, JC, KC, 2) = X_AD + Y_AD + Z_AD;
                }
                {   /* z */
                    const double COE1 = g->dxi * CV, COE3 = -2.0 / 3.0 * g->dzi * CV, COE2 = g->dyfi[JJ - 1] * CV;
                    double DUDX = (WG(Q, IP, JC, KC, 1) - WG(Q, IC, JC, KC, 1)) * g->dxi;
                    double DVDY = (WG(Q, IC, JP, KC, 2) - WG(Q, IC, JC, KC, 2)) * g->dyfi[JJ - 1];
                    double DWDZ = (WG(Q, IC, JC, KP, 3) - WG(Q, IC, JC, KC, 3)) * g->dzi;
                    const double F_EX = W3(MU, IC, JC, KC) * (DUDX + DVDY), F_IM = W3(MU, IC, JC, KC) * DWDZ;
                    DWDZ = (WG(G, IC, JC, KP, 3) * WG(DR, IC, JC, KP, 3) - WG(G, IC, JC, KC, 3) * WG(DR, IC, JC, KC, 3)) * g->dzi;
                    const double F_AD = W3(MU, IC, JC, KC) * DWDZ;
                    DUDX = (WG(Q, IP, JC, KM, 1) - WG(Q, IC, JC, KM, 1)) * g->dxi;
                    DVDY = (WG(Q, IC, JP, KM, 2) - WG(Q, IC, JC, KM, 2)) * g->dyfi[JJ - 1];
                    DWDZ = (WG(Q, IC, JC, KC, 3) - WG(Q, IC, JC, KM, 3)) * g->dzi;
                    const double B_EX = W3(MU, IC, JC, KM) * (DUDX + DVDY), B_IM = W3(MU, IC, JC, KM) * DWDZ;
                    DWDZ = (WG(G, IC, JC, KC, 3) * WG(DR, IC, JC, KC, 3) - WG(G, IC, JC, KM, 3) * WG(DR, IC, JC, KM, 3)) * g->dzi;
                    const double B_AD = W3(MU, IC, JC, KM) * DWDZ;
                    const double Z_EX = (F_EX - B_EX) * COE3, Z_IM = (F_IM - B_IM) * COE3 * (-2.0);
                    const double Z_AD = (F_AD - B_AD) * COE3 * (-2.0);
                    double VP = WG(MS, IP, JC, KC, 2);
                    double DWDX = (WG(Q, IP, JC, KC, 3) - WG(Q, IC, JC, KC, 3)) * g->dxi;
                    double DUDZ = (WG(Q, IP, JC, KC, 1) - WG(Q, IP, JC, KM, 1)) * g->dzi;
                    const double F1_EX = VP * DUDZ, F1_IM = VP * DWDX;
                    DWDX = (WG(G, IP, JC, KC, 3) * WG(DR, IP, JC, KC, 3) - WG(G, IC, JC, KC, 3) * WG(DR, IC, JC, KC, 3)) * g->dxi;
                    const double F1_AD = VP * DWDX;
                    double VC = WG(MS, IC, JC, KC, 2);
                    DWDX = (WG(Q, IC, JC, KC, 3) - WG(Q, IM, JC, KC, 3)) * g->dxi;
                    DUDZ = (WG(Q, IC, JC, KC, 1) - WG(Q, IC, JC, KM, 1)) * g->dzi;
                    const double B1_EX = VC * DUDZ, B1_IM = VC * DWDX;
                    DWDX = (WG(G, IC, JC, KC, 3) * WG(DR, IC, JC, KC, 3) - WG(G, IM, JC, KC, 3) * WG(DR, IM, JC, KC, 3)) * g->dxi;
                    const double B1_AD = VC * DWDX;
                    const double X_EX = (F1_EX - B1_EX) * COE1, X_IM = (F1_IM - B1_IM) * COE1, X_AD = (F1_AD - B1_AD) * COE1;
                    VP = WG(MS, IC, JP, KC, 3);
                    double DWDY = (WG(Q, IC, JP, KC, 3) - WG(Q, IC, JC, KC, 3)) * g->dyci[JJP - 1];
                    double DVDZ = (WG(Q, IC, JP, KC, 2) - WG(Q, IC, JP, KM, 2)) * g->dzi;
                    const double F2_EX = VP * DVDZ, F2_IM = VP * DWDY;
                    DWDY = (WG(G, IC, JP, KC, 3) * WG(DR, IC, JP, KC, 3) - WG(G, IC, JC, KC, 3) * WG(DR, IC, JC, KC, 3)) *
                           g->dyci[JJP - 1];
                    const double F2_AD = VP * DWDY;
                    VC = WG(MS, IC, JC, KC, 3);
                    DWDY = (WG(Q, IC, JC, KC, 3) - WG(Q, IC, JM, KC, 3)) * g->dyci[JJ - 1];
                    DVDZ = (WG(Q, IC, JC, KC, 2) - WG(Q, IC, JC, KM, 2)) * g->dzi;
                    const double B2_EX = VC * DVDZ, B2_IM = VC * DWDY;
                    DWDY = (WG(G, IC, JC, KC, 3) * WG(DR, IC, JC, KC, 3) - WG(G, IC, JM, KC, 3) * WG(DR, IC, JM, KC, 3)) *
                           g->dyci[JJ - 1];
                    const double B2_AD = VC * DWDY;
                    const double Y_EX = (F2_EX - B2_EX) * COE2, Y_IM = (F2_IM - B2_IM) * COE2, Y_AD = (F2_AD - B2_AD) * COE2;
                    WF3(s->hz, IC, JC, KC) = X_EX + Y_EX + Z_EX + WF3(s->hz, IC, JC, KC);
                    WE(s->im, IC, JC, KC, 3) = X_IM + Y_IM + Z_IM;
                    WE(s->ad, IC, JC, KC, 3) = X_AD + Y_AD + Z_AD;
                    W3(s->divu, IC, JC, KC) = WE(s->ad, IC, JC, KC, 3);
                }
            }
        }
    }
}

#define WAMV(JC, IDR) g->amvr[((IDR)-1) * (WM + 1) + (JC)-1]
#define WAPV(JC, IDR) g->apvr[((IDR)-1) * (WM + 1) + (JC)-1]
/* MOMFA_io(NS, IDR) (src/MOMFA_io.f90:17-48): MOMFA1_X_io, MOMFA2_Z_io,
 * MOMFA3_Y_io (src/MOMFA_io.f90:51-377) on r, then G += r from row N2I. */
static void io_momfa(const orc_grid *g, orc_io *s, int NS, int IDR, double DT, double *r)
{
    const int N2I = IDR == 2 ? 2 : 1, JSZ = WM - N2I + 1;
    const double FACOE = 0.5 * talp(NS) * DT * g->cvisc;
    const double *MU = s->mu, *MS = s->mu_stg, *RS = s->rho_stg;
    /* x: cyclic lines along i, one set per k (TDMAIJI_CYC) */
    {
        const double COE1 = FACOE * 4.0 / 3.0 * g->dxqi, COE2 = FACOE * g->dxqi;
        const size_t pl = (size_t)WL * JSZ;
        double *A = (double *)malloc(4 * pl * sizeof(double)), *B = A + pl, *Cc = B + pl, *F = Cc + pl;
        for (int KC = 1; KC <= WN; ++KC) {
            for (int JC = N2I; JC <= WM; ++JC)
                for (int IC = 1; IC <= WL; ++IC) {
                    const int IM = WIMV(IC), IP = WIPV(IC);
                    double AMU, APU;
                    if (IDR == 1) {
                        AMU = -COE1 * W3(MU, IM, JC, KC);
                        APU = -COE1 * W3(MU, IC, JC, KC);
                    } else if (IDR == 2) {
                        AMU = -COE2 * WG(MS, IC, JC, KC, 1);
                        APU = -COE2 * WG(MS, IP, JC, KC, 1);
                    } else {
                        AMU = -COE2 * WG(MS, IC, JC, KC, 2);
                        APU = -COE2 * WG(MS, IP, JC, KC, 2);
                    }
                    const size_t o = (size_t)(IC - 1) + (size_t)WL * (JC - N2I);
                    A[o] = AMU / WG(RS, IM, JC, KC, IDR);
                    Cc[o] = APU / WG(RS, IP, JC, KC, IDR);
                    B[o] = 1.0 - (AMU + APU) / WG(RS, IC, JC, KC, IDR);
                    F[o] = WF3(r, IC, JC, KC);
                }
            orc_tdma_cyc(A, B, Cc, F, WL, JSZ, 1, WL);
            for (int JC = N2I; JC <= WM; ++JC)
                for (int IC = 1; IC <= WL; ++IC) WF3(r, IC, JC, KC) = F[(size_t)(IC - 1) + (size_t)WL * (JC - N2I)];
        }
        free(A);
    }
    /* z: cyclic lines along k, one set per i */
    {
        const double COE1 = FACOE * 4.0 / 3.0 * g->dzqi, COE2 = FACOE * g->dzqi;
        const size_t pl = (size_t)WN * JSZ;
        double *A = (double *)malloc(4 * pl * sizeof(double)), *B = A + pl, *Cc = B + pl, *F = Cc + pl;
        for (int IC = 1; IC <= WL; ++IC) {
            for (int KC = 1; KC <= WN; ++KC) {
                const int KM = WKMV(KC), KP = WKPV(KC);
                for (int JC = N2I; JC <= WM; ++JC) {
                    double AMU, APU;
                    if (IDR == 1) {
                        AMU = -COE2 * WG(MS, IC, JC, KC, 2) * 1.0;
                        APU = -COE2 * WG(MS, IC, JC, KP, 2) * 1.0;
                    } else if (IDR == 2) {
                        AMU = -COE2 * WG(MS, IC, JC, KC, 3) * 1.0;
                        APU = -COE2 * WG(MS, IC, JC, KP, 3) * 1.0;
                    } else {
                        AMU = -COE1 * W3(MU, IC, JC, KM) * 1.0;
                        APU = -COE1 * W3(MU, IC, JC, KC) * 1.0;
                    }
                    const size_t o = (size_t)(KC - 1) + (size_t)WN * (JC - N2I);
                    A[o] = AMU / WG(RS, IC, JC, KM, IDR);
                    Cc[o] = APU / WG(RS, IC, JC, KP, IDR);
                    B[o] = 1.0 - (AMU + APU) / WG(RS, IC, JC, KC, IDR);
                    F[o] = WF3(r, IC, JC, KC);
                }
            }
            orc_tdma_cyc(A, B, Cc, F, WN, JSZ, 1, WN);
            for (int KC = 1; KC <= WN; ++KC)
                for (int JC = N2I; JC <= WM; ++JC) WF3(r, IC, JC, KC) = F[(size_t)(KC - 1) + (size_t)WN * (JC - N2I)];
        }
        free(A);
    }
    /* y: non-cyclic lines along j (rows 1..M), one set per k (TDMAIJJ_nonCYC,
     * zero boundary values); the wall rows of u, w use AMVR1 / APVRN, the
     * coefficients TDMA_COEF sets aside (src/TDMA_COEF.f90:86-96) */
    {
        const double COE1 = FACOE * 4.0 / 3.0, COE2 = FACOE;
        const double AMVR1 = 4.0 / (2.0 / g->dyci[1] + 1.0 / g->dyfi[0]) * (2.0 * g->dyfi[0]);
        const double APVRN = 4.0 / (2.0 / g->dyci[WM - 1] + 1.0 / g->dyfi[WM - 1]) * (2.0 * g->dyfi[WM - 1]);
        const size_t pl = (size_t)WL * WM;
        double *A = (double *)malloc(4 * pl * sizeof(double)), *B = A + pl, *Cc = B + pl, *F = Cc + pl;
        for (int KC = 1; KC <= WN; ++KC) {
            for (int JC = 1; JC <= WM; ++JC) {
                const int JM = JC - 1, JP = JC + 1;
                for (int IC = 1; IC <= WL; ++IC) {
                    const size_t o = (size_t)(IC - 1) + (size_t)WL * (JC - 1);
                    double AMU, APU, a, c, b;
                    if (IDR == 1 || IDR == 3) {
                        const double *FM = MS + (size_t)(IDR - 1) * WL * (WM + 2) * WN;
                        const double *FD = RS + (size_t)(IDR - 1) * WL * (WM + 2) * WN;
                        if (JC == 1) {
                            const double MUQT = (W3(MU, IC, JC, KC) + W3(MU, IC, JM, KC)) * 0.5;
                            AMU = -COE2 * AMVR1 * MUQT;
                            APU = -COE2 * WAPV(JC, IDR) * W3(FM, IC, JP, KC);
                            a = 0.0;
                            c = APU / W3(FD, IC, JP, KC);
                        } else if (JC == WM) {
                            const double MUQT = (W3(MU, IC, WM, KC) + W3(MU, IC, WM + 1, KC)) * 0.5;
                            AMU = -COE2 * WAMV(JC, IDR) * W3(FM, IC, JC, KC);
                            APU = -COE2 * APVRN * MUQT;
                            a = AMU / W3(FD, IC, JM, KC);
                            c = 0.0;
                        } else {
                            AMU = -COE2 * WAMV(JC, IDR) * W3(FM, IC, JC, KC);
                            APU = -COE2 * WAPV(JC, IDR) * W3(FM, IC, JP, KC);
                            a = AMU / W3(FD, IC, JM, KC);
                            c = APU / W3(FD, IC, JP, KC);
                        }
                        b = 1.0 - (AMU + APU) / W3(FD, IC, JC, KC);
                    } else {
                        if (JC == 1) {
                            AMU = -COE1 * WAMV(JC, 2) * W3(MU, IC, JM, KC);
                            APU = -COE1 * WAPV(JC, 2) * W3(MU, IC, JC, KC);
                            a = AMU / WG(RS, IC, JM, KC, 2);
                            c = APU / WG(RS, IC, JP, KC, 2);
                            b = 1.0 - (AMU + APU) / WG(RS, IC, JC, KC, 2);
                        } else {
                            AMU = -COE1 * WAMV(JC, 2) * W3(MU, IC, JM, KC);
                            APU = -COE1 * WAPV(JC, 2) * W3(MU, IC, JC, KC);
                            a = AMU / WG(RS, IC, JM, KC, 2);
                            c = APU / WG(RS, IC, JC, KC, 2);
                            b = 1.0 - (AMU + APU) / WG(RS, IC, JC, KC, 2);
                        }
                    }
                    A[o] = a;
                    B[o] = b;
                    Cc[o] = c;
                    F[o] = WF3(r, IC, JC, KC);
                }
            }
            orc_tdma_noncyc(A, B, Cc, F, NULL, WM, WL, WL, 1);
            for (int JC = 1; JC <= WM; ++JC)
                for (int IC = 1; IC <= WL; ++IC) WF3(r, IC, JC, KC) = F[(size_t)(IC - 1) + (size_t)WL * (JC - 1)];
        }
        free(A);
    }
    for (int KC = 1; KC <= WN; ++KC)
        for (int JC = N2I; JC <= WM; ++JC)
            for (int IC = 1; IC <= WL; ++IC) WG(s->g, IC, JC, KC, IDR) = WF3(r, IC, JC, KC) + WG(s->g, IC, JC, KC, IDR);
}

/* MOMFA_io(NS, IDR) alone (src/MOMFA_io.f90:17-48): the factorised x, z, y
 * solves of the right-hand side r (L, M, N), overwritten, then G(IDR) += r on
 * rows N2I..M.  Exported so that the tests can check the line solves against
 * dense operators. */
void orc_io_momfa(const orc_grid *g, orc_io *s, int NS, int IDR, double DT, double *r)
{
    io_momfa(g, s, NS, IDR, DT, r);
}

void orc_io_substep(const orc_grid *g, orc_io *s, const orc_io_params *p, int NS, double DT)
{
    const size_t ngh = (size_t)WL * (WM + 2) * WN;
    memcpy(s->g0, s->g, 3 * ngh * sizeof(double));
    if (p->visimplicit) {
        /* VISCOUS_PAR_EXPLT_*_io, RHS_MOM_EXPLICIT_io and MOMFA_io per IDR
         * (src/SOLVERRK3_MOM_IO.f90:76-90); MOMFA_io adds to G itself */
        io_convection(g, s);
        io_viscous_par(g, s);
        double *h[3] = {s->hx, s->hy, s->hz};
        for (int IDR = 1; IDR <= 3; ++IDR) {
            io_rhs_mom(g, s, p, NS, IDR, DT, h[IDR - 1]);
            io_momfa(g, s, NS, IDR, DT, h[IDR - 1]);
        }
    } else {
        orc_io_terms(g, s);
        io_rhs_mom(g, s, p, NS, 1, DT, s->hx);
        io_rhs_mom(g, s, p, NS, 2, DT, s->hy);
        io_rhs_mom(g, s, p, NS, 3, DT, s->hz);
        /* MASSFLUX_CALC_IO */
        for (int KC = 1; KC <= WN; ++KC)
            for (int JC = 1; JC <= WM; ++JC)
                for (int IC = 1; IC <= WL; ++IC) {
                    WG(s->g, IC, JC, KC, 1) = WG(s->g, IC, JC, KC, 1) + WF3(s->hx, IC, JC, KC);
                    WG(s->g, IC, JC, KC, 3) = WG(s->g, IC, JC, KC, 3) + WF3(s->hz, IC, JC, KC);
                    if (JC >= 2) WG(s->g, IC, JC, KC, 2) = WG(s->g, IC, JC, KC, 2) + WF3(s->hy, IC, JC, KC);
                }
    }
    /* INTFC_VARS3 + BC_WALL_G_io */
    io_wrap3(g, s->g);
    io_bc_wall_vec(g, s->g);
    /* DIVG_io(NS) */
    const double COE0 = 1.0 / DT / talp(NS);
#pragma omp parallel for schedule(static)
    for (int JC = 1; JC <= WM; ++JC) {
        const int JP = JC + 1;
        const double COE1 = g->dxi / 1.0, COE2 = g->dyfi[JC - 1] / 1.0;
        for (int KC = 1; KC <= WN; ++KC) {
            const int KP = WKPV(KC);
            for (int IC = 1; IC <= WL; ++IC) {
                const int IP = WIPV(IC);
                double DIVX = (WG(s->g, IP, JC, KC, 1) - WG(s->g, IC, JC, KC, 1)) * COE1;
                double DIVY = (WG(s->g, IC, JP, KC, 2) - WG(s->g, IC, JC, KC, 2)) * COE2;
                double DIVZ = (WG(s->g, IC, JC, KP, 3) - WG(s->g, IC, JC, KC, 3)) * g->dzi;
                double DTRHO = W3(s->drhodtp, IC, JC, KC) / 1.0;
                double DQCAP = DIVX + DIVY + DIVZ + DTRHO;
                WF3(s->rhsllphi, IC, JC, KC) = DQCAP * COE0;
            }
        }
    }
    orc_poisson_fft99(g, s->rhsllphi, s->dph, 0);
    orc_bc_wall_neumann(g, s->dph);
    /* PRCALC_io; INTFC_VARS1 + BC_WALL_PR_io */
    const double BE = 0.5 * talp(NS) * DT * g->cvisc;
    for (int KC = 1; KC <= WN; ++KC)
        for (int JC = 1; JC <= WM; ++JC)
            for (int IC = 1; IC <= WL; ++IC) {
                if (p->visimplicit)
                    W3(s->pr, IC, JC, KC) =
                        W3(s->pr, IC, JC, KC) + W3(s->dph, IC, JC, KC) - BE * WF3(s->rhsllphi, IC, JC, KC);
                else
                    W3(s->pr, IC, JC, KC) = W3(s->pr, IC, JC, KC) + W3(s->dph, IC, JC, KC);
            }
    io_wrap(g, s->pr);
    orc_bc_wall_neumann(g, s->pr);
    /* MASSFLUX_UPDATE_IO */
    {
        const double CX = DT * talp(NS) * g->dxi, CZ = DT * talp(NS) * g->dzi;
#pragma omp parallel for schedule(static)
        for (int JC = 1; JC <= WM; ++JC)
            for (int KC = 1; KC <= WN; ++KC) {
                const int KM = WKMV(KC);
                const double CY = DT * talp(NS) * g->dyci[JC - 1] / 1.0;
                for (int IC = 1; IC <= WL; ++IC) {
                    const int IM = WIMV(IC);
                    double DFX11 = W3(s->dph, IC, JC, KC) - W3(s->dph, IM, JC, KC);
                    WG(s->g, IC, JC, KC, 1) = WG(s->g, IC, JC, KC, 1) - DFX11 * CX;
                    double DFX33 = W3(s->dph, IC, JC, KC) - W3(s->dph, IC, JC, KM);
                    WG(s->g, IC, JC, KC, 3) = WG(s->g, IC, JC, KC, 3) - DFX33 * CZ;
                    if (JC >= 2) {
                        double DFX22 = W3(s->dph, IC, JC, KC) - W3(s->dph, IC, JC - 1, KC);
                        WG(s->g, IC, JC, KC, 2) = WG(s->g, IC, JC, KC, 2) - DFX22 * CY;
                    }
                }
            }
    }
    io_wrap3(g, s->g);
    io_bc_wall_vec(g, s->g);
}

/* DIVGCK_io (src/DIVGCK.f90:61-124, periodic x): max |div G + DrhoDtP| and
 * max |div Q| over the interior. */
void orc_divgck_io(const orc_grid *g, const orc_io *s, double *out)
{
    double DG = 0.0, DQ = 0.0;
#pragma omp parallel for reduction(max : DG, DQ) schedule(static)
    for (int JC = 1; JC <= WM; ++JC) {
        const int JP = JC + 1;
        const double COE2 = g->dyfi[JC - 1] * 1.0, COE3 = g->dzi * 1.0;
        for (int KC = 1; KC <= WN; ++KC) {
            const int KP = WKPV(KC);
            for (int IC = 1; IC <= WL; ++IC) {
                const int IP = WIPV(IC);
                double d = (WG(s->g, IP, JC, KC, 1) - WG(s->g, IC, JC, KC, 1)) * g->dxi +
                           (WG(s->g, IC, JP, KC, 2) - WG(s->g, IC, JC, KC, 2)) * COE2 +
                           (WG(s->g, IC, JC, KP, 3) - WG(s->g, IC, JC, KC, 3)) * COE3 + W3(s->drhodtp, IC, JC, KC);
                if (fabs(d) > DG) DG = fabs(d);
                d = (WG(s->q, IP, JC, KC, 1) - WG(s->q, IC, JC, KC, 1)) * g->dxi +
                    (WG(s->q, IC, JP, KC, 2) - WG(s->q, IC, JC, KC, 2)) * COE2 +
                    (WG(s->q, IC, JC, KP, 3) - WG(s->q, IC, JC, KC, 3)) * COE3;
                if (fabs(d) > DQ) DQ = fabs(d);
            }
        }
    }
    out[0] = DG;
    out[1] = DQ;
}
#undef WAMV
#undef WAPV
#undef WL
#undef WM
#undef WN
#undef W3
#undef WF3
#undef WG
#undef WE
#undef WIPV
#undef WIMV
#undef WKPV
#undef WKMV

/* ------------------------------------------------------------------------
 * PP_MEAN_ZX_FLOW_Xperiodic_IO (src/processing/PP_MEAN_ZX.f90:299-800), the
 * moments of the channel (ICASE = 1, RCCI1 = 1) that the heated channel's
 * acceptance profiles use, per row J = 1..M of the io arrays (L, 0:M+1, N[,3]):
 * cell-centred U_cct, G_cct (:379-413), then the sums of :419-437 and
 * :655-667, :735-739, each times VL1313_io = 1/(L N) (:772-795).
 * out[37 (J-1) + m]:
 *   0..3   U1xzL_io(J,1:4)    <u>, <v>, <w>, <p>
 *   4..6   G1xzL_io(J,1:3)    <g_m>
 *   7..9   UPxzL_io(J,1:3)    <u_m p>
 *   10..15 U2xzL_IO(J,1:6)    <u_m u_n>, (m,n) = 11 12 13 22 23 33 (L = (M(7-M))/2+N-3)
 *   16..21 UGxzL_IO(J,1:6)    <u_m g_n>, same order
 *   22..30 T1 D1 H1 M1 T2 D2 H2 DH PH   (T, rho, h, mu, T^2, rho^2, h^2, h rho, h p)
 *   31..33 UHxzL_io(J,1:3)    <h u_m>
 *   34..36 GHxzL_io(J,1:3)    <h g_m>
 * thermal = 0 leaves 22..36 at zero (the reference skips them).
 * ------------------------------------------------------------------------ */
#define ML (g->L)
#define MM (g->M)
#define MN (g->N)
#define M3(arr, i, j, k) (arr)[(size_t)((i)-1) + (size_t)ML * ((size_t)(j) + (size_t)(MM + 2) * ((k)-1))]
#define MG(arr, i, j, k, c) (arr)[(size_t)((i)-1) + (size_t)ML * ((size_t)(j) + (size_t)(MM + 2) * ((size_t)((k)-1) + (size_t)MN * ((c)-1)))]
void orc_mean_xz_io(const orc_grid *g, const double *q, const double *gf, const double *pr, const double *T,
                    const double *rho, const double *h, const double *mu, int thermal, double *out)
{
    const double VL = 1.0 / (double)(ML * MN);
#pragma omp parallel for schedule(static)
    for (int J = 1; J <= MM; ++J) {
        double acc[37];
        for (int m = 0; m < 37; ++m) acc[m] = 0.0;
        const int JP = J + 1;
        for (int I = 1; I <= ML; ++I) {
            const int IP = I == ML ? 1 : I + 1;
            for (int K = 1; K <= MN; ++K) {
                const int KP = K == MN ? 1 : K + 1;
                double U[3], G[3];
                G[0] = (MG(gf, I, J, K, 1) + MG(gf, IP, J, K, 1)) * 0.5;
                U[0] = (MG(q, I, J, K, 1) + MG(q, IP, J, K, 1)) * 0.5;
                G[2] = ((MG(gf, I, J, K, 3) + MG(gf, I, J, KP, 3)) * 1.0) * 0.5;
                U[2] = ((MG(q, I, J, K, 3) + MG(q, I, J, KP, 3)) * 1.0) * 0.5;
                G[1] = ((MG(gf, I, J, K, 2) + MG(gf, I, JP, K, 2)) * 1.0) * 0.5;
                U[1] = ((MG(q, I, J, K, 2) + MG(q, I, JP, K, 2)) * 1.0) * 0.5;
                const double P = M3(pr, I, J, K);
                for (int m = 0; m < 3; ++m) {
                    acc[m] += U[m];
                    acc[4 + m] += G[m];
                    acc[7 + m] += U[m] * P;
                }
                acc[3] += P;
                for (int m = 0; m < 3; ++m)
                    for (int n = m; n < 3; ++n) {
                        const int l = ((m + 1) * (7 - (m + 1))) / 2 + (n + 1) - 3 - 1;
                        acc[16 + l] += U[m] * G[n];
                        acc[10 + l] += U[m] * U[n];
                    }
                if (thermal) {
                    const double t = M3(T, I, J, K), d = M3(rho, I, J, K), e = M3(h, I, J, K), v = M3(mu, I, J, K);
                    acc[22] += t;
                    acc[23] += d;
                    acc[24] += e;
                    acc[25] += v;
                    acc[26] += t * t;
                    acc[27] += d * d;
                    acc[28] += e * e;
                    acc[29] += e * d;
                    acc[30] += e * P;
                    for (int m = 0; m < 3; ++m) {
                        acc[31 + m] += e * U[m];
                        acc[34 + m] += e * G[m];
                    }
                }
            }
        }
        for (int m = 0; m < 37; ++m) out[37 * (J - 1) + m] = acc[m] * VL;
    }
}
#undef ML
#undef MM
#undef MN
#undef M3
#undef MG
