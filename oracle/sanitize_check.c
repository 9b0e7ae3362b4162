/* The oracle's routines under AddressSanitizer and UndefinedBehaviorSanitizer
 * (make -C oracle sanitize; run by tests/test_sanitize.py): grid set-up, the
 * four line solvers, the Poisson solve, NS = 0..3 tg substeps, the xz moments,
 * the property table lookup and the io moments on small grids with odd and
 * even sizes, so that every array is walked to its ends.  Test
 * infrastructure, like the oracle itself. */
#include "chapsim_oracle.c"

#include <stdio.h>

static double urand(unsigned long long *s)
{
    *s = *s * 6364136223846793005ULL + 1442695040888963407ULL;
    return (double)(*s >> 11) * (1.0 / 9007199254740992.0);
}

static double *vec(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

static void run_grid(int L, int M, int N, unsigned long long *seed)
{
    orc_grid g;
    memset(&g, 0, sizeof g);
    g.L = L; g.M = M; g.N = N; g.icase = 1;
    g.ynd = vec(M + 1); g.ycc = vec(M); g.dyfi = vec(M); g.dyci = vec(M + 1); g.wff = vec(M + 1); g.wfb = vec(M + 1);
    g.amph = vec(M); g.acph = vec(M); g.apph = vec(M);
    g.amvr = vec(3 * (M + 1)); g.acvr = vec(3 * (M + 1)); g.apvr = vec(3 * (M + 1));
    g.ak1 = vec(L); g.ak3 = vec(N);
    orc_grid_setup(&g, 6.283185307179586, 3.141592653589793, 2.0, 2800.0);
    const size_t ngh = (size_t)L * (M + 2) * N, nc = (size_t)L * M * N;
    orc_state s;
    s.q = vec(3 * ngh); s.pr = vec(ngh); s.convh0 = vec(3 * nc);
    s.dph = vec(ngh); s.rhs = vec(3 * nc); s.rhsllphi = vec(nc); s.qtmp = vec(nc);
    for (size_t q = 0; q < 3 * ngh; ++q) s.q[q] = 0.1 * (urand(seed) - 0.5);
    for (int c = 0; c < 3; ++c)
        for (int k = 0; k < N; ++k)
            for (int i = 0; i < L; ++i) {
                s.q[i + (size_t)L * (0 + (size_t)(M + 2) * (k + (size_t)N * c))] = 0.0;
                s.q[i + (size_t)L * ((M + 1) + (size_t)(M + 2) * (k + (size_t)N * c))] = 0.0;
            }
    double *f = vec(nc), *out = vec(13 * (size_t)M);
    for (size_t q = 0; q < nc; ++q) f[q] = urand(seed) - 0.5;
    orc_poisson_fft99(&g, f, s.dph, 0);
    (void)orc_poisson_residual(&g, s.dph, f, s.rhsllphi);
    for (int ns = 0; ns <= 3; ++ns) orc_substep_tg(&g, &s, ns, 1e-3, 1, 1, 0.0);
    for (int ns = 1; ns <= 3; ++ns) orc_substep_tg(&g, &s, ns, 1e-3, 2, 3, 0.01);
    (void)orc_divgck_tg(&g, &s);
    (void)orc_cfl_tg(&g, &s);
    orc_mean_xz(&g, &s, out);
    double *io = vec(37 * (size_t)M), *gf = vec(3 * ngh), *t = vec(ngh), *rho = vec(ngh), *h = vec(ngh), *mu = vec(ngh);
    for (size_t q = 0; q < ngh; ++q) { t[q] = 1.0 + 0.1 * urand(seed); rho[q] = 1.0; h[q] = urand(seed); mu[q] = 1.0; }
    orc_mean_xz_io(&g, s.q, gf, s.pr, t, rho, h, mu, 1, io);
    printf("grid %dx%dx%d: max|div| %.3e\n", L, M, N, orc_divgck_tg(&g, &s));
    free(io); free(gf); free(t); free(rho); free(h); free(mu);
    free(f); free(out);
    free(s.q); free(s.pr); free(s.convh0); free(s.dph); free(s.rhs); free(s.rhsllphi); free(s.qtmp);
    free(g.ynd); free(g.ycc); free(g.dyfi); free(g.dyci); free(g.wff); free(g.wfb);
    free(g.amph); free(g.acph); free(g.apph); free(g.amvr); free(g.acvr); free(g.apvr); free(g.ak1); free(g.ak3);
}

static void run_tdma(int n, int nl, unsigned long long *seed)
{
    const size_t sz = (size_t)n * nl;
    double *a = vec(sz), *b = vec(sz), *c = vec(sz), *r = vec(sz), *bc = vec(2 * (size_t)nl);
    for (size_t q = 0; q < sz; ++q) {
        a[q] = -0.1 - 0.8 * urand(seed); c[q] = -0.1 - 0.8 * urand(seed);
        b[q] = 2.0 + urand(seed); r[q] = urand(seed) - 0.5;
    }
    for (size_t q = 0; q < 2 * (size_t)nl; ++q) bc[q] = urand(seed);
    orc_tdma_noncyc(a, b, c, r, bc, n, nl, nl, 1);   /* IJJ */
    orc_tdma_noncyc(a, b, c, r, bc, n, nl, 1, n);    /* IJI */
    orc_tdma_cyc(a, b, c, r, n, nl, nl, 1);
    orc_tdma_cyc(a, b, c, r, n, nl, 1, n);
    free(a); free(b); free(c); free(r); free(bc);
}

static void run_props(unsigned long long *seed)
{
    enum { NK = 64 };
    double H[NK], Tt[NK];
    for (int i = 0; i < NK; ++i) {
        const double T = 0.8 + 0.8 * i / (NK - 1);
        Tt[i] = T;
        H[i] = T - 1.0 + 0.01 * T * T;
    }
    double *b = vec(NK), *c = vec(NK), *d = vec(NK);
    orc_cubic_spline(NK, H, Tt, b, c, d);
    for (int q = 0; q < 100; ++q) {
        const double u = H[0] + (H[NK - 1] - H[0]) * urand(seed);
        double y = 0.0;
        (void)orc_sleval(NK, H, Tt, b, c, d, u, 0, &y);
    }
    free(b); free(c); free(d);
}

int main(void)
{
    unsigned long long seed = 1973;
    run_tdma(3, 5, &seed);
    run_tdma(7, 4, &seed);
    run_tdma(12, 3, &seed);
    run_grid(8, 6, 8, &seed);
    run_grid(12, 9, 10, &seed);
    run_grid(16, 12, 20, &seed);
    run_props(&seed);
    printf("sanitize_check: done\n");
    return 0;
}
