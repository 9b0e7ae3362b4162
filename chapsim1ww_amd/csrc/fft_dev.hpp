// LDS-staged mixed-radix (2, 3, 4, 5) Stockham FFT for gfx950.
//
// Replaces the Temperton/NCAR FFT99/CFFT99 passes (libs/fft99/VPASSM.f90,
// CFFT99.f90) used by the Poisson solver.  Same transform definitions:
// forward exp(-2 pi i jk/n), inverse exp(+2 pi i jk/n), both unnormalised
// here (the 1/(L N) factor of FFT99's forward real transform and of the
// reference's UN3M is applied once, in the wall-normal solve).
//
// A block transforms `howmany` vectors held in LDS; each Stockham stage reads
// one buffer and writes the other (autosort, no bit reversal), so any length
// whose factors are 2, 3, 4 and 5 works, as FFT99 allows.
#pragma once

#include "chs_common.hpp"

namespace chs {

struct FftPlan {
    int n;
    int nst;
    int rad[20];
};

// radix-R DFT of v[0..R-1] in registers; dir = -1 forward, +1 inverse
template <int R, typename C, typename Rl>
__device__ __forceinline__ void dft_small(C (&v)[5], int dir) {
    if constexpr (R == 2) {
        C t = v[0];
        v[0] = cadd(t, v[1]);
        v[1] = csub(t, v[1]);
    } else if constexpr (R == 4) {
        C t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        C t2 = cadd(v[1], v[3]), t3 = cmul_i(csub(v[1], v[3]), dir);
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    } else if constexpr (R == 3) {
        const Rl c = Rl(-0.5), s = Rl(0.86602540378443864676372317075294);  // cos, sin(2 pi/3)
        C sum = cadd(v[1], v[2]);
        C dif = csub(v[1], v[2]);
        C m = C{v[0].x + c * sum.x, v[0].y + c * sum.y};
        C t = cmul_i(C{s * dif.x, s * dif.y}, dir);
        v[0] = cadd(v[0], sum);
        v[1] = cadd(m, t);
        v[2] = csub(m, t);
    } else if constexpr (R == 5) {
        const Rl c1 = Rl(0.30901699437494742410229341718282);   // cos(2 pi/5)
        const Rl c2 = Rl(-0.80901699437494742410229341718282);  // cos(4 pi/5)
        const Rl s1 = Rl(0.95105651629515357211643933337938);   // sin(2 pi/5)
        const Rl s2 = Rl(0.58778525229247312916870595463907);   // sin(4 pi/5)
        C a14 = cadd(v[1], v[4]), d14 = csub(v[1], v[4]);
        C a23 = cadd(v[2], v[3]), d23 = csub(v[2], v[3]);
        C m1 = C{v[0].x + c1 * a14.x + c2 * a23.x, v[0].y + c1 * a14.y + c2 * a23.y};
        C m2 = C{v[0].x + c2 * a14.x + c1 * a23.x, v[0].y + c2 * a14.y + c1 * a23.y};
        C t1 = cmul_i(C{s1 * d14.x + s2 * d23.x, s1 * d14.y + s2 * d23.y}, dir);
        C t2 = cmul_i(C{s2 * d14.x - s1 * d23.x, s2 * d14.y - s1 * d23.y}, dir);
        v[0] = cadd(v[0], cadd(a14, a23));
        v[1] = cadd(m1, t1);
        v[4] = csub(m1, t1);
        v[2] = cadd(m2, t2);
        v[3] = csub(m2, t2);
    }
}

// One Stockham stage of radix R over `howmany` vectors: x -> y.
template <int R, typename Rl>
__device__ __forceinline__ void stockham_stage(const cpx_t<Rl>* x, cpx_t<Rl>* y, int n, int howmany, int bstride,
                                               int Ns, const cpx_t<Rl>* tw, int dir) {
    using C = cpx_t<Rl>;
    const int nb = n / R;
    const int tstep = n / (Ns * R);
    for (int w = threadIdx.x; w < howmany * nb; w += blockDim.x) {
        const int bidx = w / nb, j = w - bidx * nb;
        const C* src = x + bidx * bstride;
        C* dst = y + bidx * bstride;
        const int k = j % Ns;
        C v[5];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = src[j + r * nb];
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                C t = tw[r * k * tstep];
                if (dir > 0) t.y = -t.y;
                v[r] = cmul(v[r], t);
            }
        }
        dft_small<R, C, Rl>(v, dir);
        const int idxD = (j / Ns) * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) dst[idxD + r * Ns] = v[r];
    }
}

// Full transform; returns the buffer holding the result (x or y).
template <typename Rl>
__device__ cpx_t<Rl>* fft_lds(cpx_t<Rl>* x, cpx_t<Rl>* y, int howmany, int bstride, const FftPlan& p,
                              const cpx_t<Rl>* tw, int dir) {
    int Ns = 1;
    for (int s = 0; s < p.nst; ++s) {
        const int R = p.rad[s];
        switch (R) {
            case 4: stockham_stage<4, Rl>(x, y, p.n, howmany, bstride, Ns, tw, dir); break;
            case 2: stockham_stage<2, Rl>(x, y, p.n, howmany, bstride, Ns, tw, dir); break;
            case 3: stockham_stage<3, Rl>(x, y, p.n, howmany, bstride, Ns, tw, dir); break;
            default: stockham_stage<5, Rl>(x, y, p.n, howmany, bstride, Ns, tw, dir); break;
        }
        __syncthreads();
        cpx_t<Rl>* t = x;
        x = y;
        y = t;
        Ns *= R;
    }
    return x;
}

// Host: factor n into radices 4, 2, 3, 5 (FFTFAX-style factor list).
inline bool fft_plan(int n, FftPlan* p) {
    p->n = n;
    p->nst = 0;
    int m = n;
    while (m % 4 == 0) { p->rad[p->nst++] = 4; m /= 4; }
    while (m % 2 == 0) { p->rad[p->nst++] = 2; m /= 2; }
    while (m % 3 == 0) { p->rad[p->nst++] = 3; m /= 3; }
    while (m % 5 == 0) { p->rad[p->nst++] = 5; m /= 5; }
    return m == 1 && p->nst <= 20;
}

}  // namespace chs
