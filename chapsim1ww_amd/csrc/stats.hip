// xz-plane statistics of the heated channel and the time means of both
// domains:
//   chs_mean_xz_io  PP_MEAN_ZX_FLOW_Xperiodic_IO (src/processing/PP_MEAN_ZX.f90:299-800):
//                   the first and second moments of u, g = rho u, p and the
//                   thermal moments of T, rho, h, mu (T1..PH, UH, GH)
//   chs_mean_t      PP_MEAN_T_TG / PP_MEAN_T_Xperiodic_IO
//                   (src/processing/PP_MEAN_T.f90:322-370, 436-470)
//   chs_mean_t_reset PP_TMEAN_INI_ZERO_TG / _IO (src/processing/PP_MEAN_T.f90:29-54, 110-319)
// One block per row j sums its L x N cells in a fixed order (per-thread
// strided partial sums, then a tree over the block), so every result is
// deterministic.  The last xz means stay on the device (chs_handle::stats)
// for the time means.
#include "chs_common.hpp"

namespace chs {

constexpr int NIO = 37;   // moments per row of chs_mean_xz_io
constexpr int NTG = 13;   // moments per row of chs_mean_xz

struct StatGeo {
    int L, M, N;
};

// Row J = blockIdx.x + 1 of the (L, 0:M+1, N[,3]) io arrays.  Column order
// as in include/chapsim_gpu.h (chs_mean_xz_io).  The channel has RCCI1 = 1,
// so v and w's cell averages carry the factor 1.0 the reference multiplies.
__global__ void __launch_bounds__(128) k_mean_xz_io(StatGeo g, const double* __restrict__ Q,
                                                    const double* __restrict__ Gf, const double* __restrict__ PR,
                                                    const double* __restrict__ T, const double* __restrict__ RHO,
                                                    const double* __restrict__ H, const double* __restrict__ MU,
                                                    int thermal, double* __restrict__ out) {
    __shared__ double red[NIO][128];
    const int L = g.L, N = g.N, J = blockIdx.x + 1;
    const int LP = L * (g.M + 2), CS = LP * N;
    double acc[NIO];
#pragma unroll
    for (int m = 0; m < NIO; ++m) acc[m] = 0.0;
    for (int e = threadIdx.x; e < L * N; e += blockDim.x) {
        const int k = e / L, i = e - k * L;
        const int ip = i == L - 1 ? 0 : i + 1, kp = k == N - 1 ? 0 : k + 1;
        const int o = i + L * J + LP * k, oip = o - i + ip, okp = o + LP * (kp - k);
        double U[3], G[3];
        G[0] = (Gf[o] + Gf[oip]) * 0.5;
        U[0] = (Q[o] + Q[oip]) * 0.5;
        G[2] = ((Gf[2 * CS + o] + Gf[2 * CS + okp]) * 1.0) * 0.5;
        U[2] = ((Q[2 * CS + o] + Q[2 * CS + okp]) * 1.0) * 0.5;
        G[1] = ((Gf[CS + o] + Gf[CS + o + L]) * 1.0) * 0.5;
        U[1] = ((Q[CS + o] + Q[CS + o + L]) * 1.0) * 0.5;
        const double P = PR[o];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            acc[m] += U[m];
            acc[4 + m] += G[m];
            acc[7 + m] += U[m] * P;
        }
        acc[3] += P;
        int l = 0;
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int n = m; n < 3; ++n, ++l) {
                acc[16 + l] += U[m] * G[n];
                acc[10 + l] += U[m] * U[n];
            }
        if (thermal) {
            const double t = T[o], d = RHO[o], hh = H[o], v = MU[o];
            acc[22] += t;
            acc[23] += d;
            acc[24] += hh;
            acc[25] += v;
            acc[26] += t * t;
            acc[27] += d * d;
            acc[28] += hh * hh;
            acc[29] += hh * d;
            acc[30] += hh * P;
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                acc[31 + m] += hh * U[m];
                acc[34 + m] += hh * G[m];
            }
        }
    }
#pragma unroll
    for (int m = 0; m < NIO; ++m) red[m][threadIdx.x] = acc[m];
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
#pragma unroll
            for (int m = 0; m < NIO; ++m) red[m][threadIdx.x] += red[m][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x < NIO) {
        const double VL = 1.0 / (double)(L * N);
        out[NIO * (J - 1) + threadIdx.x] = red[threadIdx.x][0] * VL;
    }
}

// acc = coe3 acc + coe2 val over n values
__global__ void k_time_mean(double* __restrict__ acc, const double* __restrict__ val, int n, double coe2,
                            double coe3) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) acc[i] = coe3 * acc[i] + coe2 * val[i];
}

// a zero fill that a caller may capture in a graph (no memset node)
__global__ void k_zero(double* __restrict__ a, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0;
}

static double* xz_of(chs_handle* h, int dom) { return h->stats + (dom == 0 ? 0 : (size_t)NTG * h->M); }
static double* t_of(chs_handle* h, int dom) {
    return h->stats + (size_t)(NTG + NIO) * h->M + (dom == 0 ? 0 : (size_t)NTG * h->M);
}
static int width_of(int dom) { return dom == 0 ? NTG : NIO; }

int mean_xz_io(chs_handle* h, const chs_thermal_fields* tf, const chs_io_fields* f, int thermal, double* out_host) {
    CHS_CHECK(h && f && f->q && f->g && f->pr, CHS_E_ARG, "chs_mean_xz_io: NULL argument");
    CHS_CHECK(thermal == 0 || thermal == 1, CHS_E_ARG, "chs_mean_xz_io: thermal must be 0 or 1");
    if (thermal)
        CHS_CHECK(tf && tf->temperature && tf->density && tf->enthalpy && tf->viscosity, CHS_E_ARG,
                  "chs_mean_xz_io: thermal = 1 needs TEMPERATURE, DENSITY, ENTHALPY and VISCOUSITY");
    const hipStream_t s = h->stream;
    double* xz = xz_of(h, 1);
    const StatGeo g{h->L, h->M, h->N};
    CHS_LAUNCH(k_mean_xz_io, dim3(h->M), dim3(128), 0, s, g, f->q, f->g, f->pr, thermal ? tf->temperature : nullptr,
               thermal ? tf->density : nullptr, thermal ? tf->enthalpy : nullptr, thermal ? tf->viscosity : nullptr,
               thermal, xz);
    if (out_host) {
        CHS_HIP(hipMemcpyAsync(out_host, xz, (size_t)NIO * h->M * sizeof(double), hipMemcpyDeviceToHost, s));
        CHS_HIP(hipStreamSynchronize(s));
    }
    return CHS_OK;
}

int mean_t(chs_handle* h, int dom, int tp, int pp_instn_sz, int* nstatis, double* out_host) {
    CHS_CHECK(h, CHS_E_ARG, "chs_mean_t: NULL handle");
    CHS_CHECK(dom == 0 || dom == 1, CHS_E_ARG, "chs_mean_t: domain must be 0 (tg) or 1 (io)");
    CHS_CHECK(tp == 1 || tp == 2, CHS_E_ARG, "chs_mean_t: TP must be 1 (asymptotic) or 2 (summing)");
    CHS_CHECK(tp == 1 || pp_instn_sz > 0, CHS_E_ARG, "chs_mean_t: TP = 2 needs pp_instn_sz > 0");
    const int ns = ++h->nstatis[dom];
    double coe2, coe3;
    if (tp == 1) {
        coe2 = 1.0 / (double)ns;
        coe3 = (double)(ns - 1) / (double)ns;
    } else {
        coe2 = 1.0 / (double)pp_instn_sz;
        coe3 = 1.0;
    }
    const int n = width_of(dom) * h->M;
    const hipStream_t s = h->stream;
    CHS_LAUNCH(k_time_mean, dim3((n + 255) / 256), dim3(256), 0, s, t_of(h, dom), xz_of(h, dom), n, coe2, coe3);
    if (nstatis) *nstatis = ns;
    if (out_host) {
        CHS_HIP(hipMemcpyAsync(out_host, t_of(h, dom), (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s));
        CHS_HIP(hipStreamSynchronize(s));
    }
    return CHS_OK;
}

int mean_t_reset(chs_handle* h, int dom) {
    CHS_CHECK(h, CHS_E_ARG, "chs_mean_t_reset: NULL handle");
    CHS_CHECK(dom == 0 || dom == 1, CHS_E_ARG, "chs_mean_t_reset: domain must be 0 (tg) or 1 (io)");
    h->nstatis[dom] = 0;
    const int n = width_of(dom) * h->M;
    CHS_LAUNCH(k_zero, dim3((n + 255) / 256), dim3(256), 0, h->stream, t_of(h, dom), n);
    return CHS_OK;
}

}  // namespace chs

extern "C" {

int chs_mean_xz_io(chs_handle* h, const chs_thermal_fields* tf, const chs_io_fields* f, int thermal,
                   double* out_host) {
    return chs::mean_xz_io(h, tf, f, thermal, out_host);
}

int chs_mean_t(chs_handle* h, int domain, int tp, int pp_instn_sz, int* nstatis, double* out_host) {
    return chs::mean_t(h, domain, tp, pp_instn_sz, nstatis, out_host);
}

int chs_mean_t_reset(chs_handle* h, int domain) { return chs::mean_t_reset(h, domain); }

}  // extern "C"
