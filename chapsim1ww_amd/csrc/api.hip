// C ABI of the library: handle life cycle, state transfer, and the exports
// declared in include/chapsim_gpu.h.
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "chs_common.hpp"

namespace chs {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

int rk3_substep(chs_handle* h, int ns, double dt);
int divgck_cfl(chs_handle* h, double* div_out, double* cfl_out);
int mean_xz(chs_handle* h, double* out_host);
int mixed_check(chs_handle* h, bool wait);
int convh0_store(chs_handle* h, bool f32);
int poisson_solve(chs_handle* h, const double* f, double* dph, int ghost, double abs_tol);
bool substep_supported(const chs_handle* h);
void release_stencils(chs_handle* h);
int cn_last_path(chs_handle* h, double facoe, int* kx, int* kz, int* ky);
bool tdma_seg_config(int n, bool cyclic, int* MS, int* P, int* LPB, int maxthr, bool lpb_small = false);

void KernelTimer::begin(const char* name, hipStream_t s, hipEvent_t* out_a) {
    (void)name;
    *out_a = nullptr;
    if (!on) return;
    hipEvent_t a;
    if (hipEventCreate(&a) != hipSuccess) return;
    (void)hipEventRecord(a, s);
    *out_a = a;
}

void KernelTimer::end(hipEvent_t a, const char* name, hipStream_t s) {
    if (!on || !a) return;
    hipEvent_t b;
    if (hipEventCreate(&b) != hipSuccess) return;
    (void)hipEventRecord(b, s);
    pending.push_back({name, a, b});
}

void KernelTimer::collect() {
    for (auto& r : pending) {
        (void)hipEventSynchronize(r.b);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, r.a, r.b);
        bool found = false;
        for (auto& t : totals)
            if (t.first == r.name) {
                t.second.first += ms;
                t.second.second += 1;
                found = true;
            }
        if (!found) totals.push_back({r.name, {(double)ms, 1L}});
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    pending.clear();
}

template <typename T>
static int dalloc(T** p, size_t n) {
    *p = nullptr;
    if (n == 0) return CHS_OK;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc failed: ") + hipGetErrorString(e));
        *p = nullptr;
        return CHS_E_NOMEM;
    }
    return CHS_OK;
}

static int upload(double** dst, const double* src, size_t n, hipStream_t s) {
    CHS_TRY(dalloc(dst, n));
    CHS_HIP(hipMemcpyAsync(*dst, src, n * sizeof(double), hipMemcpyHostToDevice, s));
    return CHS_OK;
}

// exp(-2 pi i e / n), e = 0..n-1, from long-double sin/cos (correctly rounded
// to double for these arguments)
static std::vector<double2> twiddles(int n) {
    std::vector<double2> w(n);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int e = 0; e < n; ++e) {
        const long double th = 2.0L * pi * (long double)e / (long double)n;
        w[e] = make_double2((double)cosl(th), (double)-sinl(th));
    }
    return w;
}

static bool fft_len_ok(int n) {
    if (n < 4 || n % 2) return false;
    int m = n;
    for (int p : {2, 3, 5})
        while (m % p == 0) m /= p;
    return m == 1;
}

static void free_all(chs_handle* h) {
    void* ptrs[] = {h->dyfi, h->dyci, h->wff, h->wfb, h->amph, h->acph, h->apph, h->amvr, h->acvr, h->apvr,
                    h->ak1, h->ak3, h->twx, h->twz, h->twx32, h->twz32, h->q, h->pr, h->convh0, h->dph, h->rhs,
                    h->rhsllphi, h->spec, h->spec32, h->rhs_j1, h->partials, h->scalars, h->resid, h->f32a, h->f32b,
                    h->rhs2, h->stats, h->h0f};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
}

}  // namespace chs

using namespace chs;

extern "C" {

const char* chs_last_error(void) { return g_last_error.c_str(); }
const char* chs_version(void) { return "chapsim1ww_amd 0.1 (gfx950)"; }

int chs_create(const chs_grid_desc* d, void* stream, chs_handle** out) {
    CHS_CHECK(d && out, CHS_E_ARG, "chs_create: null argument");
    *out = nullptr;
    CHS_CHECK(d->icase == 1, CHS_E_ARG, "chs_create: only the channel (ICASE=1) is supported");
    CHS_CHECK(d->flowdrvtp == 1 || d->flowdrvtp == 2, CHS_E_ARG,
              "chs_create: FLOWDRVTP must be 1 (constant mass flux) or 2 (constant pressure gradient)");
    CHS_CHECK(d->nflow == 0 || d->nflow == 1 || d->nflow == 3, CHS_E_ARG,
              "chs_create: NFLOW must be 1 (x) or 3 (z): the wall-normal component has no RHS on its wall row");
    CHS_CHECK(std::isfinite(d->cfgv), CHS_E_ARG, "chs_create: CFGV must be finite");
    CHS_CHECK(d->precision == CHS_FP64 || d->precision == CHS_MIXED, CHS_E_ARG, "chs_create: bad precision");
    CHS_CHECK(fft_len_ok(d->ncl1) && fft_len_ok(d->ncl3), CHS_E_ARG,
              "chs_create: NCL1 and NCL3 must be even, >= 4, with prime factors 2, 3, 5 only (FFTFAX)");
    CHS_CHECK(d->ncl2 >= 4, CHS_E_ARG, "chs_create: NCL2 must be >= 4");
    CHS_CHECK(3.0 * d->ncl1 * (d->ncl2 + 2.0) * d->ncl3 < 2147483647.0, CHS_E_ARG,
              "chs_create: grid too large for 32-bit element offsets (3*NCL1*(NCL2+2)*NCL3 >= 2^31)");
    int ms, p, lpb;
    CHS_CHECK(tdma_seg_config(d->ncl2, false, &ms, &p, &lpb, 768), CHS_E_ARG, "chs_create: unsupported NCL2");
    const double* arrs[] = {d->dyfi, d->dyci, d->wff, d->wfb, d->amph, d->acph, d->apph,
                            d->amvr, d->acvr, d->apvr, d->ak1, d->ak3};
    for (const double* a : arrs) CHS_CHECK(a != nullptr, CHS_E_ARG, "chs_create: null coefficient array");

    chs_handle* h = new chs_handle();  // value-initialised: every pointer starts null
    h->L = d->ncl1;
    h->M = d->ncl2;
    h->N = d->ncl3;
    h->LH = h->L / 2 + 1;
    h->LHP = ((h->LH + 7) / 8) * 8;
    h->LHP32 = ((h->LH + 15) / 16) * 16;
    h->icase = d->icase;
    h->flowdrvtp = d->flowdrvtp;
    h->nflow = d->nflow == 0 ? 1 : d->nflow;
    h->cfgv = d->cfgv;
    h->precision = d->precision;
    h->dxi = d->dxi; h->dzi = d->dzi; h->dxqi = d->dxqi; h->dzqi = d->dzqi;
    h->cvisc = d->cvisc; h->volm = d->volm;
    h->refine_tol = 1e-12;
    h->refine_max = 8;
    if (stream) {
        h->stream = (hipStream_t)stream;
        h->own_stream = false;
    } else {
        if (hipStreamCreate(&h->stream) != hipSuccess) {
            delete h;
            set_error("chs_create: hipStreamCreate failed");
            return CHS_E_HIP;
        }
        h->own_stream = true;
    }
    {
        int dev = 0, v = 0;
        h->lds_max = hipGetDevice(&dev) == hipSuccess &&
                             hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess
                         ? v
                         : 65536;
    }
    const size_t L = h->L, M = h->M, N = h->N;
    const size_t ncell = L * M * N, ngh = L * (M + 2) * N;
    int st = CHS_OK;
    auto chk = [&](int s) { if (st == CHS_OK) st = s; };
    const hipStream_t s = h->stream;
    chk(upload(&h->dyfi, d->dyfi, M, s));
    chk(upload(&h->dyci, d->dyci, M + 1, s));
    chk(upload(&h->wff, d->wff, M + 1, s));
    chk(upload(&h->wfb, d->wfb, M + 1, s));
    chk(upload(&h->amph, d->amph, M, s));
    chk(upload(&h->acph, d->acph, M, s));
    chk(upload(&h->apph, d->apph, M, s));
    chk(upload(&h->amvr, d->amvr, 3 * (M + 1), s));
    chk(upload(&h->acvr, d->acvr, 3 * (M + 1), s));
    chk(upload(&h->apvr, d->apvr, 3 * (M + 1), s));
    h->h_amvr.assign(d->amvr, d->amvr + 3 * (M + 1));
    h->h_acvr.assign(d->acvr, d->acvr + 3 * (M + 1));
    h->h_apvr.assign(d->apvr, d->apvr + 3 * (M + 1));
    chk(upload(&h->ak1, d->ak1, L, s));
    chk(upload(&h->ak3, d->ak3, N, s));
    std::vector<double2> wx = twiddles(h->L), wz = twiddles(h->N);
    std::vector<float2> wx32(wx.size()), wz32(wz.size());
    for (size_t i = 0; i < wx.size(); ++i) wx32[i] = make_float2((float)wx[i].x, (float)wx[i].y);
    for (size_t i = 0; i < wz.size(); ++i) wz32[i] = make_float2((float)wz[i].x, (float)wz[i].y);
    chk(dalloc(&h->twx, wx.size()));
    chk(dalloc(&h->twz, wz.size()));
    chk(dalloc(&h->twx32, wx.size()));
    chk(dalloc(&h->twz32, wz.size()));
    chk(dalloc(&h->q, ngh * 3));
    chk(dalloc(&h->pr, ngh));
    chk(dalloc(&h->convh0, ncell * 3));
    chk(dalloc(&h->dph, ngh));
    chk(dalloc(&h->rhs, ncell * 3));
    chk(dalloc(&h->rhs2, ncell * 3));
    chk(dalloc(&h->rhsllphi, ncell));
    const size_t nspec = (size_t)h->LHP * M * N;
    if (h->precision == CHS_FP64)
        chk(dalloc(&h->spec, nspec));
    else {
        chk(dalloc(&h->spec32, (size_t)h->LHP32 * M * N));
        chk(dalloc(&h->resid, ncell));
        chk(dalloc(&h->h0f, ncell * 3));   // CONVH0 of the FP32 stencils
    }
    h->npart = 65536;
    chk(dalloc(&h->partials, (size_t)h->npart * 2));
    chk(dalloc(&h->scalars, 16));
    chk(dalloc(&h->stats, 100 * (size_t)M));
    h->chk_pending = false;
    h->chk_tol = 0.0;
    if (st == CHS_OK && hipHostMalloc((void**)&h->chk_host, 4 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
        set_error("chs_create: pinned host allocation failed");
        h->chk_host = nullptr;
        st = CHS_E_NOMEM;
    }
    if (st == CHS_OK && hipEventCreateWithFlags(&h->chk_ev, hipEventDisableTiming) != hipSuccess) {
        set_error("chs_create: event creation failed");
        h->chk_ev = nullptr;
        st = CHS_E_HIP;
    }
    if (st != CHS_OK) {
        if (h->chk_host) (void)hipHostFree(h->chk_host);
        if (h->chk_ev) (void)hipEventDestroy(h->chk_ev);
        free_all(h);
        if (h->own_stream) (void)hipStreamDestroy(h->stream);
        delete h;
        return st;
    }
    CHS_HIP(hipMemcpyAsync(h->twx, wx.data(), wx.size() * sizeof(double2), hipMemcpyHostToDevice, s));
    CHS_HIP(hipMemcpyAsync(h->twz, wz.data(), wz.size() * sizeof(double2), hipMemcpyHostToDevice, s));
    CHS_HIP(hipMemcpyAsync(h->twx32, wx32.data(), wx32.size() * sizeof(float2), hipMemcpyHostToDevice, s));
    CHS_HIP(hipMemcpyAsync(h->twz32, wz32.data(), wz32.size() * sizeof(float2), hipMemcpyHostToDevice, s));
    CHS_HIP(hipMemsetAsync(h->q, 0, ngh * 3 * sizeof(double), s));
    CHS_HIP(hipMemsetAsync(h->pr, 0, ngh * sizeof(double), s));
    CHS_HIP(hipMemsetAsync(h->convh0, 0, ncell * 3 * sizeof(double), s));
    CHS_HIP(hipMemsetAsync(h->dph, 0, ngh * sizeof(double), s));
    CHS_HIP(hipMemsetAsync(h->rhs, 0, ncell * 3 * sizeof(double), s));
    CHS_HIP(hipMemsetAsync(h->rhsllphi, 0, ncell * sizeof(double), s));
    CHS_HIP(hipMemsetAsync(h->scalars, 0, 16 * sizeof(double), s));
    CHS_HIP(hipMemsetAsync(h->stats, 0, 100 * (size_t)M * sizeof(double), s));
    h->nstatis[0] = h->nstatis[1] = 0;
    CHS_HIP(hipStreamSynchronize(s));
    *out = h;
    return CHS_OK;
}

int chs_destroy(chs_handle* h) {
    if (!h) return CHS_OK;
    (void)hipStreamSynchronize(h->stream);
    h->timer.collect();
    release_stencils(h);
    free_all(h);
    if (h->chk_host) (void)hipHostFree(h->chk_host);
    if (h->chk_ev) (void)hipEventDestroy(h->chk_ev);
    if (h->own_stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return CHS_OK;
}

int chs_upload_state(chs_handle* h, const double* q, const double* pr, const double* convh0) {
    CHS_CHECK(h, CHS_E_ARG, "chs_upload_state: null handle");
    const size_t ncell = (size_t)h->L * h->M * h->N, ngh = (size_t)h->L * (h->M + 2) * h->N;
    if (q) CHS_HIP(hipMemcpyAsync(h->q, q, 3 * ngh * sizeof(double), hipMemcpyHostToDevice, h->stream));
    if (pr) CHS_HIP(hipMemcpyAsync(h->pr, pr, ngh * sizeof(double), hipMemcpyHostToDevice, h->stream));
    if (convh0) {
        CHS_HIP(hipMemcpyAsync(h->convh0, convh0, 3 * ncell * sizeof(double), hipMemcpyHostToDevice, h->stream));
        h->h0_f32 = false;   // the FP64 store holds CONVH0 now
    }
    CHS_HIP(hipStreamSynchronize(h->stream));
    return CHS_OK;
}

int chs_download_state(chs_handle* h, double* q, double* pr, double* convh0) {
    CHS_CHECK(h, CHS_E_ARG, "chs_download_state: null handle");
    const size_t ncell = (size_t)h->L * h->M * h->N, ngh = (size_t)h->L * (h->M + 2) * h->N;
    if (q) CHS_HIP(hipMemcpyAsync(q, h->q, 3 * ngh * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    if (pr) CHS_HIP(hipMemcpyAsync(pr, h->pr, ngh * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    if (convh0) {
        CHS_TRY(chs::convh0_store(h, false));   // widened from the mixed form's FP32 store
        CHS_HIP(hipMemcpyAsync(convh0, h->convh0, 3 * ncell * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    }
    CHS_HIP(hipStreamSynchronize(h->stream));
    return mixed_check(h, true);
}

int chs_device_arrays(chs_handle* h, double** q, double** pr, double** convh0, double** dph, double** rhsllphi) {
    CHS_CHECK(h, CHS_E_ARG, "chs_device_arrays: null handle");
    if (q) *q = h->q;
    if (pr) *pr = h->pr;
    if (convh0) {
        CHS_TRY(chs::convh0_store(h, false));   // widened from the mixed form's FP32 store
        *convh0 = h->convh0;
    }
    if (dph) *dph = h->dph;
    if (rhsllphi) *rhsllphi = h->rhsllphi;
    return CHS_OK;
}

int chs_rk3_substep(chs_handle* h, int ns, double dt) {
    CHS_CHECK(h, CHS_E_ARG, "chs_rk3_substep: null handle");
    CHS_CHECK(ns >= 0 && ns <= 3, CHS_E_ARG, "chs_rk3_substep: NS must be 0..3");
    CHS_CHECK(std::isfinite(dt) && dt > 0.0, CHS_E_ARG, "chs_rk3_substep: dt must be positive and finite");
    CHS_CHECK(substep_supported(h), CHS_E_ARG,
              "chs_rk3_substep: NCL1 must be a multiple of 8 (>= 16, <= 1024) and NCL3 a multiple of 4");
    return rk3_substep(h, ns, dt);
}

int chs_set_cvisc(chs_handle* h, double cvisc) {
    CHS_CHECK(h, CHS_E_ARG, "chs_set_cvisc: null handle");
    CHS_CHECK(std::isfinite(cvisc) && cvisc > 0.0, CHS_E_ARG, "chs_set_cvisc: CVISC must be positive and finite");
    // the CN stencils are cached by FACOE = alpha dt CVISC / 2, so the next
    // substep builds the ones of the new viscosity
    h->cvisc = cvisc;
    return CHS_OK;
}

int chs_set_cn_solver(chs_handle* h, int mode) {
    CHS_CHECK(h, CHS_E_ARG, "chs_set_cn_solver: null handle");
    CHS_CHECK(mode >= 0 && mode <= 3, CHS_E_ARG,
              "chs_set_cn_solver: mode must be 0 (auto), 1 (segmented Thomas), 2 (auto, per-row RHS kernel) or 3 "
              "(auto, y-marching RHS kernel wherever it fits)");
    h->cn_mode = mode;
    return CHS_OK;
}

int chs_cn_info(chs_handle* h, int ns, double dt, int* path, int* kx, int* kz, int* ky) {
    CHS_CHECK(h && path && kx && kz && ky, CHS_E_ARG, "chs_cn_info: null argument");
    CHS_CHECK(ns >= 0 && ns <= 3 && dt > 0.0, CHS_E_ARG, "chs_cn_info: bad NS or dt");
    static const double TG[4] = {1.0, 8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
    static const double TR[4] = {0.0, 0.0, -17.0 / 60.0, -5.0 / 12.0};
    const double facoe = 0.5 * (TG[ns] + TR[ns]) * dt * h->cvisc;
    const int p = cn_last_path(h, facoe, kx, kz, ky);
    CHS_CHECK(p >= 0, p, "chs_cn_info failed");
    *path = p;
    return CHS_OK;
}

int chs_poisson_fft99(chs_handle* h, const double* f, double* dph) {
    CHS_CHECK(h && f && dph, CHS_E_ARG, "chs_poisson_fft99: null argument");
    return poisson_solve(h, f, dph, 0, 0.0);
}

int chs_set_refinement(chs_handle* h, double rel_tol, int max_iter) {
    CHS_CHECK(h, CHS_E_ARG, "chs_set_refinement: null handle");
    CHS_CHECK(rel_tol > 0.0 && max_iter >= 0 && max_iter <= 100, CHS_E_ARG, "chs_set_refinement: bad values");
    h->refine_tol = rel_tol;
    h->refine_max = max_iter;
    return CHS_OK;
}

int chs_poisson_stats(chs_handle* h, int* iters, double* rel) {
    CHS_CHECK(h, CHS_E_ARG, "chs_poisson_stats: null handle");
    const int st = mixed_check(h, true);
    if (iters) *iters = h->last_refine_iters;
    if (rel) *rel = h->last_refine_resid;
    return st;
}

int chs_divgck(chs_handle* h, double* maxdiv) {
    CHS_CHECK(h && maxdiv, CHS_E_ARG, "chs_divgck: null argument");
    return divgck_cfl(h, maxdiv, nullptr);
}

int chs_cfl(chs_handle* h, double* rate) {
    CHS_CHECK(h && rate, CHS_E_ARG, "chs_cfl: null argument");
    return divgck_cfl(h, nullptr, rate);
}

int chs_mean_xz(chs_handle* h, double* out) {
    CHS_CHECK(h && out, CHS_E_ARG, "chs_mean_xz: null argument");
    return mean_xz(h, out);
}

int chs_last_dpgrns(chs_handle* h, double* v) {
    CHS_CHECK(h && v, CHS_E_ARG, "chs_last_dpgrns: null argument");
    CHS_HIP(hipMemcpyAsync(v, h->scalars, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    CHS_HIP(hipStreamSynchronize(h->stream));
    return CHS_OK;
}

int chs_profile_enable(chs_handle* h, int on) {
    CHS_CHECK(h, CHS_E_ARG, "chs_profile_enable: null handle");
    h->timer.collect();
    h->timer.on = on != 0;
    return CHS_OK;
}

int chs_profile_reset(chs_handle* h) {
    CHS_CHECK(h, CHS_E_ARG, "chs_profile_reset: null handle");
    h->timer.collect();
    h->timer.totals.clear();
    return CHS_OK;
}

int chs_profile_read(chs_handle* h, int cap, char* names, double* ms, long* calls, int* count) {
    CHS_CHECK(h && count, CHS_E_ARG, "chs_profile_read: null argument");
    h->timer.collect();
    const int n = (int)h->timer.totals.size();
    *count = n;
    for (int i = 0; i < n && i < cap; ++i) {
        const auto& t = h->timer.totals[i];
        if (names) {
            std::strncpy(names + 48 * i, t.first.c_str(), 47);
            names[48 * i + 47] = 0;
        }
        if (ms) ms[i] = t.second.first;
        if (calls) calls[i] = t.second.second;
    }
    return CHS_OK;
}

}  // extern "C"
