// THERM_PROP_UPDATE (src/THERM_PROP_UPDATE.f90:129-205) by table lookup on
// the device: per cell T = S_HT(h), then mu = S_TM(T), kappa = S_TK(T),
// rho = S_TD(T) and c_p = S_TCp(T) from the natural cubic splines of the
// scaled property table (NIST_SPLINE_H/T, CUBIC_SPLINE,
// src/Flow_State.f90:916-1014, 1693-1788).
//
// Layout: the knots of H and T (16 KB each at 2048 knots) and the bucket
// starts of the two interval searches are staged in LDS once per persistent
// block; the spline coefficients are read per cell as one record per interval
// (32 B for H -> T, 128 B = one cache line for T -> D, M, K, CP), so a cell
// costs its 8 B of h, its 40 B of outputs and two L2-resident records.
// The reference's four separate bisections over the same T knots collapse
// into one: they find the same interval.
#include <algorithm>
#include <cmath>

#include "props.hpp"

namespace chs {

// Natural-end cubic spline of Forsythe, Malcolm and Moler (the CUBIC_SPLINE
// of src/Flow_State.f90:1693-1788): s(u) = y_i + dx (b_i + dx (c_i + dx d_i))
// on [x_i, x_{i+1}), end conditions from third divided differences.  n >= 4.
static void fmm_spline(int n, const double* x, const double* y, double* b, double* c, double* d) {
#pragma clang fp contract(off)
    d[0] = x[1] - x[0];
    c[1] = (y[1] - y[0]) / d[0];
    for (int i = 1; i < n - 1; ++i) {
        d[i] = x[i + 1] - x[i];
        b[i] = 2.0 * (d[i - 1] + d[i]);
        c[i + 1] = (y[i + 1] - y[i]) / d[i];
        c[i] = c[i + 1] - c[i];
    }
    b[0] = -d[0];
    b[n - 1] = -d[n - 2];
    c[0] = c[2] / (x[3] - x[1]) - c[1] / (x[2] - x[0]);
    c[n - 1] = c[n - 2] / (x[n - 1] - x[n - 3]) - c[n - 3] / (x[n - 2] - x[n - 4]);
    c[0] = c[0] * d[0] * d[0] / (x[3] - x[0]);
    c[n - 1] = -c[n - 1] * d[n - 2] * d[n - 2] / (x[n - 1] - x[n - 4]);
    for (int i = 1; i < n; ++i) {
        const double t = d[i - 1] / b[i - 1];
        b[i] = b[i] - t * d[i - 1];
        c[i] = c[i] - t * c[i - 1];
    }
    c[n - 1] = c[n - 1] / b[n - 1];
    for (int i = n - 2; i >= 0; --i) c[i] = (c[i] - d[i] * c[i + 1]) / b[i];
    b[n - 1] = (y[n - 1] - y[n - 2]) / d[n - 2] + d[n - 2] * (c[n - 2] + 2.0 * c[n - 1]);
    for (int i = 0; i < n - 1; ++i) {
        b[i] = (y[i + 1] - y[i]) / d[i] - d[i] * (c[i + 1] + 2.0 * c[i]);
        d[i] = (c[i + 1] - c[i]) / d[i];
        c[i] = 3.0 * c[i];
    }
    c[n - 1] = 3.0 * c[n - 1];
    d[n - 1] = d[n - 2];
}

// bucket b covers [lo + b/bi, lo + (b+1)/bi); start[b] = last interval whose
// left knot is at or below the bucket's left edge
static void bucket_starts(int n, const double* x, int nb, double lo, double bi, std::vector<int>& start) {
    start.assign(nb, 0);
    int i = 0;
    for (int b = 0; b < nb; ++b) {
        const double edge = lo + (double)b / bi;
        while (i < n - 2 && x[i + 1] <= edge) ++i;
        start[b] = i;
    }
}

__global__ void k_zero_u64(unsigned long long* c, int n) {
    if ((int)threadIdx.x < n) c[threadIdx.x] = 0ull;
}

constexpr int PT = 1024;     // threads per lookup block
constexpr int PCELL = 8;     // cells per thread and chunk (measured: 4 -> 8, 0.252 -> 0.240 ms)

template <int KW>
__global__ void __launch_bounds__(PT) k_props_table(PropTab t, Rows rw, const double* __restrict__ h,
                                                    double* __restrict__ To, double* __restrict__ rho,
                                                    double* __restrict__ mu, double* __restrict__ kap,
                                                    double* __restrict__ cp, unsigned long long* __restrict__ cnt,
                                                    int count_lookups) {
    extern __shared__ unsigned char smem[];
    double* xh = reinterpret_cast<double*>(smem);
    double* xt = xh + t.n;
    int* sh = reinterpret_cast<int*>(xt + t.n);
    int* st = sh + t.nbh;
    for (int e = threadIdx.x; e < t.n; e += blockDim.x) {
        xh[e] = t.xh[e];
        xt[e] = t.xt[e];
    }
    for (int e = threadIdx.x; e < t.nbh; e += blockDim.x) sh[e] = t.sh[e];
    for (int e = threadIdx.x; e < t.nbt; e += blockDim.x) st[e] = t.st[e];
    __syncthreads();
    const long long chunk = (long long)PT * PCELL;
    const long long cpr = (rw.rowlen + chunk - 1) / chunk;   // chunks per row
    const long long nch = cpr * rw.nrow;
    unsigned long long bad = 0, looked = 0;
    for (long long ch = blockIdx.x; ch < nch; ch += gridDim.x) {
        const long long r = rw.nrow == 1 ? 0 : ch / cpr;       // block-uniform
        const long long c0 = (ch - r * cpr) * chunk;
        const long long base = r * rw.stride;
        // all loads of h first (independent), then the lookups
        double hv[PCELL];
        bool in[PCELL];
#pragma unroll
        for (int u = 0; u < PCELL; ++u) {
            const long long c = c0 + u * PT + threadIdx.x;
            in[u] = c < rw.rowlen;
            hv[u] = in[u] ? h[base + c] : 0.0;
        }
        // S_HT for every cell at once: the H intervals (LDS) and their
        // coefficient records (L2), loaded together
        int ih[PCELL];
        double4 rr[PCELL];
#pragma unroll
        for (int u = 0; u < PCELL; ++u) {
            // NaN and out-of-range values search at h_lo (a float-to-int
            // conversion of them would be undefined); their results are unused
            const double hq = hv[u] > t.h_lo && hv[u] < t.h_hi ? hv[u] : t.h_lo;
            ih[u] = find_iv<KW>(xh, sh, t.n, t.nbh, t.h_lo, t.hbi, hq);
            rr[u] = reinterpret_cast<const double4*>(t.rh)[ih[u]];
        }
        double tv[PCELL];
        bool okh[PCELL];
#pragma unroll
        for (int u = 0; u < PCELL; ++u) {
            const bool nan = hv[u] != hv[u];   // a NaN passes the range tests and gives NaN
            okh[u] = nan || (hv[u] > t.h_lo && hv[u] < t.h_hi);
            tv[u] = nan ? hv[u] : cubic_nc(rr[u].x, rr[u].y, rr[u].z, rr[u].w, hv[u] - xh[ih[u]]);
        }
        // S_TD/TM/TK/TCp two cells at a time: both cells' T records are
        // loaded before either is evaluated
#pragma unroll
        for (int u0 = 0; u0 < PCELL; u0 += 2) {
            int kt[2];
            double4 rq[2][4];
#pragma unroll
            for (int v = 0; v < 2; ++v) {
                const double tq = tv[u0 + v] > t.t_lo && tv[u0 + v] < t.t_hi ? tv[u0 + v] : t.t_lo;
                kt[v] = find_iv<KW>(xt, st, t.n, t.nbt, t.t_lo, t.tbi, tq);
                const double4* q = reinterpret_cast<const double4*>(t.rt) + 4 * (size_t)kt[v];
#pragma unroll
                for (int m = 0; m < 4; ++m) rq[v][m] = q[m];
            }
#pragma unroll
            for (int v = 0; v < 2; ++v) {
                const int u = u0 + v;
                // THERM_PROP_UPDATE leaves |h| < 1e-10 (and cells past the row) as they are
                if (!in[u] || fabs(hv[u]) < 1.0e-10) continue;
                ++looked;
                const double T = tv[u];
                const bool nan = T != T;
                if (!okh[u] || (!nan && (T <= t.t_lo || T >= t.t_hi))) {
                    ++bad;
                    continue;
                }
                double D, Mu, K, CP;
                if (nan) {
                    D = Mu = K = CP = T;
                } else {
                    const double dt = T - xt[kt[v]];
                    D = cubic_nc(rq[v][0].x, rq[v][0].y, rq[v][0].z, rq[v][0].w, dt);
                    Mu = cubic_nc(rq[v][1].x, rq[v][1].y, rq[v][1].z, rq[v][1].w, dt);
                    K = cubic_nc(rq[v][2].x, rq[v][2].y, rq[v][2].z, rq[v][2].w, dt);
                    CP = cubic_nc(rq[v][3].x, rq[v][3].y, rq[v][3].z, rq[v][3].w, dt);
                    if (fabs(T - 1.0) < 1.0e-10) D = Mu = K = CP = 1.0;
                }
                const long long q = base + c0 + u * PT + threadIdx.x;
                To[q] = T;
                mu[q] = Mu;
                kap[q] = K;
                rho[q] = D;
                if (cp) cp[q] = CP;
            }
        }
    }
    // out-of-range count (and, on request, the lookups): one atomic per wave
    for (int o = 32; o > 0; o >>= 1) {
        bad += __shfl_xor(bad, o);
        looked += __shfl_xor(looked, o);
    }
    if ((threadIdx.x & 63) == 0) {
        if (bad) atomicAdd(cnt, bad);
        if (count_lookups && looked) atomicAdd(cnt + 1, looked);
    }
}

int props_table_launch(chs_props* p, const double* h, Rows rw, double* T, double* rho, double* mu, double* kap,
                       double* cp, unsigned long long* cnt, int count_lookups, hipStream_t stream) {
    const long long chunk = (long long)PT * PCELL;
    const long long nch = ((rw.rowlen + chunk - 1) / chunk) * rw.nrow;
    if (nch == 0) return CHS_OK;
#define CHS_PROPS_K(KK)                                                                                          \
    do {                                                                                                         \
        CHS_TRY(props_lds_attr((const void*)k_props_table<KK>, p));                                              \
        const int grid = props_grid_occ(p, (const void*)k_props_table<KK>, PT, nch);                             \
        CHS_LAUNCH(k_props_table<KK>, dim3(grid), dim3(PT), props_lds_bytes(p), stream, tab_of(p), rw, h, T, rho, \
                   mu, kap, cp, cnt, count_lookups);                                                             \
    } while (0)
    switch (p->kwalk) {
        case 2: CHS_PROPS_K(2); break;
        case 3: CHS_PROPS_K(3); break;
        case 4: CHS_PROPS_K(4); break;
        default: CHS_PROPS_K(0); break;
    }
#undef CHS_PROPS_K
    return CHS_OK;
}

int props_create(int n, const double* H, const double* T, const double* D, const double* M, const double* K,
                 const double* CP, void* stream, chs_props** out) {
    CHS_CHECK(out != nullptr, CHS_E_ARG, "chs_props_create: out is NULL");
    *out = nullptr;
    CHS_CHECK(n >= 4 && n <= 4096, CHS_E_ARG, "chs_props_create: need 4 <= n <= 4096 knots");
    CHS_CHECK(H && T && D && M && K && CP, CHS_E_ARG, "chs_props_create: NULL table column");
    for (int i = 0; i + 1 < n; ++i) {
        CHS_CHECK(H[i] < H[i + 1], CHS_E_ARG, "chs_props_create: H must increase strictly");
        CHS_CHECK(T[i] < T[i + 1], CHS_E_ARG, "chs_props_create: T must increase strictly");
    }
    for (int i = 0; i < n; ++i)
        CHS_CHECK(std::isfinite(H[i]) && std::isfinite(T[i]) && std::isfinite(D[i]) && std::isfinite(M[i]) &&
                      std::isfinite(K[i]) && std::isfinite(CP[i]),
                  CHS_E_ARG, "chs_props_create: non-finite table value");
    chs_props* p = new chs_props();
    p->n = n;
    p->nbh = p->nbt = n;
    p->h_lo = H[0];
    p->h_hi = H[n - 1];
    p->t_lo = T[0];
    p->t_hi = T[n - 1];
    p->hbi = (double)p->nbh / (p->h_hi - p->h_lo);
    p->tbi = (double)p->nbt / (p->t_hi - p->t_lo);
    // splines H -> T and T -> D, M, K, CP
    std::vector<double> b(n), c(n), d(n);
    p->hH.assign(H, H + n);
    p->hT.assign(T, T + n);
    p->hrh.assign(4 * (size_t)n, 0.0);
    p->hrt.assign(16 * (size_t)n, 0.0);
    fmm_spline(n, H, T, b.data(), c.data(), d.data());
    for (int i = 0; i < n; ++i) {
        p->hrh[4 * i + 0] = T[i];
        p->hrh[4 * i + 1] = b[i];
        p->hrh[4 * i + 2] = c[i];
        p->hrh[4 * i + 3] = d[i];
    }
    const double* ys[4] = {D, M, K, CP};
    for (int f = 0; f < 4; ++f) {
        fmm_spline(n, T, ys[f], b.data(), c.data(), d.data());
        for (int i = 0; i < n; ++i) {
            p->hrt[16 * i + 4 * f + 0] = ys[f][i];
            p->hrt[16 * i + 4 * f + 1] = b[i];
            p->hrt[16 * i + 4 * f + 2] = c[i];
            p->hrt[16 * i + 4 * f + 3] = d[i];
        }
    }
    // RHOHmax as NIST_SPLINE_H finds it (src/Flow_State.f90:960-990): the
    // largest D(H) * H of the H -> D spline over 3n - 1 equidistant points
    {
        fmm_spline(n, H, D, b.data(), c.data(), d.data());
        double best = -1.0e-14;
        for (int I = 2; I <= 3 * n; ++I) {
#pragma clang fp contract(off)
            const double hv = H[0] + (double)(I - 1) * (H[n - 1] - H[0]) / (double)(3 * n);
            const int i = (int)(std::upper_bound(H, H + n, hv) - H) - 1;   // H[i] <= hv < H[i+1]
            const double dx = hv - H[i];
            double dv = D[i] + dx * (b[i] + dx * (c[i] + dx * d[i]));
            if (std::fabs(hv) < 1.0e-10) dv = 1.0;
            if (dv * hv > best) best = dv * hv;
        }
        p->rhoh_max = best;
    }
    // bucket counts: n each, doubled while a bucket holds more than two
    // knots and the staged tables stay within 64 KB of LDS; the lookups then
    // walk a fixed 1 + (most knots in a bucket) steps (kwalk <= 4), or to the
    // end of the walk when the knots are too uneven for that
    std::vector<int> sh, st;
    auto most = [&](const std::vector<int>& start) {
        int m = 0;
        for (size_t b = 0; b < start.size(); ++b) {
            const int nx = b + 1 < start.size() ? start[b + 1] : n - 2;
            m = std::max(m, nx - start[b]);
        }
        return m;
    };
    int mh = 0, mt = 0;
    for (;;) {
        bucket_starts(n, H, p->nbh, p->h_lo, p->hbi, sh);
        bucket_starts(n, T, p->nbt, p->t_lo, p->tbi, st);
        mh = most(sh);
        mt = most(st);
        const bool gh = mh > 2, gt = mt > 2;
        const size_t lds = sizeof(double) * 2 * (size_t)n +
                           sizeof(int) * ((size_t)p->nbh * (gh ? 2 : 1) + (size_t)p->nbt * (gt ? 2 : 1));
        if ((!gh && !gt) || lds > 65536) break;
        if (gh) p->nbh *= 2;
        if (gt) p->nbt *= 2;
        p->hbi = (double)p->nbh / (p->h_hi - p->h_lo);
        p->tbi = (double)p->nbt / (p->t_hi - p->t_lo);
    }
    const int kw = std::max(mh, mt) + 1;
    p->kwalk = kw <= 4 ? std::max(kw, 2) : 0;
    auto fail = [&](int code) {
        chs_props_destroy(p);
        return code;
    };
    if (stream) {
        p->stream = (hipStream_t)stream;
    } else {
        if (hipStreamCreate(&p->stream) != hipSuccess) return fail(CHS_E_HIP);
        p->own_stream = true;
    }
    {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
            p->lds_max = v;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
            p->ncu = v;
    }
    if (props_lds_bytes(p) > (size_t)p->lds_max) {
        set_error("chs_props_create: the search tables of " + std::to_string(n) + " knots (" +
                  std::to_string(props_lds_bytes(p)) + " B) exceed the device's LDS per block (" +
                  std::to_string(p->lds_max) + " B)");
        return fail(CHS_E_ARG);
    }
    const size_t bytes = sizeof(double) * (2 * (size_t)n + 20 * (size_t)n) + sizeof(int) * ((size_t)p->nbh + p->nbt);
    unsigned char* blk = nullptr;
    if (hipMalloc((void**)&blk, bytes + 256) != hipSuccess) {
        set_error("chs_props_create: device allocation failed");
        return fail(CHS_E_NOMEM);
    }
    p->rh = reinterpret_cast<double*>(blk);   // 32 B records first (256 B aligned)
    p->rt = p->rh + 4 * (size_t)n;
    p->xh = p->rt + 16 * (size_t)n;
    p->xt = p->xh + n;
    p->sh = reinterpret_cast<int*>(p->xt + n);
    p->st = p->sh + p->nbh;
    if (hipMalloc((void**)&p->counters, 8 * sizeof(unsigned long long)) != hipSuccess ||
        hipHostMalloc((void**)&p->ecnt_host, 2 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess) {
        set_error("chs_props_create: allocation of the counters failed");
        return fail(CHS_E_NOMEM);
    }
    p->ecnt = p->counters + 4;
    p->ecnt_host[0] = p->ecnt_host[1] = 0;
    if (hipEventCreateWithFlags(&p->ecnt_ev, hipEventDisableTiming) != hipSuccess) {
        set_error("chs_props_create: event creation failed");
        return fail(CHS_E_HIP);
    }
    if (hipMemcpy(p->rh, p->hrh.data(), p->hrh.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->rt, p->hrt.data(), p->hrt.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->xh, H, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->xt, T, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->sh, sh.data(), sh.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->st, st.data(), st.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
        set_error("chs_props_create: table upload failed");
        return fail(CHS_E_HIP);
    }
    *out = p;
    return CHS_OK;
}

void mlp_release(chs_props* p);
bool mlp_ready(const chs_props* p);
int mlp_update(chs_props* p, const double* h, Rows rw, double* T, double* rho, double* mu, double* kap, double* cp,
               unsigned long long* cnt, hipStream_t stream);

int props_destroy(chs_props* p) {
    if (!p) return CHS_OK;
    if (p->ecnt_pending) (void)hipEventSynchronize(p->ecnt_ev);   // the copy into ecnt_host
    mlp_release(p);
    if (p->rh) (void)hipFree(p->rh);
    if (p->counters) (void)hipFree(p->counters);
    if (p->ecnt_host) (void)hipHostFree(p->ecnt_host);
    if (p->ecnt_ev) (void)hipEventDestroy(p->ecnt_ev);
    if (p->own_stream && p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
    return CHS_OK;
}

// The range check of the last chs_energy_substep: waits for its counters'
// copy (once), then reports cells outside the table once, as CHS_E_RANGE.
void energy_range_resolve(chs_props* p) {
    if (!p->ecnt_pending) return;
    if (hipEventSynchronize(p->ecnt_ev) != hipSuccess) {
        p->ecnt_bad = ~0ull;   // unknown: reported as an error below
    } else {
        p->ecnt_bad = p->ecnt_host[0];
        p->last_fallback = p->ecnt_sur ? (long long)p->ecnt_host[1] : 0;
    }
    p->ecnt_pending = false;
}

int energy_range_check(chs_props* p, const char* who) {
    energy_range_resolve(p);
    const unsigned long long bad = p->ecnt_bad;
    if (!bad) return CHS_OK;
    p->ecnt_bad = 0;
    if (bad == ~0ull) {
        set_error(std::string(who) + ": the range check of the last chs_energy_substep could not be read");
        return CHS_E_HIP;
    }
    set_error(std::string(who) + ": the last chs_energy_substep had " + std::to_string(bad) +
              " property lookups outside the table");
    return CHS_E_RANGE;
}

int props_update(chs_props* p, int mode, const double* h, long long nrow, long long rowlen, long long stride,
                 double* T, double* rho, double* mu, double* kap, double* cp, long long* n_fallback) {
    CHS_CHECK(p != nullptr, CHS_E_ARG, "chs_therm_prop_update: NULL handle");
    CHS_CHECK(nrow >= 0 && rowlen >= 0 && (nrow <= 1 || stride >= rowlen), CHS_E_ARG,
              "chs_therm_prop_update: bad row description");
    CHS_CHECK(mode == CHS_PROPS_TABLE || mode == CHS_PROPS_SURROGATE, CHS_E_ARG,
              "chs_therm_prop_update: unknown mode");
    if (n_fallback) *n_fallback = 0;
    if (nrow == 0 || rowlen == 0) return CHS_OK;
    CHS_CHECK(h && T && rho && mu && kap, CHS_E_ARG, "chs_therm_prop_update: NULL array");
    Rows rw{nrow, rowlen, nrow > 1 ? stride : rowlen};
    CHS_LAUNCH(k_zero_u64, dim3(1), dim3(64), 0, p->stream, p->counters, 4);
    if (mode == CHS_PROPS_TABLE)
        CHS_TRY(props_table_launch(p, h, rw, T, rho, mu, kap, cp, p->counters, 0, p->stream));
    else
        CHS_TRY(mlp_update(p, h, rw, T, rho, mu, kap, cp, p->counters, p->stream));
    unsigned long long host[2] = {0, 0};
    CHS_HIP(hipMemcpyAsync(host, p->counters, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, p->stream));
    CHS_HIP(hipStreamSynchronize(p->stream));
    if (n_fallback) *n_fallback = (long long)host[1];
    if (host[0]) {
        set_error("chs_therm_prop_update: " + std::to_string(host[0]) +
                  " cells outside the property table (left unchanged)");
        return CHS_E_RANGE;
    }
    return CHS_OK;
}

}  // namespace chs

extern "C" {

int chs_props_create(int n, const double* H, const double* T, const double* D, const double* M, const double* K,
                     const double* CP, void* stream, chs_props** out) {
    return chs::props_create(n, H, T, D, M, K, CP, stream, out);
}

int chs_props_destroy(chs_props* p) { return chs::props_destroy(p); }

int chs_props_set_mode(chs_props* p, int mode) {
    CHS_CHECK(p != nullptr, CHS_E_ARG, "chs_props_set_mode: NULL handle");
    CHS_CHECK(mode == CHS_PROPS_TABLE || mode == CHS_PROPS_SURROGATE, CHS_E_ARG, "chs_props_set_mode: unknown mode");
    CHS_CHECK(mode == CHS_PROPS_TABLE || (p->mlp && chs::mlp_ready(p)), CHS_E_ARG,
              "chs_props_set_mode: the surrogate is not trained and certified");
    p->mode = mode;
    return CHS_OK;
}

int chs_props_last_fallback(chs_props* p, long long* n_fallback) {
    CHS_CHECK(p != nullptr && n_fallback != nullptr, CHS_E_ARG, "chs_props_last_fallback: NULL argument");
    chs::energy_range_resolve(p);
    *n_fallback = p->last_fallback;
    return chs::energy_range_check(p, "chs_props_last_fallback");
}

int chs_props_check(chs_props* p) {
    CHS_CHECK(p != nullptr, CHS_E_ARG, "chs_props_check: NULL handle");
    return chs::energy_range_check(p, "chs_props_check");
}

int chs_props_rhoh_max(chs_props* p, double* rhoh_max) {
    CHS_CHECK(p != nullptr && rhoh_max != nullptr, CHS_E_ARG, "chs_props_rhoh_max: NULL argument");
    *rhoh_max = p->rhoh_max;
    return CHS_OK;
}

int chs_therm_prop_update(chs_props* p, int mode, const double* h_dev, long long nrow, long long rowlen,
                          long long row_stride, double* T_dev, double* rho_dev, double* mu_dev, double* kap_dev,
                          double* cp_dev, long long* n_fallback) {
    return chs::props_update(p, mode, h_dev, nrow, rowlen, row_stride, T_dev, rho_dev, mu_dev, kap_dev, cp_dev,
                             n_fallback);
}

}  // extern "C"
