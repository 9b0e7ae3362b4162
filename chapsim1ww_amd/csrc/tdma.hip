// Standalone batched tridiagonal solves: the C-ABI chs_tdma_batched, which
// stands in for the reference's four line solvers
//   TDMAIJJ_nonCYC (src/TDMAIJJ_nonCYC.f90:67-123)   TDMAIJI_nonCYC (src/TDMAIJI_nonCYC.f90:67-123)
//   TDMAIJJ_CYC    (src/TDMAIJJ_CYC.f90:67-157)      TDMAIJI_CYC    (src/TDMAIJI_CYC.f90:67-157)
// on arbitrary line/element strides.
#include "chs_common.hpp"
#include "tdma_dev.hpp"

namespace chs {

// Segmented solver: LPB lines per block, P = ceil(n / MS) segments per line.
template <int MS>
__global__ void __launch_bounds__(1024) k_tdma_seg(const double* __restrict__ A, const double* __restrict__ B,
                                                   const double* __restrict__ Cc, double* __restrict__ Rr,
                                                   const double* __restrict__ bc, int n, int nlines,
                                                   long long es, long long ls, int LPB, int P, int cyclic) {
    extern __shared__ unsigned char smem[];
    const int tid = threadIdx.x;
    // consecutive lanes take consecutive lines, except for cyclic lines with
    // contiguous rows (es = 1, IJI), where they take consecutive segments of
    // one line (measured at 65536 x 192: 0.252 -> 0.183 ms; 0.203 -> 0.319 ms
    // for non-cyclic lines, which keep the line-major order)
    const bool rowc = es == 1 && cyclic;
    const int ll = rowc ? tid / P : tid % LPB, seg = rowc ? tid % P : tid / LPB;
    const long long line = (long long)blockIdx.x * LPB + ll;
    const bool valid = line < nlines;
    double a[MS], b[MS], c[MS], d[MS];
    const long long base = valid ? line * ls : 0;
    // branch-free loads (rows past n read row n - 1 of a valid line), all
    // issued before the first wait; padding rows become identity rows
#pragma unroll
    for (int l = 0; l < MS; ++l) {
        const long long o = base + (long long)min(seg * MS + l, n - 1) * es;
        a[l] = A[o];
        b[l] = B[o];
        c[l] = Cc[o];
        d[l] = Rr[o];
    }
#pragma unroll
    for (int l = 0; l < MS; ++l) {
        if (!(valid && seg * MS + l < n)) {   // padding rows: identity, decoupled
            a[l] = 0.0;
            b[l] = 1.0;
            c[l] = 0.0;
            d[l] = 0.0;
        }
    }
    if (!cyclic && valid) {
        // the end rows take the boundary values BC(line,1:2) (TDMAIJJ_nonCYC.f90:93-97,108-113)
        if (seg == 0) {
            if (bc) d[0] = d[0] - a[0] * bc[line];
            a[0] = 0.0;
        }
        const int jl = n - 1;
        if (jl / MS == seg) {
            const int l = jl % MS;
#pragma unroll
            for (int q = 0; q < MS; ++q)
                if (q == l) {
                    if (bc) d[q] = d[q] - c[q] * bc[nlines + line];
                    c[q] = 0.0;
                }
        }
    }
    seg_solve<MS, double>(a, b, c, d, LPB, P, ll, seg, valid, cyclic != 0, smem);
    if (valid) {
#pragma unroll
        for (int l = 0; l < MS; ++l) {
            const int j = seg * MS + l;
            if (j < n) Rr[base + (long long)j * es] = d[l];
        }
    }
}

// Fallback for shapes the segmented kernel does not cover: one thread per
// line, the reference recurrences, H/G in a device workspace.
__global__ void k_tdma_seq(const double* __restrict__ A, const double* __restrict__ B,
                           const double* __restrict__ Cc, double* __restrict__ Rr, const double* __restrict__ bc,
                           int n, int nlines, long long es, long long ls, int cyclic, double* __restrict__ ws) {
    const long long line = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (line >= nlines) return;
    const double* a = A + line * ls;
    const double* b = B + line * ls;
    const double* c = Cc + line * ls;
    double* r = Rr + line * ls;
    // workspace rows are interleaved over lines: coalesced
    double* H = ws + line;
    double* G = ws + (long long)n * nlines + line;
    double* G2 = ws + 2LL * n * nlines + line;
    const long long wst = nlines;
    if (!cyclic) {
        const double blo = bc ? bc[line] : 0.0, bhi = bc ? bc[nlines + line] : 0.0;
        H[0] = c[0] / b[0];
        G[0] = (r[0] - a[0] * blo) / b[0];
        for (int j = 1; j < n - 1; ++j) {
            const double pp = 1.0 / (b[j * es] - H[(j - 1) * wst] * a[j * es]);
            H[j * wst] = c[j * es] * pp;
            G[j * wst] = (r[j * es] - a[j * es] * G[(j - 1) * wst]) * pp;
        }
        const int j = n - 1;
        const double pp = 1.0 / (b[j * es] - H[(j - 1) * wst] * a[j * es]);
        double x = (r[j * es] - c[j * es] * bhi - a[j * es] * G[(j - 1) * wst]) * pp;
        r[j * es] = x;
        for (int jj = n - 2; jj >= 0; --jj) {
            x = G[jj * wst] - H[jj * wst] * x;
            r[jj * es] = x;
        }
        return;
    }
    H[(n - 2) * wst] = 0.0;
    H[0] = c[0] / b[0];
    G[0] = r[0] / b[0];
    G2[0] = -a[0] / b[0];
    for (int i = 1; i <= n - 3; ++i) {
        const double pp = 1.0 / (b[i * es] - H[(i - 1) * wst] * a[i * es]);
        const double q1 = G[(i - 1) * wst] * a[i * es], q2 = G2[(i - 1) * wst] * a[i * es];
        H[i * wst] = c[i * es] * pp;
        G[i * wst] = (r[i * es] - q1) * pp;
        G2[i * wst] = -q2 * pp;
    }
    {
        const int i = n - 2;
        const double pp = 1.0 / (b[i * es] - H[(i - 1) * wst] * a[i * es]);
        const double q1 = G[(i - 1) * wst] * a[i * es], q2 = G2[(i - 1) * wst] * a[i * es];
        G[i * wst] = (r[i * es] - q1) * pp;
        G2[i * wst] = (-c[i * es] - q2) * pp;
    }
    for (int i = n - 3; i >= 0; --i) {
        G[i * wst] = -H[i * wst] * G[(i + 1) * wst] + G[i * wst];
        G2[i * wst] = -H[i * wst] * G2[(i + 1) * wst] + G2[i * wst];
    }
    const int i = n - 1;
    const double pp = r[i * es] - c[i * es] * G[0] - a[i * es] * G[(i - 1) * wst];
    const double qq = b[i * es] + c[i * es] * G2[0] + a[i * es] * G2[(i - 1) * wst];
    const double xn = pp / qq;
    r[i * es] = xn;
    for (int k = n - 2; k >= 0; --k) r[k * es] = G[k * wst] + G2[k * wst] * xn;
}

// Choose the segmented configuration for (n, cyclic) within maxthr threads per
// block, or return false.  8-row segments keep the solver within 128 VGPRs.
// lpb_small: blocks of 8 lines (several blocks per CU overlap their load,
// reduce and store phases; measured faster for the complex-valued Poisson
// lines, slower for the real batched solver)
#ifndef CHS_TDMA_LPB
#define CHS_TDMA_LPB 64
#endif
bool tdma_seg_config(int n, bool cyclic, int* MS, int* P, int* LPB, int maxthr, bool lpb_small) {
    const int cand[4] = {8, 12, 16, 4};
    const int lpbs_big[2] = {CHS_TDMA_LPB, CHS_TDMA_LPB == 64 ? 32 : CHS_TDMA_LPB}, lpbs_small[2] = {8, 8};
    const int* lpbs = lpb_small ? lpbs_small : lpbs_big;
    for (int ms : cand) {
        for (int q = 0; q < 2; ++q) {
            const int lpb = lpbs[q];
            const int p = (n + ms - 1) / ms;
            if (p * lpb > maxthr) continue;
            // cyclic lines cannot be padded (the wrap couples the last real row)
            if (cyclic && (n % ms != 0 || p < 2)) continue;
            *MS = ms;
            *P = p;
            *LPB = lpb;
            return true;
        }
    }
    return false;
}

// ws: the sequential fallback's workspace of 3 n nlines doubles, or nullptr to
// take it from the stream-ordered pool for this launch
int tdma_launch(const double* a, const double* b, const double* c, double* r, const double* bc, int n, int nlines,
                long long es, long long ls, int cyclic, hipStream_t s, double* ws_given = nullptr) {
    int MS = 0, P = 0, LPB = 0;
    bool seg = tdma_seg_config(n, cyclic != 0, &MS, &P, &LPB, 1024, false);
    // measured at 65536 lines x 192 rows: lines contiguous (ls = 1, the IJJ
    // layout) fastest in 12-row segments, 32 lines per block (0.141 -> 0.116
    // ms non-cyclic, 0.168 -> 0.126 ms cyclic); rows contiguous (IJI) in
    // 8-row segments, 16 lines per block (0.229 -> 0.200 ms non-cyclic) or 8
    // for cyclic lines
    {
        const int ms = ls == 1 ? 12 : 8, lpb = ls == 1 ? 32 : (cyclic ? 8 : 16);
        const int p = (n + ms - 1) / ms;
        if (seg && p * lpb <= 1024 && (!cyclic || (n % ms == 0 && p >= 2))) {
            MS = ms;
            P = p;
            LPB = lpb;
        }
    }
    if (seg) {
        const dim3 grid((nlines + LPB - 1) / LPB), block(LPB * P);
        const size_t sm = seg_smem_bytes<double>(LPB, P);
        switch (MS) {
            case 16: CHS_LAUNCH(k_tdma_seg<16>, grid, block, sm, s, a, b, c, r, bc, n, nlines, es, ls, LPB, P, cyclic); break;
            case 12: CHS_LAUNCH(k_tdma_seg<12>, grid, block, sm, s, a, b, c, r, bc, n, nlines, es, ls, LPB, P, cyclic); break;
            case 4: CHS_LAUNCH(k_tdma_seg<4>, grid, block, sm, s, a, b, c, r, bc, n, nlines, es, ls, LPB, P, cyclic); break;
            default: CHS_LAUNCH(k_tdma_seg<8>, grid, block, sm, s, a, b, c, r, bc, n, nlines, es, ls, LPB, P, cyclic); break;
        }
        return CHS_OK;
    }
    const int bs = 256;
    if (ws_given) {
        CHS_LAUNCH(k_tdma_seq, dim3((nlines + bs - 1) / bs), dim3(bs), 0, s, a, b, c, r, bc, n, nlines, es, ls, cyclic,
                   ws_given);
        return CHS_OK;
    }
    double* ws = nullptr;
    CHS_HIP(hipMallocAsync((void**)&ws, sizeof(double) * 3 * (size_t)n * nlines, s));
    CHS_LAUNCH(k_tdma_seq, dim3((nlines + bs - 1) / bs), dim3(bs), 0, s, a, b, c, r, bc, n, nlines, es, ls, cyclic, ws);
    CHS_HIP(hipFreeAsync(ws, s));
    return CHS_OK;
}

// Workspace of the allocation-free entry point: one copy per GPU, held until
// the process ends.  Host code never passes `scratch` itself to a kernel (its
// host-side name is not a device address); it uses the address that
// hipGetSymbolAddress returns on each device, kept in scratch_dev[].
constexpr size_t TDMA_SCRATCH_DOUBLES = size_t(4) << 20;   // 32 MiB
static __device__ __align__(256) double scratch[TDMA_SCRATCH_DOUBLES];

static std::mutex scratch_mu;
static double* scratch_dev[64] = {};

// The current device's scratch address, looked up once per device.
static int scratch_address(double** out) {
    int dev = 0;
    CHS_HIP(hipGetDevice(&dev));
    CHS_CHECK(dev >= 0 && dev < 64, CHS_E_ARG, "chs_tdma_batched_noalloc: device ordinal above 63");
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (!scratch_dev[dev]) {
        void* p = nullptr;
        CHS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        scratch_dev[dev] = static_cast<double*>(p);
    }
    *out = scratch_dev[dev];
    return CHS_OK;
}

}  // namespace chs

extern "C" int chs_tdma_batched(const double* a_dev, const double* b_dev, const double* c_dev, double* r_dev,
                                const double* bc_dev, int n, int nlines, long long elem_stride,
                                long long line_stride, int cyclic, void* stream) {
    CHS_CHECK(a_dev && b_dev && c_dev && r_dev, CHS_E_ARG, "chs_tdma_batched: null array");
    CHS_CHECK(n >= 3, CHS_E_ARG, "chs_tdma_batched: n must be >= 3 (the reference reads outside its arrays below that)");
    CHS_CHECK(nlines >= 0, CHS_E_ARG, "chs_tdma_batched: nlines < 0");
    CHS_CHECK(elem_stride != 0 || n == 1, CHS_E_ARG, "chs_tdma_batched: zero element stride");
    if (nlines == 0) return CHS_OK;
    return chs::tdma_launch(a_dev, b_dev, c_dev, r_dev, cyclic ? nullptr : bc_dev, n, nlines, elem_stride,
                            line_stride, cyclic, (hipStream_t)stream);
}

// chs_tdma_batched without any allocation: the segmented kernel needs only
// LDS; the sequential fallback (lines the segmented kernel cannot split) takes
// its 3 n nlines-double workspace from this GPU's static scratch array.
extern "C" int chs_tdma_batched_noalloc(const double* a_dev, const double* b_dev, const double* c_dev, double* r_dev,
                                        const double* bc_dev, int n, int nlines, long long elem_stride,
                                        long long line_stride, int cyclic, void* stream) {
    CHS_CHECK(a_dev && b_dev && c_dev && r_dev, CHS_E_ARG, "chs_tdma_batched_noalloc: null array");
    CHS_CHECK(n >= 3, CHS_E_ARG, "chs_tdma_batched_noalloc: n must be >= 3 (the reference reads outside its arrays below that)");
    CHS_CHECK(nlines >= 0, CHS_E_ARG, "chs_tdma_batched_noalloc: nlines < 0");
    CHS_CHECK(elem_stride != 0, CHS_E_ARG, "chs_tdma_batched_noalloc: zero element stride");
    if (nlines == 0) return CHS_OK;
    int MS = 0, P = 0, LPB = 0;
    double* ws = nullptr;
    if (!chs::tdma_seg_config(n, cyclic != 0, &MS, &P, &LPB, 1024, false)) {
        CHS_CHECK(3 * (size_t)n * (size_t)nlines <= chs::TDMA_SCRATCH_DOUBLES, CHS_E_ARG,
                  "chs_tdma_batched_noalloc: these lines need a workspace of 3*n*nlines doubles, above the "
                  "4194304 of the static scratch array (use chs_tdma_batched)");
        const int st = chs::scratch_address(&ws);
        if (st != CHS_OK) return st;
    }
    return chs::tdma_launch(a_dev, b_dev, c_dev, r_dev, cyclic ? nullptr : bc_dev, n, nlines, elem_stride,
                            line_stride, cyclic, (hipStream_t)stream, ws);
}

// Host-array form for a drop-in of the reference subroutines called on host
// (Fortran) arrays: copies the lines to the device, solves, copies r back.
// Element (l, j) at l*line_stride + j*elem_stride within each host array of
// span = (nlines-1)*line_stride + (n-1)*elem_stride + 1 doubles.
extern "C" int chs_tdma_batched_host(const double* a, const double* b, const double* c, double* r,
                                     const double* bc, int n, int nlines, long long elem_stride,
                                     long long line_stride, int cyclic) {
    CHS_CHECK(a && b && c && r, CHS_E_ARG, "chs_tdma_batched_host: null array");
    CHS_CHECK(n >= 3 && nlines >= 0 && elem_stride > 0 && line_stride > 0, CHS_E_ARG,
              "chs_tdma_batched_host: need n >= 3 and positive strides");
    if (nlines == 0) return CHS_OK;
    const size_t span = (size_t)(nlines - 1) * line_stride + (size_t)(n - 1) * elem_stride + 1;
    double* d = nullptr;
    const size_t nb = bc && !cyclic ? 2 * (size_t)nlines : 0;
    CHS_HIP(hipMalloc((void**)&d, sizeof(double) * (4 * span + nb)));
    int st = CHS_OK;
    const double* src[4] = {a, b, c, r};
    for (int q = 0; q < 4 && st == CHS_OK; ++q)
        if (hipMemcpy(d + q * span, src[q], span * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) st = CHS_E_HIP;
    if (st == CHS_OK && nb && hipMemcpy(d + 4 * span, bc, nb * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        st = CHS_E_HIP;
    if (st == CHS_OK)
        st = chs::tdma_launch(d, d + span, d + 2 * span, d + 3 * span, nb ? d + 4 * span : nullptr, n, nlines,
                              elem_stride, line_stride, cyclic, nullptr);
    if (st == CHS_OK && hipMemcpy(r, d + 3 * span, span * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        st = CHS_E_HIP;
    (void)hipFree(d);
    if (st == CHS_E_HIP) chs::set_error("chs_tdma_batched_host: HIP copy or launch failed");
    return st;
}
