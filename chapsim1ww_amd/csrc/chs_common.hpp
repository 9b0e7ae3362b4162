// Shared definitions of the CHAPSim MI355X library: error handling, the
// handle, complex helpers and launch utilities.  gfx950 only.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/chapsim_gpu.h"

namespace chs {

// ---------------------------------------------------------------------------
// errors: every export returns 0 or a CHS_E_* code; the last message is kept
// per thread and returned by chs_last_error().
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);

#define CHS_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ::chs::set_error(std::string(#call) + ": " + hipGetErrorString(e_));        \
            return CHS_E_HIP;                                                           \
        }                                                                               \
    } while (0)

#define CHS_CHECK(cond, code, msg)                                                      \
    do {                                                                                \
        if (!(cond)) {                                                                  \
            ::chs::set_error(msg);                                                      \
            return (code);                                                              \
        }                                                                               \
    } while (0)

#define CHS_TRY(call)                                                                   \
    do {                                                                                \
        int s_ = (call);                                                                \
        if (s_ != CHS_OK) return s_;                                                    \
    } while (0)

// launch + immediate launch-error check
#define CHS_LAUNCH(kernel, grid, block, shmem, stream, ...)                             \
    do {                                                                                \
        hipLaunchKernelGGL(kernel, grid, block, shmem, stream, __VA_ARGS__);            \
        hipError_t e_ = hipGetLastError();                                              \
        if (e_ != hipSuccess) {                                                         \
            ::chs::set_error(std::string("launch " #kernel ": ") + hipGetErrorString(e_)); \
            return CHS_E_HIP;                                                           \
        }                                                                               \
    } while (0)

// ---------------------------------------------------------------------------
// complex arithmetic on HIP vector types
// ---------------------------------------------------------------------------
template <typename R> struct cpx;
template <> struct cpx<double> { using t = double2; };
template <> struct cpx<float> { using t = float2; };
template <typename R> using cpx_t = typename cpx<R>::t;

__host__ __device__ inline double2 mk(double r, double i) { return make_double2(r, i); }
__host__ __device__ inline float2 mk(float r, float i) { return make_float2(r, i); }
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ inline double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline float2 cmul(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ inline float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ inline double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ inline float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by -i (dir = -1) or +i (dir = +1)
template <typename C> __device__ inline C cmul_i(C a, int dir) {
    return dir < 0 ? C{a.y, -a.x} : C{-a.y, a.x};
}

// scalar-or-complex helpers used by the tridiagonal solver (real matrices,
// real or complex right-hand sides)
__device__ inline double vscale(double v, double s) { return v * s; }
__device__ inline double2 vscale(double2 v, double s) { return make_double2(v.x * s, v.y * s); }
__device__ inline double vaxpy(double v, double a, double x) { return v - a * x; }        // v - a x
__device__ inline double2 vaxpy(double2 v, double a, double2 x) { return make_double2(v.x - a * x.x, v.y - a * x.y); }
__device__ inline double vfma(double v, double a, double x) { return v + a * x; }         // v + a x
__device__ inline double2 vfma(double2 v, double a, double2 x) { return make_double2(v.x + a * x.x, v.y + a * x.y); }
template <typename V> __device__ inline V vzero();
template <> __device__ inline double vzero<double>() { return 0.0; }
template <> __device__ inline double2 vzero<double2>() { return make_double2(0.0, 0.0); }

// ---------------------------------------------------------------------------
// MI355X has 8 XCDs with private L2s; consecutive workgroup ids are dealt
// round-robin over the XCDs.  Remap so that each XCD walks a contiguous range
// of tiles: neighbouring tiles (which share halo rows) then share an L2.
// ---------------------------------------------------------------------------
// A read of a 1-D coefficient array, which no kernel writes, through the
// constant address space: with a uniform index it becomes a scalar load
// (SGPRs, counted apart from the wave's vector loads, so that using the
// value does not wait for the loads issued before it).
__device__ __forceinline__ double ldc(const double* p, int idx) {
    return ((const __attribute__((address_space(4))) double*)p)[idx];
}

// One array around the cell (IC, JC, KC): a(di, dj, dk) is its value at
// (IC+di, JC+dj, KC+dk), each of di, dj, dk in -1..1 and a constant at every
// call site (x and z periodic).  CellAcc reads a ghost-aware global array;
// CellRing reads the LDS ring of the marching kernels, which holds the rows
// JC-1, JC, JC+1 as separate planes with the z neighbours as halo rows.
struct CellAcc {
    const double* __restrict__ b;  // the array (uniform), and the cell's offset in it
    int c, ip, im, kp, km, L;
    __device__ __forceinline__ double operator()(int di, int dj, int dk) const {
        return b[c + (di > 0 ? ip : di < 0 ? im : 0) + dj * L + (dk > 0 ? kp : dk < 0 ? km : 0)];
    }
};
struct CellRing {
    const double *m, *c, *p;  // rows JC-1, JC, JC+1, at the cell
    int ip, im, L;
    __device__ __forceinline__ double operator()(int di, int dj, int dk) const {
        return (dj < 0 ? m : dj > 0 ? p : c)[(di > 0 ? ip : di < 0 ? im : 0) + dk * L];
    }
};

__device__ __forceinline__ CellAcc cell_acc(const double* base, int c, int ip, int im, int kp, int km, int L) {
    return CellAcc{base, c, ip, im, kp, km, L};
}

__device__ inline int xcd_tile(int bid, int nblk) {
    const int nx = 8;
    int per = (nblk + nx - 1) / nx;
    int t = (bid % nx) * per + bid / nx;
    return t;  // may be >= nblk for the padding blocks when nblk % 8 != 0
}

// ---------------------------------------------------------------------------
// per-kernel timing with HIP events (enabled by chs_profile_enable)
// ---------------------------------------------------------------------------
struct KernelTimer {
    bool on = false;
    struct Rec { std::string name; hipEvent_t a, b; };
    std::vector<Rec> pending;
    std::vector<std::pair<std::string, std::pair<double, long>>> totals;  // name -> (ms, count)
    void begin(const char* name, hipStream_t s, hipEvent_t* out_a);
    void end(hipEvent_t a, const char* name, hipStream_t s);
    void collect();
};

// Tiles of 64 (i) x 4 (k) cells of one row j, run one per block of a 1-D
// grid of 8 * ceil(ntiles / 8) blocks.  Block b runs on XCD b % 8; tile
// t = (b % 8) * per + b / 8 gives each XCD a contiguous range of tiles,
// ordered i-tile fastest, then j, then k-tile, so the tiles an XCD runs
// back to back are j-neighbours and their j +- 1 rows are in its L2.
// Returns false for the padding blocks.
__device__ __forceinline__ bool row_tile(int L, int N, int nj, int* it, int* j, int* kt, int* tile) {
    const int ni = (L + 63) / 64, nk = (N + 3) / 4;
    const int nt = ni * nj * nk, per = (nt + 7) / 8;
    const int b = blockIdx.x;
    const int t = (b % 8) * per + b / 8;
    if (t >= nt) return false;
    *it = t % ni;
    const int r = t / ni;
    *j = r % nj;
    *kt = r / nj;
    *tile = t;
    return true;
}

// blocks of the 1-D grid for row_tile over nj rows
inline unsigned row_tile_blocks(int L, int N, int nj) {
    const long nt = (long)((L + 63) / 64) * ((N + 3) / 4) * nj;
    return (unsigned)(8 * ((nt + 7) / 8));
}

// Rows per y chunk of a kernel that marches along y with an LDS ring, one
// block per group of k rows (nkg groups) and chunk of rows: about one block
// per CU (256 on MI355X), or, with fill, one for each block that the ring's
// LDS (smem bytes of lds_max) lets a CU hold, at most four; chunks of at
// least 4 rows, none empty.
// A march shorter than MARCH_MIN rows (small grids) loses to the row-tile
// kernels, which need no ring fill: measured at 64^3, FP64 substep 74.9 ->
// 71.1 us, heated substep 122.1 -> 111.3 us (4-row chunks); at 128^3 (8 and
// 16 rows) the marching kernels win (244 against 251 us, 492 against 533 us).
constexpr int MARCH_MIN = 8;
inline int march_chunk_rows(int rows, int nkg, size_t smem, int lds_max, bool fill) {
    const int bpc = fill ? std::max(1, std::min(4, (int)((size_t)lds_max / std::max<size_t>(smem, 1)))) : 1;
    const int nyc = std::max(1, std::min(rows / 4, (256 * bpc + nkg - 1) / nkg));
    return (rows + nyc - 1) / nyc;
}

struct CnCache;   // truncated-inverse CN stencils of one handle (substep.hip)

}  // namespace chs

// The handle behind the opaque chs_handle of the C ABI.
struct chs_handle {
    int L, M, N, LH, LHP;
    int LHP32;           // row length of the FP32 spectrum (LH rounded up to 16)
    int icase, flowdrvtp, precision;
    int nflow;      // streamwise direction: 1 (x) or 3 (z)
    double cfgv;    // FLOWDRVTP = 2: DPGRNS = -CFGV/2 TALP DT
    double dxi, dzi, dxqi, dzqi, cvisc, volm;
    hipStream_t stream;
    bool own_stream;
    // 1-D coefficients on the device (copied at create)
    double *dyfi, *dyci, *wff, *wfb, *amph, *acph, *apph, *amvr, *acvr, *apvr, *ak1, *ak3;
    // FFT twiddle tables W_n^e = exp(-2 pi i e / n), FP64 and FP32 copies
    double2 *twx, *twz;
    float2 *twx32, *twz32;
    // state in the reference layouts
    double *q, *pr, *convh0;
    // CONVH0 in FP32 (mixed form, FP32 stencils): h0f holds the current
    // CONVH0 while h0_f32 is set, convh0 otherwise
    float* h0f;
    bool h0_f32;
    // work: DPH(L,0:M+1,N), RHS x3 (L,M,N), RHSLLPHI(L,M,N), spectral S
    double *dph, *rhs, *rhsllphi;
    double *rhs2;        // second RHS buffer of the stencil CN path (z output)
    int cn_mode;         // 0 = auto (stencils when their truncation bound holds), 1 = segmented Thomas,
                         // 2 = auto with the per-row RHS kernel instead of the y-marching one
    int lds_max;         // LDS bytes a block may use on the handle's device
    double2 *spec;
    float2 *spec32;
    double *rhs_j1;      // the u-increment on j = 1, which becomes v's row-1 RHS
    double *partials;    // per-block partial sums / maxima
    double *scalars;     // [0] DPGRNS, [1] max|div|, [2] CFL, [3..] scratch
    double *resid;       // mixed-precision residual / correction fields
    float *f32a, *f32b;
    int npart;
    chs::KernelTimer timer;
    // host copies of AMVR/ACVR/APVR (chs_create) and the handle's cache of
    // truncated-inverse CN stencils, keyed by FACOE (substep.hip)
    std::vector<double> h_amvr, h_acvr, h_apvr;
    std::mutex cn_mu;
    chs::CnCache* cn_cache;
    // xz-plane means of the last chs_mean_xz (13 M, tg) and chs_mean_xz_io
    // (37 M) call and their time means (PP_MEAN_T_*), in one allocation:
    // stats + [0, 13M) tg xz, [13M, 50M) io xz, [50M, 63M) tg time, [63M, 100M) io time
    double *stats;
    int nstatis[2];      // NSTATIS_TG, NSTATIS_IO
    // the mixed-precision substep's convergence check, fused into its
    // projection and read back without a wait: the largest max|r| / bound
    // since the last reported error and the latest max|r|, max|f| land in
    // pinned chk_host when chk_ev completes; chk_tol is the latest bound
    bool chk_pending;
    double chk_tol;
    double* chk_host;
    hipEvent_t chk_ev;
    // mixed-precision Poisson refinement: tolerance, cap, statistics of the last call
    double refine_tol;
    int refine_max;
    int last_refine_iters;
    double last_refine_resid;
};
