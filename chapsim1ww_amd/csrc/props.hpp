// Thermophysical-property state shared by the table lookup (props.hip) and
// the learned surrogate (mlp.hip).
#pragma once

#include "chs_common.hpp"

namespace chs {
struct Mlp;
}

// The property handle of the C ABI (declared opaque in chapsim_gpu.h).
struct chs_props {
    int n = 0;                 // knots of the scaled table
    int nbh = 0, nbt = 0;      // buckets of the H and T interval searches
    int kwalk = 0;             // fixed walk steps that make both searches exact (0: walk to the end)
    double h_lo = 0, h_hi = 0, t_lo = 0, t_hi = 0;
    double hbi = 0, tbi = 0;   // 1 / bucket width
    double rhoh_max = 0;       // RHOHmax of NIST_SPLINE_H (the limiter of the energy update)
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int lds_max = 65536;       // LDS bytes a block may use on the table's device
    int ncu = 256;             // compute units of that device
    // device table: knots, bucket starts, per-interval coefficient records
    double* xh = nullptr;      // H knots [n]
    double* xt = nullptr;      // T knots [n]
    int* sh = nullptr;         // H bucket -> first candidate interval [nbh]
    int* st = nullptr;         // T bucket -> first candidate interval [nbt]
    double* rh = nullptr;      // [n][4]: T, b, c, d of the H -> T spline
    double* rt = nullptr;      // [n][16]: (y, b, c, d) of T -> D, M, K, CP
    unsigned long long* counters = nullptr;   // [4] device counters
    // host copies used by the surrogate's certification
    std::vector<double> hH, hT, hrh, hrt;
    // learned surrogate (mlp.hip)
    chs::Mlp* mlp = nullptr;
    // the THERM_PROP_UPDATE form of chs_energy_substep (CHS_PROPS_TABLE or
    // CHS_PROPS_SURROGATE) and the cells its last call handed to the table
    int mode = 0;
    long long last_fallback = 0;
    // the energy substep's range check, read without a stream sync: its
    // device counters (after the four of the other calls), their pinned host
    // copy, the event that follows the copy, and a resolved count not yet
    // reported
    unsigned long long* ecnt = nullptr;
    unsigned long long* ecnt_host = nullptr;
    hipEvent_t ecnt_ev = nullptr;
    bool ecnt_pending = false, ecnt_sur = false;
    unsigned long long ecnt_bad = 0;
};

namespace chs {

// Rows of cells: row r holds rowlen consecutive values starting at r*stride
// (the interior j-block of one k plane of a ghost-aware array, or one flat run).
struct Rows {
    long long nrow, rowlen, stride;
};

// The range check of the last chs_energy_substep on p (props.hip):
// energy_range_resolve waits for its pending copy; energy_range_check also
// reports cells outside the table, once, as CHS_E_RANGE.
void energy_range_resolve(chs_props* p);
int energy_range_check(chs_props* p, const char* who);

// LDS bytes of the staged search tables of one lookup block
inline size_t props_lds_bytes(const chs_props* p) {
    return sizeof(double) * 2 * (size_t)p->n + sizeof(int) * ((size_t)p->nbh + p->nbt);
}

// Blocks of a persistent lookup kernel of `threads` threads with the staged
// tables in LDS: as many as are resident at once on the device (registers and
// LDS), at most `want`.
inline int props_grid_occ(const chs_props* p, const void* kernel, int threads, long long want) {
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, props_lds_bytes(p)) != hipSuccess ||
        per < 1)
        per = 1;
    const long long g = (long long)p->ncu * per;
    return (int)(want < g ? want : g);
}

// Allow a lookup kernel the staged tables' LDS when they exceed the default
// 64 KB per block (props_create has checked them against the device limit).
inline int props_lds_attr(const void* kernel, const chs_props* p) {
    const size_t b = props_lds_bytes(p);
    if (b > 65536) {
        hipError_t e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)b);
        if (e != hipSuccess) {
            set_error(std::string("hipFuncSetAttribute (property tables in LDS): ") + hipGetErrorString(e));
            return CHS_E_HIP;
        }
    }
    return CHS_OK;
}

// Interval i (0-based) with x[i] <= u < x[i+1], for x[0] < u < x[n-1]: the
// result of the reference's bisection (NIST_SLEVAL_*, src/Flow_State.f90:
// 1050-1061).  A bucket gives the first candidate; short walks make it exact
// whatever the rounding of the bucket index.
__device__ __forceinline__ int find_interval(const double* x, const int* start, int n, int nb, double lo,
                                             double bi, double u) {
    int b = (int)((u - lo) * bi);
    b = b < 0 ? 0 : (b >= nb ? nb - 1 : b);
    int i = start[b];
    while (i < n - 2 && x[i + 1] <= u) ++i;
    while (i > 0 && x[i] > u) --i;
    return i;
}

// find_interval with a fixed number of steps and no branch: at most one step
// back (the rounding of the bucket index) and K forward, K being one more
// than the most knots any bucket holds (props_create checks it), so that the
// searches of several cells interleave.
template <int K>
__device__ __forceinline__ int find_interval_k(const double* x, const int* start, int n, int nb, double lo,
                                               double bi, double u) {
    int b = (int)((u - lo) * bi);
    b = b < 0 ? 0 : (b >= nb ? nb - 1 : b);
    int i = start[b];
    i -= (i > 0 && x[i] > u) ? 1 : 0;
#pragma unroll
    for (int s = 0; s < K; ++s) i += (i < n - 2 && x[i + 1] <= u) ? 1 : 0;
    return i;
}

template <int K>
__device__ __forceinline__ int find_iv(const double* x, const int* start, int n, int nb, double lo, double bi,
                                       double u) {
    if constexpr (K > 0)
        return find_interval_k<K>(x, start, n, nb, lo, bi, u);
    else
        return find_interval(x, start, n, nb, lo, bi, u);
}

// y + dx*(b + dx*(c + dx*d)) evaluated as written, without contraction into
// fused multiply-adds (the same roundings as the reference's FP64 code).
__device__ __forceinline__ double cubic_nc(double y, double b, double c, double d, double dx) {
#pragma clang fp contract(off)
    return y + dx * (b + dx * (c + dx * d));
}

// device view of the table
struct PropTab {
    int n, nbh, nbt;
    double h_lo, h_hi, t_lo, t_hi, hbi, tbi;
    const double *xh, *xt, *rh, *rt;
    const int *sh, *st;
};

inline PropTab tab_of(const chs_props* p) {
    PropTab t;
    t.n = p->n;
    t.nbh = p->nbh;
    t.nbt = p->nbt;
    t.h_lo = p->h_lo;
    t.h_hi = p->h_hi;
    t.t_lo = p->t_lo;
    t.t_hi = p->t_hi;
    t.hbi = p->hbi;
    t.tbi = p->tbi;
    t.xh = p->xh;
    t.xt = p->xt;
    t.rh = p->rh;
    t.rt = p->rt;
    t.sh = p->sh;
    t.st = p->st;
    return t;
}

// S_HT (NIST_SLEVAL_HT, src/Flow_State.f90:1015-1070): false when h leaves
// the table (the reference STOPs); hack: |h| < 1e-10 gives T = 1.  A NaN
// passes the range tests and gives NaN, as there.
template <int KW = 0>
__device__ __forceinline__ bool spline_ht(const PropTab& t, const double* xh, const int* sh, double hv, double& T,
                                          bool hack) {
    if (hv != hv) {
        T = hv;
        return true;
    }
    if (hv <= t.h_lo || hv >= t.h_hi) return false;
    const int i = find_iv<KW>(xh, sh, t.n, t.nbh, t.h_lo, t.hbi, hv);
    const double4 r = reinterpret_cast<const double4*>(t.rh)[i];
    T = cubic_nc(r.x, r.y, r.z, r.w, hv - xh[i]);
    if (hack && fabs(hv) < 1.0e-10) T = 1.0;
    return true;
}

// S_TD, S_TM, S_TK, S_TCp at T (NIST_SLEVAL_TD/TM/TK/TCp, src/Flow_State.f90:
// 1415-1692): one interval search for all four; |T - 1| < 1e-10 gives
// D = M = K = CP = 1 (each evaluator's own rule, :1470, :1524, :1578, :1686).
template <int KW = 0>
__device__ __forceinline__ bool spline_t(const PropTab& t, const double* xt, const int* st, double tv, double& D,
                                         double& Mu, double& K, double& CP) {
    if (tv != tv) {
        D = Mu = K = CP = tv;
        return true;
    }
    if (tv <= t.t_lo || tv >= t.t_hi) return false;
    const int k = find_iv<KW>(xt, st, t.n, t.nbt, t.t_lo, t.tbi, tv);
    const double4* q = reinterpret_cast<const double4*>(t.rt) + 4 * (size_t)k;
    const double4 cd = q[0], cm = q[1], ck = q[2], cc = q[3];
    const double dt = tv - xt[k];
    D = cubic_nc(cd.x, cd.y, cd.z, cd.w, dt);
    Mu = cubic_nc(cm.x, cm.y, cm.z, cm.w, dt);
    K = cubic_nc(ck.x, ck.y, ck.z, ck.w, dt);
    CP = cubic_nc(cc.x, cc.y, cc.z, cc.w, dt);
    if (fabs(tv - 1.0) < 1.0e-10) D = Mu = K = CP = 1.0;
    return true;
}

// THERM_PROP_UPDATE of one cell (the caller skips |h| < 1e-10); false
// (outputs untouched) when h or T leaves the table.
// S_TD alone (DENSITY_UPDATE needs only the density): the same interval
// search and the same evaluation as spline_t, reading only the density's
// coefficients of the interval's record.
template <int KW = 0>
__device__ __forceinline__ bool spline_t_density(const PropTab& t, const double* xt, const int* st, double tv,
                                                 double& D) {
    if (tv != tv) {
        D = tv;
        return true;
    }
    if (tv <= t.t_lo || tv >= t.t_hi) return false;
    const int k = find_iv<KW>(xt, st, t.n, t.nbt, t.t_lo, t.tbi, tv);
    const double4 cd = reinterpret_cast<const double4*>(t.rt)[4 * (size_t)k];
    D = cubic_nc(cd.x, cd.y, cd.z, cd.w, tv - xt[k]);
    if (fabs(tv - 1.0) < 1.0e-10) D = 1.0;
    return true;
}

template <int KW = 0>
__device__ __forceinline__ bool lookup_cell(const PropTab& t, const double* xh, const double* xt, const int* sh,
                                            const int* st, double hv, double& T, double& D, double& Mu, double& K,
                                            double& CP) {
    double tv;
    if (!spline_ht<KW>(t, xh, sh, hv, tv, false)) return false;
    if (!spline_t<KW>(t, xt, st, tv, D, Mu, K, CP)) return false;
    T = tv;
    return true;
}

// Stage the search tables of t in LDS (all threads of the block; ends with a
// barrier).  Layout: xh[n] xt[n] sh[nbh] st[nbt], props_lds_bytes() bytes.
__device__ __forceinline__ void stage_tables(const PropTab& t, unsigned char* smem, const double*& xh,
                                             const double*& xt, const int*& sh, const int*& st) {
    double* a = reinterpret_cast<double*>(smem);
    double* b = a + t.n;
    int* c = reinterpret_cast<int*>(b + t.n);
    int* d = c + t.nbh;
    for (int e = threadIdx.x; e < t.n; e += blockDim.x) {
        a[e] = t.xh[e];
        b[e] = t.xt[e];
    }
    for (int e = threadIdx.x; e < t.nbh; e += blockDim.x) c[e] = t.sh[e];
    for (int e = threadIdx.x; e < t.nbt; e += blockDim.x) d[e] = t.st[e];
    __syncthreads();
    xh = a;
    xt = b;
    sh = c;
    st = d;
}

}  // namespace chs
