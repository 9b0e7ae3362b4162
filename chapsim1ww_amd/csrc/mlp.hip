// Learned property surrogate for THERM_PROP_UPDATE: a 1 -> 64 -> 64 -> 5 tanh
// network h -> (T, rho, mu, kappa, c_p) with BF16 weights and FP32
// accumulation on the gfx950 matrix cores, trained on the device by the
// kernels below (forward on MFMA, backward in FP32, Adam), certified against
// the table per enthalpy bin, and used by THERM_PROP_UPDATE where certified
// (the table lookup elsewhere).
//
// Network (float32 master parameters, flat, NPAR = 4613):
//   w1[64] b1[64] W2[64][64] b2[64] W3[5][64] b3[5]        W[out][in]
// Every weight is stored as a pair of bfloat16 values, q(w) = hi(w) + lo(w)
// with hi(v) = bf(v) and lo(v) = bf(v - hi(v)) (bf = round to bfloat16,
// nearest even); activations enter the matrix cores the same way:
//   x = (h - hc) * hs;  a1 = tanh(q(w1) x + b1)
//   a2 = tanh(hi(W2) (hi(a1) + lo(a1)) + lo(W2) hi(a1) + b2)
//   o  = hi(W3) (hi(a2) + lo(a2)) + lo(W3) hi(a2) + b3;   y = m + s o
// -- three BF16 products with FP32 accumulation per weight block, keeping
// ~16 significant bits of weights and activations (a single BF16 weight
// limits the fit to ~1e-3 relative, above the 1e-4 the fallback demands).
//
// Matrix-core tiling (v_mfma_f32_16x16x32_bf16, wave64): a tile is 16 samples.
// The result of a 16-unit x 16-sample product sits in lane l = 16 g + i as
// units 4g + r (r = 0..3) of sample i; that is exactly the operand layout the
// next layer needs once the 32 reduction slots of a lane group are taken as
// units {4g..4g+3} u {16+4g..16+4g+3} of a 32-unit block -- the weights are
// loaded in that slot order, so layers chain in registers.
#include <algorithm>
#include <cmath>
#include <vector>

#include "props.hpp"

namespace chs {

constexpr int H1 = 64, H2 = 64, NOUT = 5;
constexpr int O_W1 = 0, O_B1 = 64, O_W2 = 128, O_B2 = O_W2 + H2 * H1, O_W3 = O_B2 + H2, O_B3 = O_W3 + NOUT * H2;
constexpr int NPAR = O_B3 + NOUT;   // 4613
constexpr int TB = 64;              // samples per training block (one tile per wave)
constexpr int NPT = 33;             // certification points per bin (edges included)

typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef float f4v __attribute__((ext_vector_type(4)));


// tanh(z) = sign(z) (1 - 2 / (exp(2|z|) + 1)); absolute error ~1e-7
__device__ inline float tanh_f(float z) {
    const float e = __expf(2.0f * fabsf(z));
    const float t = 1.0f - 2.0f / (e + 1.0f);
    return copysignf(t, z);
}

// unit of reduction slot s (0..7) of lane group g within a 32-unit block
__host__ __device__ constexpr int slot_unit(int g, int s) { return s < 4 ? 4 * g + s : 16 + 4 * g + (s - 4); }

struct MlpNorm {
    double hc, hs;        // x = (h - hc) * hs
    double h_lo, h_hi;    // trained range
    float m[NOUT], s[NOUT];
};

// The network's matrix operands in LDS, each lane's 8 values in its slot
// order: op 0..7 = W2 hi [unit tile ut][block kb] at 2 ut + kb, 8..15 = W2 lo,
// 16..17 = W3 hi [kb], 18..19 = W3 lo; plus layer 1 and the biases.
constexpr int NOPS = 20;
struct WLds {
    bf8v op[NOPS][64];
    float w1[H1], b1[H1], b2[H2], b3[8];
};

// wq holds every weight as hi + lo (two bfloat16 values, exact in float):
// split it back into its matrix operands
__device__ inline void stage_weights(WLds& W, const float* __restrict__ wq) {
    for (int e = threadIdx.x; e < NOPS * 64; e += blockDim.x) {
        const int op = e >> 6, lane = e & 63, i = lane & 15, g = lane >> 4;
        const bool lo = (op >= 8 && op < 16) || op >= 18;
        bf8v v;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            float w;
            if (op < 16) {
                const int q = op & 7, ut = q >> 1, kb = q & 1;
                w = wq[O_W2 + (16 * ut + i) * H1 + 32 * kb + slot_unit(g, s)];
            } else {
                const int kb = (op - 16) & 1;
                w = i < NOUT ? wq[O_W3 + i * H2 + 32 * kb + slot_unit(g, s)] : 0.0f;
            }
            const __bf16 h = (__bf16)w;
            v[s] = lo ? (__bf16)(w - (float)h) : h;
        }
        W.op[op][lane] = v;
    }
    for (int e = threadIdx.x; e < H1; e += blockDim.x) {
        W.w1[e] = wq[O_W1 + e];
        W.b1[e] = wq[O_B1 + e];
        W.b2[e] = wq[O_B2 + e];
    }
    if (threadIdx.x < 8) W.b3[threadIdx.x] = threadIdx.x < NOUT ? wq[O_B3 + threadIdx.x] : 0.0f;
    __syncthreads();
}

// hi/lo bfloat16 operands of the 32-unit block kb of a (16 values per lane)
__device__ inline void split_block(const float* a, int kb, bf8v& hi, bf8v& lo) {
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const float v = a[8 * kb + s];
        const __bf16 h = (__bf16)v;
        hi[s] = h;
        lo[s] = (__bf16)(v - (float)h);
    }
}

// W a with W = Wh + Wl and a = ah + al: Wh ah + Wh al + Wl ah on the matrix cores
__device__ inline f4v mm3(const bf8v& wh, const bf8v& wl, const bf8v& ah, const bf8v& al, f4v acc) {
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wh, ah, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wh, al, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wl, ah, acc, 0, 0, 0);
    return acc;
}

// Forward of one 16-sample tile: lane (g, i) holds sample i's x; returns
// o for outputs 4g + r, and a1 / a2 for units 16t + 4g + r (index 4t + r).
__device__ inline f4v forward_tile(const WLds& W, int lane, float x, float* a1, float* a2) {
    const int g = lane >> 4;
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int u = 16 * t + 4 * g + r;
            a1[4 * t + r] = tanh_f(fmaf(W.w1[u], x, W.b1[u]));
        }
    bf8v h0, l0, h1, l1;
    split_block(a1, 0, h0, l0);
    split_block(a1, 1, h1, l1);
#pragma unroll
    for (int ut = 0; ut < 4; ++ut) {
        f4v acc;
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[r] = W.b2[16 * ut + 4 * g + r];
        acc = mm3(W.op[2 * ut][lane], W.op[8 + 2 * ut][lane], h0, l0, acc);
        acc = mm3(W.op[2 * ut + 1][lane], W.op[8 + 2 * ut + 1][lane], h1, l1, acc);
#pragma unroll
        for (int r = 0; r < 4; ++r) a2[4 * ut + r] = tanh_f(acc[r]);
    }
    split_block(a2, 0, h0, l0);
    split_block(a2, 1, h1, l1);
    f4v o;
#pragma unroll
    for (int r = 0; r < 4; ++r) o[r] = W.b3[(4 * g + r) & 7];
    o = mm3(W.op[16][lane], W.op[18][lane], h0, l0, o);
    o = mm3(W.op[17][lane], W.op[19][lane], h1, l1, o);
    return o;
}

// weight as hi + lo bfloat16 pair (exact sum in float)
__device__ inline float wsplit(float w) {
    const __bf16 h = (__bf16)w;
    return (float)h + (float)(__bf16)(w - (float)h);
}

struct Mlp {
    float* par = nullptr;     // FP32 master parameters [NPAR]
    float* wq = nullptr;      // weights as q(w) = hi + lo, FP32 biases [NPAR] (what the forward uses)
    float* am = nullptr;      // Adam moments [NPAR] x 2
    float* av = nullptr;
    float* partial = nullptr; // per-block gradient partials [maxblk][NPAR + 1] (last: loss)
    int maxblk = 0;
    MlpNorm norm{};
    bool trained = false;
    long long step = 0;
    int nbins = 0;
    unsigned char* cert = nullptr;   // [nbins] bit 0: all five outputs within the bound; bit 1: T, rho, mu, kappa
    double cert_frac = 0.0;          // fraction of bins with bit 0
    double cert_frac4 = 0.0;         // fraction of bins with bit 1
};

// ---------------------------------------------------------------------------
// training: forward of TB samples per block on the matrix cores, backward in
// FP32 through LDS, block partial gradients; then Adam over the partials.
// ---------------------------------------------------------------------------
struct TrainLds {
    float a1[H1][TB];
    float a2[H2][TB];
    float o[NOUT][TB];     // outputs, then their gradients
    float t[NOUT][TB];
    float x[TB];
    float red[256];
};

__device__ inline unsigned long long mix64(unsigned long long z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// samples: sel == nullptr -> sample b*TB + s of (px, pt) directly (count n);
// otherwise the pool index is drawn from (seed, step, position)
__global__ void __launch_bounds__(256) k_mlp_train(const float* __restrict__ wq, MlpNorm nm,
                                                   const float* __restrict__ px, const float* __restrict__ pt,
                                                   long long npool, long long nsel, unsigned long long seed,
                                                   long long step, int random_draw, double inv5B,
                                                   float* __restrict__ partial) {
    extern __shared__ unsigned char smem[];
    TrainLds& S = *reinterpret_cast<TrainLds*>(smem);
    WLds& WL = *reinterpret_cast<WLds*>(smem + sizeof(TrainLds));
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    // stage samples and weights
    for (int s = tid; s < TB; s += 256) {
        const long long pos = (long long)blockIdx.x * TB + s;
        long long q = -1;
        if (pos < nsel) q = random_draw ? (long long)(mix64(seed ^ mix64((unsigned long long)step * 0x100000001b3ull + pos)) % (unsigned long long)npool) : pos;
        S.x[s] = q >= 0 ? px[q] : 0.0f;
        for (int k = 0; k < NOUT; ++k) S.t[k][s] = q >= 0 ? pt[q * NOUT + k] : 0.0f;   // 0 marks padding
    }
    stage_weights(WL, wq);   // ends with a barrier
    // forward: each wave one tile of 16 samples
    {
        const int i = lane & 15, g = lane >> 4;
        for (int tt = 0; tt < TB / 64; ++tt) {
            const int s = 16 * (wv * (TB / 64) + tt) + i;
            float a1[16], a2[16];
            const f4v o = forward_tile(WL, lane, S.x[s], a1, a2);
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    S.a1[16 * t + 4 * g + r][s] = a1[4 * t + r];
                    S.a2[16 * t + 4 * g + r][s] = a2[4 * t + r];
                }
#pragma unroll
            for (int r = 0; r < 4; ++r)
                if (4 * g + r < NOUT) S.o[4 * g + r][s] = o[r];
        }
    }
    __syncthreads();
    // loss and dL/do
    float lsum = 0.0f;
    if (tid < TB) {
        const int s = tid;
#pragma unroll
        for (int k = 0; k < NOUT; ++k) {
            const float t = S.t[k][s];
            float d = 0.0f;
            if (t != 0.0f) {
                const float y = nm.m[k] + nm.s[k] * S.o[k][s];
                const float e = (y - t) / t;
                lsum += e * e;
                d = (float)(2.0 * inv5B) * e / t * nm.s[k];
            }
            S.o[k][s] = d;
        }
    }
    S.red[tid] = lsum;
    __syncthreads();
    float* P = partial + (size_t)blockIdx.x * (NPAR + 1);
    if (tid == 0) {
        double l = 0.0;
        for (int q = 0; q < 256; ++q) l += S.red[q];
        P[NPAR] = (float)(l * inv5B);
    }
    // dW3, db3
    for (int e = tid; e < NOUT * H2; e += 256) {
        const int k = e / H2, u = e % H2;
        float acc = 0.0f;
        for (int s = 0; s < TB; ++s) acc = fmaf(S.o[k][s], S.a2[u][s], acc);
        P[O_W3 + e] = acc;
    }
    if (tid < NOUT) {
        float acc = 0.0f;
        for (int s = 0; s < TB; ++s) acc += S.o[tid][s];
        P[O_B3 + tid] = acc;
    }
    __syncthreads();
    // dz2 = (W3^T do) * (1 - a2^2), in place of a2
    for (int e = tid; e < H2 * TB; e += 256) {
        const int u = e / TB, s = e % TB;
        float acc = 0.0f;
#pragma unroll
        for (int k = 0; k < NOUT; ++k) acc = fmaf(wq[O_W3 + k * H2 + u], S.o[k][s], acc);
        const float a = S.a2[u][s];
        S.a2[u][s] = acc * (1.0f - a * a);
    }
    __syncthreads();
    // dW2 (4 x 4 register blocks), db2
    {
        const int u0 = 4 * (tid >> 4), v0 = 4 * (tid & 15);
        float acc[4][4] = {};
        for (int s = 0; s < TB; ++s) {
            float dz[4], av[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                dz[q] = S.a2[u0 + q][s];
                av[q] = S.a1[v0 + q][s];
            }
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = fmaf(dz[a], av[b], acc[a][b]);
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) P[O_W2 + (u0 + a) * H1 + v0 + b] = acc[a][b];
        if (tid < H2) {
            float sacc = 0.0f;
            for (int s = 0; s < TB; ++s) sacc += S.a2[tid][s];
            P[O_B2 + tid] = sacc;
        }
    }
    __syncthreads();
    // dz1 = (W2^T dz2) * (1 - a1^2), in place of a1
    for (int e = tid; e < H1 * TB; e += 256) {
        const int v = e % H1, s = e / H1;
        float acc = 0.0f;
        for (int u = 0; u < H2; ++u) acc = fmaf(wq[O_W2 + u * H1 + v], S.a2[u][s], acc);
        const float a = S.a1[v][s];
        S.a1[v][s] = acc * (1.0f - a * a);
    }
    __syncthreads();
    if (tid < H1) {
        float gw = 0.0f, gb = 0.0f;
        for (int s = 0; s < TB; ++s) {
            gw = fmaf(S.a1[tid][s], S.x[s], gw);
            gb += S.a1[tid][s];
        }
        P[O_W1 + tid] = gw;
        P[O_B1 + tid] = gb;
    }
}

// sum of the block partials -> gradient (and loss); Adam step when lr > 0;
// wq <- q(w) for weights, FP32 biases.  Block = 32 parameters x 8 threads:
// thread s of a parameter sums partials s, s+8, ..., then the 8 sums are
// added in a fixed order (deterministic).
constexpr int AD_P = 32, AD_S = 8;
__global__ void __launch_bounds__(AD_P * AD_S) k_mlp_adam(const float* __restrict__ partial, int nblk,
                                                         float* __restrict__ par, float* __restrict__ am,
                                                         float* __restrict__ av, float* __restrict__ wq,
                                                         float* __restrict__ grad_out, float lr, float b1t,
                                                         float b2t) {
    __shared__ double red[AD_S][AD_P];
    const int pl = threadIdx.x % AD_P, sl = threadIdx.x / AD_P;
    const int p = blockIdx.x * AD_P + pl;
    double gs = 0.0;
    if (p <= NPAR)
        for (int b = sl; b < nblk; b += AD_S) gs += partial[(size_t)b * (NPAR + 1) + p];
    red[sl][pl] = gs;
    __syncthreads();
    if (sl != 0 || p > NPAR) return;
    double g = 0.0;
#pragma unroll
    for (int q = 0; q < AD_S; ++q) g += red[q][pl];
    if (grad_out) grad_out[p] = (float)g;
    if (p == NPAR || lr <= 0.0f) return;
    const float gf = (float)g;
    const float m = 0.9f * am[p] + 0.1f * gf;
    const float v = 0.999f * av[p] + 0.001f * gf * gf;
    am[p] = m;
    av[p] = v;
    const float mh = m / (1.0f - b1t), vh = v / (1.0f - b2t);
    const float w = par[p] - lr * mh / (sqrtf(vh) + 1e-8f);
    par[p] = w;
    const bool bias = (p >= O_B1 && p < O_W2) || (p >= O_B2 && p < O_W3) || p >= O_B3;
    wq[p] = bias ? w : wsplit(w);
}

__global__ void k_mlp_round(const float* __restrict__ par, float* __restrict__ wq) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPAR) return;
    const bool bias = (p >= O_B1 && p < O_W2) || (p >= O_B2 && p < O_W3) || p >= O_B3;
    wq[p] = bias ? par[p] : wsplit(par[p]);
}

// training pool: x = (h - hc) hs (float), targets from the table (float)
__global__ void k_mlp_pool(const double* __restrict__ h, long long n, MlpNorm nm, const double* __restrict__ T,
                           const double* __restrict__ D, const double* __restrict__ M, const double* __restrict__ K,
                           const double* __restrict__ CP, float* __restrict__ px, float* __restrict__ pt) {
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x) {
        px[q] = (float)((h[q] - nm.hc) * nm.hs);
        pt[q * NOUT + 0] = (float)T[q];
        pt[q * NOUT + 1] = (float)D[q];
        pt[q * NOUT + 2] = (float)M[q];
        pt[q * NOUT + 3] = (float)K[q];
        pt[q * NOUT + 4] = (float)CP[q];
    }
}

// ---------------------------------------------------------------------------
// evaluation: raw network outputs y (FP64 out[5][n]) of n values of h
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_mlp_eval(const float* __restrict__ wq, MlpNorm nm,
                                                  const double* __restrict__ h, long long n,
                                                  double* __restrict__ out) {
    extern __shared__ unsigned char smem[];
    WLds& W = *reinterpret_cast<WLds*>(smem);
    stage_weights(W, wq);
    const int lane = threadIdx.x & 63, i = lane & 15, g = lane >> 4;
    const long long ntile = (n + 15) / 16;
    const long long wid = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nw = ((long long)gridDim.x * blockDim.x) >> 6;
    for (long long tl = wid; tl < ntile; tl += nw) {
        const long long q = tl * 16 + i;
        const double hv = q < n ? h[q] : nm.hc;
        const float x = (float)((hv - nm.hc) * nm.hs);
        float a1[16], a2[16];
        const f4v o = forward_tile(W, lane, x, a1, a2);
        if (q < n) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int k = 4 * g + r;
                if (k < NOUT) out[(size_t)k * n + q] = (double)(nm.m[k] + nm.s[k] * o[r]);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// THERM_PROP_UPDATE with the surrogate: certified bins use the network, all
// other cells (outside the trained range, uncertified bins, NaN) the table.
// ---------------------------------------------------------------------------
constexpr int ST = 256;     // threads per surrogate block (4 waves, 64 cells per pass)
constexpr int SPASS = 8;    // passes per chunk

__global__ void __launch_bounds__(ST) k_mlp_props(PropTab t, const float* __restrict__ wq, MlpNorm nm,
                                                  const unsigned char* __restrict__ cert, int nbins, Rows rw,
                                                  const double* __restrict__ h, double* __restrict__ To,
                                                  double* __restrict__ rho, double* __restrict__ mu,
                                                  double* __restrict__ kap, double* __restrict__ cp,
                                                  unsigned long long* __restrict__ cnt) {
    extern __shared__ unsigned char smem[];
    WLds& W = *reinterpret_cast<WLds*>(smem);
    unsigned char* sc = smem + sizeof(WLds);
    for (int e = threadIdx.x; e < nbins; e += blockDim.x) sc[e] = cert[e];
    // bin flags: bit 0 all five outputs certified, bit 1 T, rho, mu, kappa
    // (what THERM_PROP_UPDATE writes when c_p is not asked for)
    const unsigned char need = cp ? 1 : 2;
    stage_weights(W, wq);   // ends with a barrier
    // the table fallback searches the knots in global memory (L2-resident)
    const double *xh = t.xh, *xt = t.xt;
    const int *sh = t.sh, *st = t.st;
    const int lane = threadIdx.x & 63, i = lane & 15, g = lane >> 4, wv = threadIdx.x >> 6;
    const double binv = (double)nbins / (nm.h_hi - nm.h_lo);
    const long long chunk = (long long)ST / 4 * SPASS;        // cells per chunk
    const long long cpr = (rw.rowlen + chunk - 1) / chunk;
    const long long nch = cpr * rw.nrow;
    unsigned long long bad = 0, fell = 0;
    for (long long ch = blockIdx.x; ch < nch; ch += gridDim.x) {
        const long long r = ch / cpr;
        const long long c0 = (ch - r * cpr) * chunk;
        const long long base = r * rw.stride;
        for (int ps = 0; ps < SPASS; ++ps) {
            const long long c = c0 + (long long)(ps * 4 + wv) * 16 + i;
            const bool in = c < rw.rowlen;
            const long long q = base + (in ? c : 0);
            const double hv = in ? h[q] : nm.hc;
            // the network where its bin is certified
            bool use_net = hv > nm.h_lo && hv < nm.h_hi;
            if (use_net) {
                int b = (int)((hv - nm.h_lo) * binv);
                b = b < 0 ? 0 : (b >= nbins ? nbins - 1 : b);
                use_net = (sc[b] & need) != 0;
            }
            const float x = (float)((use_net ? hv - nm.hc : 0.0) * nm.hs);
            f4v o = {0.0f, 0.0f, 0.0f, 0.0f};
            if (__any(use_net)) {   // wave-uniform: the matrix ops run with every lane active
                float a1[16], a2[16];
                o = forward_tile(W, lane, x, a1, a2);
            }
            if (!in || fabs(hv) < 1.0e-10) continue;
            if (use_net) {
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) {
                    const int k = 4 * g + rr;
                    const double y = (double)(nm.m[k < NOUT ? k : 0] + nm.s[k < NOUT ? k : 0] * o[rr]);
                    if (k == 0) To[q] = y;
                    else if (k == 1) rho[q] = y;
                    else if (k == 2) mu[q] = y;
                    else if (k == 3) kap[q] = y;
                    else if (k == 4 && cp) cp[q] = y;
                }
            } else if (g == 0) {
                ++fell;
                double T, D, Mu, K, CP;
                if (!lookup_cell(t, xh, xt, sh, st, hv, T, D, Mu, K, CP)) {
                    ++bad;
                    continue;
                }
                To[q] = T;
                mu[q] = Mu;
                kap[q] = K;
                rho[q] = D;
                if (cp) cp[q] = CP;
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        bad += __shfl_xor(bad, o);
        fell += __shfl_xor(fell, o);
    }
    if (lane == 0) {
        if (bad) atomicAdd(cnt, bad);
        if (fell) atomicAdd(cnt + 1, fell);
    }
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------
int props_table_launch(chs_props* p, const double* h, Rows rw, double* T, double* rho, double* mu, double* kap,
                       double* cp, unsigned long long* cnt, int count_lookups, hipStream_t stream);

void mlp_release(chs_props* p) {
    if (!p || !p->mlp) return;
    Mlp* m = p->mlp;
    if (m->par) (void)hipFree(m->par);
    if (m->partial) (void)hipFree(m->partial);
    if (m->cert) (void)hipFree(m->cert);
    delete m;
    p->mlp = nullptr;
}

// Glorot-uniform initial weights, zero biases, from a SplitMix64 stream
static void init_params(std::vector<float>& par, unsigned long long seed) {
    par.assign(NPAR, 0.0f);
    unsigned long long z = seed;
    auto uni = [&]() {
        z += 0x9e3779b97f4a7c15ull;
        unsigned long long v = z;
        v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ull;
        v = (v ^ (v >> 27)) * 0x94d049bb133111ebull;
        v ^= v >> 31;
        return (double)(v >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
    };
    const double a1 = std::sqrt(6.0 / (1 + H1)), a2 = std::sqrt(6.0 / (H1 + H2)), a3 = std::sqrt(6.0 / (H2 + NOUT));
    for (int u = 0; u < H1; ++u) par[O_W1 + u] = (float)(a1 * uni());
    for (int e = 0; e < H2 * H1; ++e) par[O_W2 + e] = (float)(a2 * uni());
    for (int e = 0; e < NOUT * H2; ++e) par[O_W3 + e] = (float)(a3 * uni());
}

static int mlp_alloc(chs_props* p, int maxblk) {
    if (p->mlp && p->mlp->maxblk >= maxblk) return CHS_OK;
    Mlp* old = p->mlp;
    Mlp* m = new Mlp();
    if (old) {
        m->norm = old->norm;
        m->trained = old->trained;
        m->step = old->step;
    }
    m->maxblk = maxblk;
    // one allocation: par | wq | am | av
    if (hipMalloc((void**)&m->par, sizeof(float) * 4 * (size_t)(NPAR + 3)) != hipSuccess ||
        hipMalloc((void**)&m->partial, sizeof(float) * (size_t)maxblk * (NPAR + 1)) != hipSuccess) {
        if (m->par) (void)hipFree(m->par);
        delete m;
        set_error("surrogate: device allocation failed");
        return CHS_E_NOMEM;
    }
    m->wq = m->par + (NPAR + 3);
    m->am = m->wq + (NPAR + 3);
    m->av = m->am + (NPAR + 3);
    if (old) {
        CHS_HIP(hipMemcpy(m->par, old->par, sizeof(float) * 4 * (size_t)(NPAR + 3), hipMemcpyDeviceToDevice));
        if (old->cert) {
            m->cert = old->cert;
            m->nbins = old->nbins;
            m->cert_frac = old->cert_frac;
            m->cert_frac4 = old->cert_frac4;
            old->cert = nullptr;
        }
        mlp_release(p);
    }
    p->mlp = m;
    return CHS_OK;
}

static int set_params(chs_props* p, const float* host) {
    CHS_TRY(mlp_alloc(p, 64));
    Mlp* m = p->mlp;
    std::vector<float> z(2 * (size_t)(NPAR + 3), 0.0f);
    CHS_HIP(hipMemcpy(m->par, host, sizeof(float) * NPAR, hipMemcpyHostToDevice));
    CHS_HIP(hipMemcpy(m->am, z.data(), sizeof(float) * z.size(), hipMemcpyHostToDevice));
    CHS_LAUNCH(k_mlp_round, dim3((NPAR + 255) / 256), dim3(256), 0, p->stream, m->par, m->wq);
    m->step = 0;
    return CHS_OK;
}

static size_t mlp_props_lds(const chs_props* p) { return sizeof(WLds) + (size_t)(p->mlp ? p->mlp->nbins : 0); }

bool mlp_ready(const chs_props* p) { return p->mlp && p->mlp->trained && p->mlp->cert; }

int mlp_update(chs_props* p, const double* h, Rows rw, double* T, double* rho, double* mu, double* kap, double* cp,
               unsigned long long* cnt, hipStream_t stream) {
    CHS_CHECK(p->mlp && p->mlp->trained && p->mlp->cert, CHS_E_ARG,
              "chs_therm_prop_update: the surrogate is not trained and certified");
    const Mlp* m = p->mlp;
    // no certified bin: every cell takes the table, so run the table kernel
    if ((cp ? m->cert_frac : m->cert_frac4) == 0.0) return props_table_launch(p, h, rw, T, rho, mu, kap, cp, cnt, 1, stream);
    const long long chunk = (long long)ST / 4 * SPASS;
    const long long nch = ((rw.rowlen + chunk - 1) / chunk) * rw.nrow;
    if (nch == 0) return CHS_OK;
    const int grid = (int)std::min<long long>(nch, 2048);
    CHS_LAUNCH(k_mlp_props, dim3(grid), dim3(ST), mlp_props_lds(p), stream, tab_of(p), m->wq, m->norm, m->cert,
               m->nbins, rw, h, T, rho, mu, kap, cp, cnt);
    return CHS_OK;
}

__global__ void k_zero_u64_b(unsigned long long* c, int n) {
    if ((int)threadIdx.x < n) c[threadIdx.x] = 0ull;
}

// pool of training samples: x and table labels (float) on the device
struct Pool {
    float* px = nullptr;
    float* pt = nullptr;
    ~Pool() {
        if (px) (void)hipFree(px);
    }
};

static int make_pool(chs_props* p, const double* h, long long n, Pool& pool) {
    const MlpNorm& nm = p->mlp->norm;
    double* lab = nullptr;
    CHS_HIP(hipMalloc((void**)&lab, sizeof(double) * 5 * (size_t)n));
    float* buf = nullptr;
    if (hipMalloc((void**)&buf, sizeof(float) * 6 * (size_t)n) != hipSuccess) {
        (void)hipFree(lab);
        set_error("surrogate: device allocation failed");
        return CHS_E_NOMEM;
    }
    pool.px = buf;
    pool.pt = buf + n;
    Rows rw{1, n, n};
    hipLaunchKernelGGL(k_zero_u64_b, dim3(1), dim3(64), 0, p->stream, p->counters, 4);
    int st = props_table_launch(p, h, rw, lab, lab + n, lab + 2 * n, lab + 3 * n, lab + 4 * n, p->counters, 0,
                                p->stream);
    unsigned long long bad = 0;
    if (st == CHS_OK) {
        if (hipMemcpyAsync(&bad, p->counters, sizeof(bad), hipMemcpyDeviceToHost, p->stream) != hipSuccess ||
            hipStreamSynchronize(p->stream) != hipSuccess)
            st = CHS_E_HIP;
    }
    if (st == CHS_OK && bad) {
        set_error("surrogate: training samples outside the property table");
        st = CHS_E_RANGE;
    }
    if (st == CHS_OK) {
        hipLaunchKernelGGL(k_mlp_pool, dim3(1024), dim3(256), 0, p->stream, h, n, nm, lab, lab + n, lab + 2 * n,
                           lab + 3 * n, lab + 4 * n, pool.px, pool.pt);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(p->stream) != hipSuccess) st = CHS_E_HIP;
    }
    (void)hipFree(lab);
    return st;
}

static int train(chs_props* p, const double* h, long long nsamp, int steps, int batch, double lr0,
                 unsigned long long seed, double* final_loss) {
    CHS_CHECK(h && nsamp >= 16, CHS_E_ARG, "chs_surrogate_train: need at least 16 samples");
    CHS_CHECK(steps >= 0 && batch >= TB && batch % TB == 0 && batch <= 65536, CHS_E_ARG,
              "chs_surrogate_train: batch must be a multiple of 64 up to 65536");
    const int nblk = batch / TB;
    CHS_TRY(mlp_alloc(p, std::max(nblk, 64)));
    Mlp* m = p->mlp;
    // normalisation from the samples (host pass over a copy of h)
    std::vector<double> hh(nsamp);
    CHS_HIP(hipMemcpy(hh.data(), h, sizeof(double) * nsamp, hipMemcpyDeviceToHost));
    double lo = hh[0], hi = hh[0];
    for (double v : hh) {
        CHS_CHECK(std::isfinite(v), CHS_E_ARG, "chs_surrogate_train: non-finite sample");
        lo = std::min(lo, v);
        hi = std::max(hi, v);
    }
    CHS_CHECK(hi > lo, CHS_E_ARG, "chs_surrogate_train: samples span no range");
    MlpNorm nm{};
    nm.hc = 0.5 * (lo + hi);
    nm.hs = 2.0 / (hi - lo);
    nm.h_lo = lo;
    nm.h_hi = hi;
    for (int k = 0; k < NOUT; ++k) {
        nm.m[k] = 0.0f;
        nm.s[k] = 1.0f;
    }
    m->norm = nm;
    Pool pool;
    CHS_TRY(make_pool(p, h, nsamp, pool));
    // output scaling: mean and spread of each label over the pool
    {
        std::vector<float> t(5 * (size_t)nsamp);
        CHS_HIP(hipMemcpy(t.data(), pool.pt, sizeof(float) * t.size(), hipMemcpyDeviceToHost));
        for (int k = 0; k < NOUT; ++k) {
            double s1 = 0, s2 = 0;
            for (long long q = 0; q < nsamp; ++q) {
                s1 += t[q * 5 + k];
                s2 += (double)t[q * 5 + k] * t[q * 5 + k];
            }
            const double mu = s1 / nsamp, var = std::max(s2 / nsamp - mu * mu, 0.0);
            m->norm.m[k] = (float)mu;
            m->norm.s[k] = (float)std::max(std::sqrt(var), 1e-3 * std::fabs(mu) + 1e-6);
        }
    }
    if (!m->trained) {
        std::vector<float> par;
        init_params(par, seed);
        CHS_TRY(set_params(p, par.data()));
    }
    m = p->mlp;
    const size_t lds = sizeof(TrainLds) + sizeof(WLds);
    double b1t = 1.0, b2t = 1.0;
    for (long long s0 = 0; s0 < (long long)steps; ++s0) {
        // cosine decay from lr0 to lr0/100
        const double f = steps > 1 ? (double)s0 / (steps - 1) : 0.0;
        const double lr = lr0 * (0.01 + 0.99 * 0.5 * (1.0 + std::cos(3.14159265358979323846 * f)));
        ++m->step;
        b1t = std::pow(0.9, (double)m->step);
        b2t = std::pow(0.999, (double)m->step);
        CHS_LAUNCH(k_mlp_train, dim3(nblk), dim3(256), lds, p->stream, m->wq, m->norm, pool.px, pool.pt, nsamp,
                   (long long)batch, seed, m->step, 1, 1.0 / (5.0 * batch), m->partial);
        CHS_LAUNCH(k_mlp_adam, dim3((NPAR + 1 + AD_P - 1) / AD_P), dim3(AD_P * AD_S), 0, p->stream, m->partial, nblk, m->par, m->am,
                   m->av, m->wq, (float*)nullptr, (float)lr, (float)b1t, (float)b2t);
    }
    // loss over the whole pool with the final weights
    double loss = 0.0;
    {
        const long long nb = (nsamp + TB - 1) / TB;
        for (long long b0 = 0; b0 < nb; b0 += m->maxblk) {
            const int nbb = (int)std::min<long long>(m->maxblk, nb - b0);
            const long long off = b0 * TB;
            const long long cnt = std::min<long long>((long long)nbb * TB, nsamp - off);
            CHS_LAUNCH(k_mlp_train, dim3(nbb), dim3(256), lds, p->stream, m->wq, m->norm, pool.px + off,
                       pool.pt + off * NOUT, cnt, cnt, 0ull, 0ll, 0, 1.0 / (5.0 * nsamp), m->partial);
            std::vector<float> part((size_t)nbb * (NPAR + 1));
            CHS_HIP(hipMemcpyAsync(part.data(), m->partial, sizeof(float) * part.size(), hipMemcpyDeviceToHost,
                                   p->stream));
            CHS_HIP(hipStreamSynchronize(p->stream));
            for (int b = 0; b < nbb; ++b) loss += part[(size_t)b * (NPAR + 1) + NPAR];
        }
    }
    m->trained = true;
    if (m->cert) {   // a new network needs a new certificate
        (void)hipFree(m->cert);
        m->cert = nullptr;
        m->cert_frac = 0.0;
        m->cert_frac4 = 0.0;
    }
    if (final_loss) *final_loss = loss;
    return CHS_OK;
}

// gradient of the loss over exactly the n given samples (no update)
static int grad(chs_props* p, const double* h, long long n, float* grad_host, double* loss) {
    CHS_CHECK(p->mlp && p->mlp->trained, CHS_E_ARG, "chs_surrogate_grad: no network (train or set parameters first)");
    CHS_CHECK(h && n > 0 && n <= 128LL * TB, CHS_E_ARG, "chs_surrogate_grad: 1 <= n <= 8192 samples");
    Mlp* m = p->mlp;
    const int nblk = (int)((n + TB - 1) / TB);
    CHS_TRY(mlp_alloc(p, nblk));
    m = p->mlp;
    Pool pool;
    CHS_TRY(make_pool(p, h, n, pool));
    float* g = nullptr;
    CHS_HIP(hipMalloc((void**)&g, sizeof(float) * (NPAR + 1)));
    hipLaunchKernelGGL(k_mlp_train, dim3(nblk), dim3(256), sizeof(TrainLds) + sizeof(WLds), p->stream, m->wq, m->norm, pool.px,
                       pool.pt, n, n, 0ull, 0ll, 0, 1.0 / (5.0 * n), m->partial);
    hipLaunchKernelGGL(k_mlp_adam, dim3((NPAR + 1 + AD_P - 1) / AD_P), dim3(AD_P * AD_S), 0, p->stream, m->partial, nblk, m->par,
                       m->am, m->av, m->wq, g, 0.0f, 0.0f, 0.0f);
    std::vector<float> gh(NPAR + 1);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(gh.data(), g, sizeof(float) * (NPAR + 1), hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    (void)hipFree(g);
    CHS_HIP(e);
    if (grad_host) std::copy(gh.begin(), gh.begin() + NPAR, grad_host);
    if (loss) *loss = gh[NPAR];
    return CHS_OK;
}

// per-bin certification: the network against the table on NPT points per bin
__global__ void __launch_bounds__(256) k_mlp_cert_points(MlpNorm nm, int nbins, double* __restrict__ hp) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)nbins * NPT) return;
    const int b = (int)(q / NPT), k = (int)(q % NPT);
    const double w = (nm.h_hi - nm.h_lo) / nbins;
    double v = nm.h_lo + w * (b + (double)k / (NPT - 1));
    // inside the open trained range
    if (v <= nm.h_lo) v = nm.h_lo + 1e-12 * (nm.h_hi - nm.h_lo);
    if (v >= nm.h_hi) v = nm.h_hi - 1e-12 * (nm.h_hi - nm.h_lo);
    hp[q] = v;
}

static int certify(chs_props* p, double tol, int nbins, double* frac) {
    CHS_CHECK(p->mlp && p->mlp->trained, CHS_E_ARG, "chs_surrogate_certify: no network");
    CHS_CHECK(nbins >= 1 && nbins <= 16384 && tol > 0.0, CHS_E_ARG, "chs_surrogate_certify: 1 <= nbins <= 16384, tol > 0");
    Mlp* m = p->mlp;
    const long long np = (long long)nbins * NPT;
    double* buf = nullptr;
    CHS_HIP(hipMalloc((void**)&buf, sizeof(double) * 11 * (size_t)np));
    double* hp = buf;
    double* tab = buf + np;          // 5 x np
    double* net = buf + 6 * np;      // 5 x np
    int st = CHS_OK;
    hipLaunchKernelGGL(k_mlp_cert_points, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, p->stream, m->norm, nbins, hp);
    hipLaunchKernelGGL(k_zero_u64_b, dim3(1), dim3(64), 0, p->stream, p->counters, 4);
    Rows rw{1, np, np};
    st = props_table_launch(p, hp, rw, tab, tab + np, tab + 2 * np, tab + 3 * np, tab + 4 * np, p->counters, 0,
                            p->stream);
    if (st == CHS_OK) {
        hipLaunchKernelGGL(k_mlp_eval, dim3(1024), dim3(256), sizeof(WLds), p->stream, m->wq, m->norm, hp, np, net);
        if (hipGetLastError() != hipSuccess) st = CHS_E_HIP;
    }
    std::vector<double> ht(5 * (size_t)np), hn(5 * (size_t)np), hh((size_t)np);
    unsigned long long bad = 0;
    if (st == CHS_OK &&
        (hipMemcpyAsync(hh.data(), hp, sizeof(double) * np, hipMemcpyDeviceToHost, p->stream) != hipSuccess ||
         hipMemcpyAsync(ht.data(), tab, sizeof(double) * 5 * np, hipMemcpyDeviceToHost, p->stream) != hipSuccess ||
         hipMemcpyAsync(hn.data(), net, sizeof(double) * 5 * np, hipMemcpyDeviceToHost, p->stream) != hipSuccess ||
         hipMemcpyAsync(&bad, p->counters, sizeof(bad), hipMemcpyDeviceToHost, p->stream) != hipSuccess ||
         hipStreamSynchronize(p->stream) != hipSuccess))
        st = CHS_E_HIP;
    (void)hipFree(buf);
    if (st != CHS_OK) {
        set_error("chs_surrogate_certify: device evaluation failed");
        return st;
    }
    CHS_CHECK(bad == 0, CHS_E_RANGE, "chs_surrogate_certify: trained range leaves the property table");
    std::vector<unsigned char> flag(nbins, 0);
    long long ok = 0, ok4 = 0;
    for (int b = 0; b < nbins; ++b) {
        double worst = 0.0, worst4 = 0.0;
        for (int k = 0; k < NPT; ++k) {
            const size_t q = (size_t)b * NPT + k;
            // THERM_PROP_UPDATE leaves cells with |h| < 1e-10 unchanged in
            // both modes, and the lookup wrote no reference value for them
            if (std::fabs(hh[q]) < 1.0e-10) continue;
            for (int f = 0; f < NOUT; ++f) {
                const double tv = ht[(size_t)f * np + q], nv = hn[(size_t)f * np + q];
                double e = std::fabs(nv - tv) / std::max(std::fabs(tv), 1e-300);
                e = e != e ? 1e300 : e;
                worst = std::max(worst, e);
                if (f < 4) worst4 = std::max(worst4, e);   // outputs T, rho, mu, kappa
            }
        }
        // margin of 2 between the checked points and the bin's interior
        const bool c5 = worst <= 0.5 * tol, c4 = worst4 <= 0.5 * tol;
        flag[b] = (unsigned char)((c5 ? 1 : 0) | (c4 ? 2 : 0));
        ok += c5;
        ok4 += c4;
    }
    if (m->cert) (void)hipFree(m->cert);
    m->cert = nullptr;
    CHS_HIP(hipMalloc((void**)&m->cert, nbins));
    CHS_HIP(hipMemcpy(m->cert, flag.data(), nbins, hipMemcpyHostToDevice));
    m->nbins = nbins;
    m->cert_frac = (double)ok / nbins;
    m->cert_frac4 = (double)ok4 / nbins;
    if (frac) *frac = m->cert_frac;
    return CHS_OK;
}

static int eval(chs_props* p, const double* h, long long n, double* out) {
    CHS_CHECK(p->mlp && p->mlp->trained, CHS_E_ARG, "chs_surrogate_eval: no network");
    CHS_CHECK(h && out && n > 0, CHS_E_ARG, "chs_surrogate_eval: bad arguments");
    CHS_LAUNCH(k_mlp_eval, dim3((unsigned)std::min<long long>((n + 63) / 64, 2048)), dim3(256), sizeof(WLds), p->stream,
               p->mlp->wq, p->mlp->norm, h, n, out);
    CHS_HIP(hipStreamSynchronize(p->stream));
    return CHS_OK;
}

static int params(chs_props* p, int op, float* host, int count, double* norm) {
    CHS_CHECK(count == NPAR, CHS_E_ARG, "chs_surrogate_params: count must be 4613");
    CHS_CHECK(host != nullptr, CHS_E_ARG, "chs_surrogate_params: NULL buffer");
    if (op == 0) {
        CHS_CHECK(p->mlp && p->mlp->trained, CHS_E_ARG, "chs_surrogate_params: no network");
        CHS_HIP(hipMemcpy(host, p->mlp->par, sizeof(float) * NPAR, hipMemcpyDeviceToHost));
        if (norm) {
            const MlpNorm& nm = p->mlp->norm;
            norm[0] = nm.hc;
            norm[1] = nm.hs;
            norm[2] = nm.h_lo;
            norm[3] = nm.h_hi;
            for (int k = 0; k < NOUT; ++k) {
                norm[4 + k] = nm.m[k];
                norm[9 + k] = nm.s[k];
            }
        }
        return CHS_OK;
    }
    CHS_CHECK(op == 1 && norm != nullptr, CHS_E_ARG, "chs_surrogate_params: op 1 needs the normalisation");
    CHS_TRY(set_params(p, host));
    Mlp* m = p->mlp;
    m->norm.hc = norm[0];
    m->norm.hs = norm[1];
    m->norm.h_lo = norm[2];
    m->norm.h_hi = norm[3];
    for (int k = 0; k < NOUT; ++k) {
        m->norm.m[k] = (float)norm[4 + k];
        m->norm.s[k] = (float)norm[9 + k];
    }
    m->trained = true;
    if (m->cert) {
        (void)hipFree(m->cert);
        m->cert = nullptr;
    }
    CHS_HIP(hipStreamSynchronize(p->stream));
    return CHS_OK;
}

}  // namespace chs

extern "C" {

int chs_surrogate_train(chs_props* p, const double* h_dev, long long nsamp, int steps, int batch, double lr,
                        unsigned long long seed, double* final_loss) {
    CHS_CHECK(p != nullptr, CHS_E_ARG, "chs_surrogate_train: NULL handle");
    return chs::train(p, h_dev, nsamp, steps, batch, lr, seed, final_loss);
}

int chs_surrogate_certified(chs_props* p, double* all_outputs, double* without_cp) {
    CHS_CHECK(p && p->mlp && p->mlp->cert, CHS_E_ARG, "chs_surrogate_certified: the surrogate is not certified");
    if (all_outputs) *all_outputs = p->mlp->cert_frac;
    if (without_cp) *without_cp = p->mlp->cert_frac4;
    return CHS_OK;
}

int chs_surrogate_certify(chs_props* p, double tol, int nbins, double* certified_fraction) {
    CHS_CHECK(p != nullptr, CHS_E_ARG, "chs_surrogate_certify: NULL handle");
    return chs::certify(p, tol, nbins, certified_fraction);
}

int chs_surrogate_grad(chs_props* p, const double* h_dev, long long n, float* grad_host, double* loss) {
    CHS_CHECK(p != nullptr, CHS_E_ARG, "chs_surrogate_grad: NULL handle");
    return chs::grad(p, h_dev, n, grad_host, loss);
}

int chs_surrogate_eval(chs_props* p, const double* h_dev, long long n, double* out_dev) {
    CHS_CHECK(p != nullptr, CHS_E_ARG, "chs_surrogate_eval: NULL handle");
    return chs::eval(p, h_dev, n, out_dev);
}

int chs_surrogate_params(chs_props* p, int op, float* host, int count, double* norm) {
    CHS_CHECK(p != nullptr, CHS_E_ARG, "chs_surrogate_params: NULL handle");
    return chs::params(p, op, host, count, norm);
}

}  // extern "C"
