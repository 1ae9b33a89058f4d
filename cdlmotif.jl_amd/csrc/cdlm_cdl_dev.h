// cdlm_cdl_dev.h -- device side of the CDL ADMM (included once, by cdlm_cdl.hip).
//
// Reference path (CDLmotif.jl): src/CDL/train.jl:1-98, src/CDL/csc_kernels.jl:8-95,
// src/CDL/cdl_kernels.jl:2-77, src/SETUP/CDL_setup.jl:15-286, src/WRAPPER/wrap.jl:4-12.
//
// Formulation.  The reference keeps ~25 E-sized arrays (E = Lg*M*N) and takes one batched
// FFT per array per step.  Here the same iteration is computed from six E-sized arrays
// with the algebra rearranged so that each CSC iteration needs O(1) transforms per
// sequence instead of O(M):
//   * the CSC right-hand side only enters through V_m = (Y2-y2)+(Y3-y3) and U = Y1-y1;
//     with S_v = sum_m G1_m V_m, the rank-1 solve gives
//       X_m = v_m / c0 + g_m (*) h,      h  = F^-1[(c0 U - S_v) / (c0 (|G1|^2 + c0))]
//       DX  = F^-1[S_v / c0 + |G1|^2 h]                      (c0 = sherman_const_csc)
//     where (*) is circular cross-correlation with the filter g_m = gp(G1)[:, m], which is zero
//     past filter_len (prox_simplex_zeros!), so it costs filter_len taps per sample;
//     S_v = F(sum_m g_m (x) v_m) is a filter_len-tap circular convolution summed over m.
//   * X is held as the pair (Xt = v/c0, h): X = Xt + g (*) h is rebuilt on the fly by the
//     next kernel that reads X.  g is always the current G1: before G1 changes (the
//     dictionary step) X is written back materialised (Xt = X, h = 0) -- by update_best's
//     pass, which rebuilds all of X anyway, or else by the dictionary step itself.
//   * the dictionary step works on half spectra (dh, g1h, g3h, as the reference keeps dh, g1h,
//     g3h), g3 through g1 - g3: every sequence's g1 and g3 receive the same d and differ by the
//     filters only (g1 += d - G1, g3 += d - G3 from g1 = g3 = 0), so g1 - g3 = sum over the
//     steps of G3 - G1 is one (Hs, M) array; its solve is
//       dh_m = Qh_m + conj(Y3h_m) kappa,   kappa = (c1 F(G2-g2) - sum_m Y3h_m Qh_m)/(|Y3h|^2 + c1)
//     with Qh_m = (G1h_m - g1h_m) + (G3h_m - g3h_m) (c1 = sherman_const_cdl; no 1/c1, as
//     sherman_CDL_combine).  Y3 = Pi4(X + y3) is nonzero on every 4th sample only, so its
//     spectrum is a length Lg/4 DFT repeated 4 times: the only per-column transform of the
//     step.  dh and the new g1h are functions of the old g1h, Y3h and kappa, so those are what
//     the step stores (with the step's filter spectra); the next step rebuilds dh and g1h per
//     bin as it streams the column, and the means over n of d + g1, d + g3 and d come from
//     sums over n of the old g1h and of conj(Y3h) kappa, transformed back once per filter.
// All quantities are exact rewrites of the reference iteration; only rounding differs.
//
// Launch model.  Every kernel reads its problem from a device table of descriptors
// (CdlDev) and maps its block index to (problem, unit) through a prefix array, so one
// launch runs one problem or a batch of problems of different N, L, M and filter length.
// Template parameters (TAPS, NSL) are the batch maxima: filter taps past a problem's own
// filter length are zero (G1 vanishes off its support) and add exact zeros, and all halos
// wrap modulo the problem's own Lg.
//
// HBM layout: the reference's column-major (Lg, M, N) with l fastest; one work-group owns
// one sequence n and streams its contiguous (Lg x M) slab.  dh, g1h are (Hs, M, N)
// complex with Hs = Lg/2 + 1 rounded up to even (bins 0..Lg/2 and a zero pad).  y3 is stored without its
// stride-4 support samples: there y3 + X - Pi4(X + y3) = (y3 + X) - (X + y3) = 0 exactly
// (IEEE addition commutes) from the first iteration on, and y3 starts at 0, so the
// support samples are identically zero and only l % 4 != 0 is kept, 3 floats per slot.
#pragma once

#include <type_traits>

#include "cdlm_internal.h"

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kMaxTaps = 32;
constexpr int kPartStride = 8;  // doubles per sequence in the objective partials
constexpr int kDpart = 4;       // doubles per (bin, filter) and sequence chunk in the dictionary means

// Device-side scalar state of one problem.
struct CdlScal {
    float obj_best;
    int best_idx;   // which X_best slice buffer holds the best iterate
    int improved;
    int g1cur;      // which G1 buffer holds the current filters
    int xmat;       // 1: X is held materialised (h == 0 for every sequence)
    int pad;
    double best_stats[3];  // count, sum, sum of squares of positive X_best entries
    double red[kPartStride];
};

// Descriptor of one problem (all device pointers).
struct CdlDev {
    int L4, N, M, Lg, gamma, L, nch, trace_cap;
    // N is the number of sequences the per-sequence kernels see; a problem streamed through a
    // device window runs them chunk by chunk: the chunk's first sequence is n_off of Nt in all,
    // and its dictionary-mean chunks are red_c0 .. red_c0 + red_cn - 1 of nch (resident
    // problems: n_off = 0, Nt = N, all nch chunks)
    int n_off, Nt, red_c0, red_cn;
    float rho, sigma, lambda, sh_csc, sh_cdl;
    const float *s;
    const int *fl;
    float *Xt, *y2;                                // (Lg, M, N)
    float2 *g1h;                                   // (Hs, M, N): g1h of the step before the last one
    float2 *kap;                                   // (Hs, N): kappa of the last dictionary step
    float2 *d13h;                                  // (Hs, M): g1h - g3h, the same for every sequence
    float2 *G1hp, *Fp;                             // (Hs, M): G1h and G3h + d13h of the last dictionary step
    float *y3;                                     // (3 Lg/4, M, N): y3 off the stride-4 support
    float *h, *y1, *Y1, *DX, *G2, *g2, *Td;       // (Lg, N)
    float *G1b[2], *G3, *meand, *Dbest, *gsq;     // (Lg, M)
    float2 *G1hb[2], *G3h;                         // (Hs, M): half spectra of G1b[i], G3
    float *cS, *sS;                                // (Lg)
    float *xs[2];                                  // (L, M, N) X_best slices
    float2 *y3s;                                   // (Lg/8+1, M, N): half spectra of the last step's Y3h
    const float2 *tw;                              // (Lg) twiddles
    double *partials, *dpart, *trace;
    CdlScal *sc;
    const float *X0;       // (Lg, M, N) initial codes, or null: N(0,1) from seed
    const float *D0;       // (Lg, M) masked initial filters
    unsigned long long seed;
    FftPlan plan, plan4;
};

// Half-spectrum bins kept per column: Lg/2 + 1 rounded up to even, so that a column is a
// whole number of float4 (two bins each) and columns stay 16-byte aligned.
__host__ __device__ __forceinline__ int half_bins(int Lg) { return ((Lg / 2 + 1) + 1) & ~1; }
// Complex entries of the dictionary kernel's per-wave copy of a stored Y3h half spectrum
// (Lg/8 + 1 bins, rounded up to even).
__host__ __device__ __forceinline__ int dict_phx(int Lg) { return ((Lg / 8 + 1) + 1) & ~1; }

// Block -> (problem, local unit): the last p with starts[p] <= b.
__device__ __forceinline__ int find_problem(const int *__restrict__ starts, int np, int b, int &local) {
    int lo = 0, hi = np - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= b) lo = mid; else hi = mid - 1;
    }
    local = b - starts[lo];
    return lo;
}

// Device pointers read from the descriptor table carry no address space the compiler can
// prove, so every access goes through gp(), which makes it a global-memory access
// (global_load/global_store instead of flat instructions that also count against the
// LDS wait counter).
// Filter taps are read through the constant address space: the same word for every lane, and
// no kernel that reads them writes them, so they come through the scalar cache into scalar
// registers instead of taking vector-memory instructions beside the streamed state.
#if defined(__HIP_DEVICE_COMPILE__)
#define CDLM_GLOBAL __attribute__((address_space(1)))
#define CDLM_CONST __attribute__((address_space(4)))
#else
#define CDLM_GLOBAL   // the host pass only parses device code
#define CDLM_CONST
#endif
template <typename T>
__device__ __forceinline__ CDLM_GLOBAL T *gp(T *p) { return (CDLM_GLOBAL T *)p; }
__device__ __forceinline__ const float *g1cur_of(const CdlDev &P) { return P.G1b[P.sc->g1cur & 1]; }

__device__ __forceinline__ float4 ld4(const float *p, int s) { return ((const CDLM_GLOBAL float4 *)p)[s]; }
__device__ __forceinline__ void st4(float *p, int s, float4 v) { ((CDLM_GLOBAL float4 *)p)[s] = v; }
__device__ __forceinline__ float f4(const float4 &v, int q) { return q == 0 ? v.x : q == 1 ? v.y : q == 2 ? v.z : v.w; }
__device__ __forceinline__ float4 mk4(const float *a) { return make_float4(a[0], a[1], a[2], a[3]); }
__device__ __forceinline__ float4 sub4(const float4 a, const float4 b) {
    return make_float4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w);
}
// Samples 1..3 of slot s of a y3 column (3 floats per slot).
__device__ __forceinline__ float4 ld3(const float *p, int s) {
    const CDLM_GLOBAL float *q = gp(p) + 3 * s;
    return make_float4(0.f, q[0], q[1], q[2]);
}
__device__ __forceinline__ void st3(float *p, int s, const float *v) {
    CDLM_GLOBAL float *q = gp(p) + 3 * s;
    q[0] = v[1];
    q[1] = v[2];
    q[2] = v[3];
}

// Between the iterations of one k_csc launch a slot's state is carried as the two shrink /
// projection inputs u2 = X + y2 (in the y2 buffer) and u3 = X + y3 (in the Xt buffer; X on
// the stride-4 support, where y3 = 0), 8 floats per slot instead of Xt, y2, y3 (11).  This
// recovers, with the same float operations the iteration used, the Xt, y2 and y3 it would
// have stored: y2 = u2 - shrink(u2), y3 = u3 off the support, Xt = v / c0 with
// v = (Y2 - y2) + (Y3 - y3), bit for bit (the product is kept out of any contraction).
__device__ __forceinline__ void csc_unpack(float4 &xt, float4 &y2, float4 &y3, float lam, float inv_c0) {
    float a[4], b[4], c[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float t = f4(y2, q), u3 = f4(xt, q);
        const float at = fabsf(t) - lam;
        const float Y2 = at > 0.f ? copysignf(at, t) : 0.f;
        const float Y3 = (q == 0) ? u3 : 0.f;
        b[q] = t - Y2;
        c[q] = (q == 0) ? 0.f : u3;
        a[q] = __fmul_rn((Y2 - b[q]) + (Y3 - c[q]), inv_c0);
    }
    xt = mk4(a);
    y2 = mk4(b);
    y3 = mk4(c);
}

// X on the 4 samples of slot s: Xt + sum_j g[j] * h[l + j] (circular; hb4 holds h plus a
// wrapped tail of TAPS+4 samples, as float4).
template <int TAPS>
__device__ __forceinline__ void recon4(const float4 xt, const float4 *hb4, int s, const float *g, float *X) {
    float hw[TAPS + 4];
#pragma unroll
    for (int u = 0; u <= TAPS / 4; ++u) {
        float4 t = hb4[s + u];
        hw[4 * u] = t.x; hw[4 * u + 1] = t.y; hw[4 * u + 2] = t.z; hw[4 * u + 3] = t.w;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        float acc = f4(xt, q);
#pragma unroll
        for (int j = 0; j < TAPS; ++j) acc = fmaf(g[j], hw[q + j], acc);
        X[q] = acc;
    }
}

// z[4s..4s+3] += sum_j g[j] * v[(l - j) mod Lg]; vb4 holds v at float4 offset TAPS/4 with
// the wrapped head of TAPS samples in front.
template <int TAPS>
__device__ __forceinline__ void conv_acc4(const float4 *vb4, int s, const float *g, float4 *zacc4, bool first = false) {
    float vw[TAPS + 4];
#pragma unroll
    for (int u = 0; u <= TAPS / 4; ++u) {
        float4 t = vb4[s + u];
        vw[4 * u] = t.x; vw[4 * u + 1] = t.y; vw[4 * u + 2] = t.z; vw[4 * u + 3] = t.w;
    }
    // vw[0] takes no part in the sums; keeping it live keeps the loads whole 16-byte LDS reads
    // (without it the compiler reads the remaining 23 words as unaligned 4-byte pairs, whose
    // 16-byte lane stride puts 4-8 lanes on one bank)
    asm volatile("" ::"v"(vw[0]));
    float4 z = first ? make_float4(0.f, 0.f, 0.f, 0.f) : zacc4[s];
    float o[4] = {z.x, z.y, z.z, z.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        float acc = o[q];
#pragma unroll
        for (int j = 0; j < TAPS; ++j) acc = fmaf(g[j], vw[TAPS + q - j], acc);
        o[q] = acc;
    }
    zacc4[s] = mk4(o);
}

template <int TAPS>
__device__ __forceinline__ void load_taps(const float *G, int m, int Lg, float *g) {
    // the column's address made visibly uniform (it is: one problem per work-group, m per wave)
    const uint64_t a = (uint64_t)(G + (size_t)m * Lg);
    const uint64_t au = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(a >> 32)) << 32) |
                        (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)a);
    const CDLM_CONST float *Gc = (const CDLM_CONST float *)au;
    const int nl = __builtin_amdgcn_readfirstlane(Lg);
#pragma unroll
    for (int j = 0; j < TAPS; ++j) g[j] = (j < nl) ? Gc[j] : 0.f;
}

// h of sequence n into LDS with its circular tail: hb[Lg + t] = h[t mod Lg], t < TAPS + 4.
template <int TAPS>
__device__ __forceinline__ void load_h(const CdlDev &P, size_t ro, float *hb, int tid, int nthr) {
    const int Lg = P.Lg;
    for (int l = tid; l < Lg + TAPS + 4; l += nthr) hb[l] = gp(P.h)[ro + (l < Lg ? l : (l - Lg) % Lg)];
}

// Wrapped head of a v column (float4 units): vb4[TAPS/4 - u] = v4[(nslot - u) mod nslot].
template <int TAPS>
__device__ __forceinline__ void v_head(float4 *vb4, int nslot, int lane) {
    if (lane < TAPS / 4) {
        const int u = lane + 1;
        vb4[TAPS / 4 - u] = vb4[TAPS / 4 + ((nslot - u % nslot) % nslot)];
    }
}

// Offset of column (n, m) in the y3 array (3 floats per slot).
__device__ __forceinline__ size_t col3_of(int n, int m, int M, int nslot) { return ((size_t)n * M + m) * 3 * nslot; }

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x ^= x >> 30; x *= 0xbf58476d1ce4e5b9ULL;
    x ^= x >> 27; x *= 0x94d049bb133111ebULL;
    x ^= x >> 31;
    return x;
}
// N(0,1) sample for element e (Box-Muller on two 24-bit uniforms).
__device__ __forceinline__ float gauss_of(uint64_t seed, uint64_t e) {
    uint64_t r = mix64(seed * 0x9e3779b97f4a7c15ULL + e + 1);
    float u1 = ((float)((r >> 40) & 0xffffff) + 0.5f) * (1.0f / 16777216.0f);
    float u2 = ((float)((r >> 16) & 0xffffff) + 0.5f) * (1.0f / 16777216.0f);
    return sqrtf(-2.0f * logf(u1)) * cosf(6.28318530717958647f * u2);
}

// ---------------------------------------------------------------------------- reset
// Column m of the initial filters into G1 (buffer 0), G3, the D_best candidate and D_best
// (CDL_setup.jl:183-189,233-236); the scalar state starts over.
__global__ __launch_bounds__(kThreads) void k_reset(const CdlDev *__restrict__ tab, const int *__restrict__ starts,
                                                    int np) {
    int m;
    const CdlDev &P = tab[find_problem(starts, np, blockIdx.x, m)];
    const int Lg = P.Lg;
    for (int l = threadIdx.x; l < Lg; l += kThreads) {
        const size_t i = (size_t)m * Lg + l;
        const float v = gp(P.D0)[i];
        gp(P.G1b[0])[i] = v;
        gp(P.G3)[i] = v;
        gp(P.meand)[i] = v;
        gp(P.Dbest)[i] = v;
    }
    if (m == 0 && threadIdx.x == 0) {
        CdlScal *sc = P.sc;
        sc->obj_best = __builtin_inff();
        sc->best_idx = 0;
        sc->improved = 0;
        sc->g1cur = 0;
        sc->xmat = 1;
        for (int i = 0; i < 3; ++i) sc->best_stats[i] = 0.0;
        for (int i = 0; i < kPartStride; ++i) sc->red[i] = 0.0;
    }
}

// Per member p: the first cnt trace entries and the scalar state, into
// out[slot[p] * (cnt + kGatherExtra) ...] (one host copy returns the whole batch).
constexpr int kGatherExtra = 8;
__global__ void k_gather(const CdlDev *__restrict__ tab, const int *__restrict__ slot, int np, int cnt,
                         double *__restrict__ out) {
    const int p = blockIdx.x;
    if (p >= np) return;
    const CdlDev &P = tab[p];
    double *o = out + (size_t)slot[p] * (cnt + kGatherExtra);
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) o[i] = (i < P.trace_cap) ? gp(P.trace)[i] : 0.0;
    if (threadIdx.x == 0) {
        const CdlScal *sc = P.sc;
        o[cnt + 0] = (double)sc->obj_best;
        o[cnt + 1] = sc->best_stats[0];
        o[cnt + 2] = sc->best_stats[1];
        o[cnt + 3] = sc->best_stats[2];
        o[cnt + 4] = (double)sc->improved;
        o[cnt + 5] = sc->red[5];
        o[cnt + 6] = (double)sc->best_idx;
        o[cnt + 7] = (double)sc->g1cur;
    }
}

// ---------------------------------------------------------------------------- init
// CDL_setup.jl:183-230 for one sequence: X = Pi4(X0); DX = sum_m D_m (x) X_m (the
// spectral product of :201-203 as a circular convolution); d = D (as F(D)); Y1 = y = g = 0;
// G2 = DX.
template <int TAPS>
__global__ __launch_bounds__(kThreads) void k_init(const CdlDev *__restrict__ tab, const int *__restrict__ starts,
                                                   int np) {
    extern __shared__ float4 smem4[];
    int n;
    const CdlDev &P = tab[find_problem(starts, np, blockIdx.x, n)];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int Lg = P.Lg, M = P.M, nslot = Lg >> 2;
    const int vb4Len = (Lg + TAPS) / 4;
    float4 *vb4 = smem4 + w * vb4Len;
    float *zbase = reinterpret_cast<float *>(smem4 + kWaves * vb4Len);
    float4 *zacc4 = reinterpret_cast<float4 *>(zbase) + w * (Lg / 4);
    for (int i = tid; i < kWaves * Lg; i += kThreads) zbase[i] = 0.f;
    __syncthreads();
    const float *G1 = g1cur_of(P);
    const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int m = w; m < M; m += kWaves) {
        const int mm = __builtin_amdgcn_readfirstlane(m);
        const size_t col = ((size_t)n * M + mm) * Lg;
        float g[TAPS];
        load_taps<TAPS>(G1, mm, Lg, g);
        for (int s = lane; s < nslot; s += 64) {
            const size_t colg = ((size_t)(P.n_off + n) * M + mm) * Lg;   // the element's index in the whole problem
            const float x = P.X0 ? gp(P.X0)[col + 4 * s] : gauss_of(P.seed, colg + 4 * (size_t)s);
            const float4 xp = make_float4(x, 0.f, 0.f, 0.f);
            st4(P.Xt + col, s, xp);
            st4(P.y2 + col, s, zero4);
            st3(P.y3 + col3_of(n, mm, M, nslot), s, reinterpret_cast<const float *>(&zero4));
            vb4[TAPS / 4 + s] = xp;
        }
        // d = D (dh = F(D), CDL_setup.jl:207-208), g1 = g3 = 0: the dictionary state A = 0,
        // Y3h = 0 and kappa = 0 with the filter spectra G1hp = F(D) and Fp = 0 (g1 - g3 = 0)
        {
            const int Hs = half_bins(Lg), Ph = (Lg >> 2) / 2 + 1;
            const size_t colh = ((size_t)n * M + mm) * Hs;
            const float2 *G1h = P.G1hb[P.sc->g1cur & 1] + (size_t)mm * Hs;
            for (int t = lane; t < Hs / 2; t += 64) {
                st4(reinterpret_cast<float *>(P.g1h + colh), t, zero4);
                if (n == 0) {
                    st4(reinterpret_cast<float *>(P.d13h + (size_t)mm * Hs), t, zero4);
                    st4(reinterpret_cast<float *>(P.Fp + (size_t)mm * Hs), t, zero4);
                    st4(reinterpret_cast<float *>(P.G1hp + (size_t)mm * Hs), t, ld4(reinterpret_cast<const float *>(G1h), t));
                }
            }
            for (int k = lane; k < Ph; k += 64) gp(P.y3s)[((size_t)n * M + mm) * Ph + k] = make_float2(0.f, 0.f);
        }
        wave_lds_sync();
        v_head<TAPS>(vb4, nslot, lane);
        wave_lds_sync();
        for (int s = lane; s < nslot; s += 64) conv_acc4<TAPS>(vb4, s, g, zacc4);
        wave_lds_sync();
    }
    __syncthreads();
    const size_t ro = (size_t)n * Lg;
    for (int l = tid; l < Lg; l += kThreads) {
        float z = zbase[l];
#pragma unroll
        for (int ww = 1; ww < kWaves; ++ww) z += zbase[ww * Lg + l];
        gp(P.DX)[ro + l] = z;
        gp(P.G2)[ro + l] = z;
        gp(P.h)[ro + l] = 0.f;
        gp(P.y1)[ro + l] = 0.f;
        gp(P.Y1)[ro + l] = 0.f;
        gp(P.g2)[ro + l] = 0.f;
        gp(P.Td)[ro + l] = 0.f;
    }
    const int Hs = half_bins(Lg);
    for (int k = tid; k < Hs; k += kThreads) gp(P.kap)[(size_t)n * Hs + k] = make_float2(0.f, 0.f);
}

// ---------------------------------------------------------------------------- CSC
// `iters` iterations of code_update! (train.jl:5-43) for one sequence.  The iterations of
// different sequences are independent (the dictionary spectrum is fixed during code_update!),
// so a work-group runs all of its sequence's iterations back to back and h stays in LDS
// between them (written to HBM after the last).  Between two iterations of the launch the
// streamed state is the pair (u2, u3) of csc_unpack, 8 floats per slot, in the Xt and y2
// buffers: the first iteration reads (Xt, y2, y3) and the last writes them, so every other
// kernel sees the usual state.  WAVES waves per work-group; wave w owns the
// columns m = w, w + WAVES, ...; a lane owns the float4 slots s = lane + 64 i (i < NSL) of a
// column, the same in every iteration.  With WAVES = 4 (three work-groups per CU) column
// m + WAVES's loads are issued before column m is processed (two register sets); with
// WAVES = 8 (two work-groups per CU where the LDS allows, so that one streams while the other
// runs its per-sequence transforms) and WAVES = 16 (one work-group per CU) there is no
// prefetch, at most 128 VGPRs, and the waves hide the latency instead.  With one or two
// sequences per CU resident, a sequence's state (u2, u3: 0.43 MB at Lg = 840, M = 64) is
// largely re-read from the last-level cache in the next iteration rather than from HBM.
// The (Lg) vectors y1, Y1, DX and the data window stay in LDS for the whole launch, and the
// next iteration's first column is loaded before the per-sequence transforms.
// LDS (float4 units): twiddles (Lg/2) | h with wrapped tail | WAVES x v column with wrapped
// head | WAVES x z partial | y1, Y1, DX, data window (Lg each); the FFT ping-pong buffers
// reuse the v columns after the column loop.
template <int WAVES, int NSL, int TAPS>
struct CscShape {
    static constexpr int kThr = 64 * WAVES;
    static constexpr bool kPrefetch = WAVES <= 4;
    static constexpr int kWavesPerSimd = WAVES <= 4 ? (NSL <= 4 && TAPS <= 20 ? 3 : 2) : 4;
};

template <int TAPS, int NSL, int WAVES>
__global__ __launch_bounds__(64 * WAVES, (CscShape<WAVES, NSL, TAPS>::kWavesPerSimd)) void k_csc(
    const CdlDev *__restrict__ tab, const int *__restrict__ starts, int np, int iters) {
    constexpr int kThr = CscShape<WAVES, NSL, TAPS>::kThr;
    constexpr bool kPf = CscShape<WAVES, NSL, TAPS>::kPrefetch;
    constexpr int kSg = kPf ? NSL : 2;     // slots loaded together without prefetch (4: no faster)
    extern __shared__ float4 smem4[];
    int n;
    const CdlDev &P = tab[find_problem(starts, np, blockIdx.x, n)];
    const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int Lg = P.Lg, M = P.M, nslot = Lg >> 2;
    const int hb4Len = (Lg + TAPS + 4) / 4, vb4Len = (Lg + TAPS) / 4;
    float2 *tw = reinterpret_cast<float2 *>(smem4);
    float4 *hb4 = smem4 + Lg / 2;
    float4 *vbase4 = hb4 + hb4Len;
    float4 *zbase4 = vbase4 + WAVES * vb4Len;
    float *hb = reinterpret_cast<float *>(hb4);
    float *zbase = reinterpret_cast<float *>(zbase4);
    float *y1s = zbase + WAVES * Lg;                 // y1, Y1, DX and the data window, kept in LDS
    float *Y1s = y1s + Lg;
    float *DXs = Y1s + Lg;
    float *bs = DXs + Lg;
    float2 *fa = reinterpret_cast<float2 *>(vbase4);
    float2 *fb = fa + Lg;
    const size_t ro = (size_t)n * Lg;
    for (int l = tid; l < Lg; l += kThr) {
        tw[l] = gp(P.tw)[l];
        y1s[l] = gp(P.y1)[ro + l];
        Y1s[l] = gp(P.Y1)[ro + l];
        DXs[l] = gp(P.DX)[ro + l];
        const bool win = (l >= P.gamma) && (l < P.gamma + P.L4);
        bs[l] = win ? gp(P.s)[(size_t)n * P.L4 + (l - P.gamma)] : 0.f;
    }
    load_h<TAPS>(P, ro, hb, tid, kThr);
    if (n == 0 && tid == 0) P.sc->xmat = 0;   // this launch leaves h != 0
    const float *G1 = g1cur_of(P);
    const float lam = P.lambda;
    const float inv_c0 = 1.0f / P.sh_csc;
    float4 *vb4 = vbase4 + w * vb4Len;
    float4 *zacc4 = zbase4 + w * (Lg / 4);
    float4 nx[kPf ? NSL : 2], n2[kPf ? NSL : 2], n3[kPf ? NSL : 2];
    bool have_next = false;       // without prefetch: the first two slots of column w are loaded
    // packed: the slots hold (u3, u2) (every iteration after a launch's first), so y3 is not read
    auto load_col = [&](int m, float4 *a, float4 *b, float4 *c, bool packed) {
        const size_t col = ((size_t)n * M + m) * Lg, col3 = col3_of(n, m, M, nslot);
#pragma unroll
        for (int i = 0; i < NSL; ++i) {
            const int s = lane + 64 * i;
            if (s < nslot) {
                a[i] = ld4(P.Xt + col, s);
                b[i] = ld4(P.y2 + col, s);
                c[i] = packed ? make_float4(0.f, 0.f, 0.f, 0.f) : ld3(P.y3 + col3, s);
            }
        }
    };
    if (kPf && w < M) load_col(w, nx, n2, n3, false);
    const int nw = M < WAVES ? M : WAVES;       // waves that own columns (and z partials)
    // The column pass of one iteration, compiled once per (rdU, wrU): (u2, u3) read / written
    // instead of (Xt, y2, y3).  With the two flags as constants the slot code has no branches
    // on them, so the compiler keeps every value in place instead of merging two paths with
    // register copies.
    auto columns = [&](auto rdU_t, auto wrU_t) {
    constexpr bool rdU = decltype(rdU_t)::value, wrU = decltype(wrU_t)::value;
    for (int m = w; m < M; m += WAVES) {
        const int mm = __builtin_amdgcn_readfirstlane(m);
        const size_t col = ((size_t)n * M + mm) * Lg, col3 = col3_of(n, mm, M, nslot);
        float4 xt[kPf ? NSL : 1], a2[kPf ? NSL : 1], a3[kPf ? NSL : 1];
        if constexpr (kPf) {
#pragma unroll
            for (int i = 0; i < NSL; ++i) { xt[i] = nx[i]; a2[i] = n2[i]; a3[i] = n3[i]; }
            if (mm + WAVES < M) load_col(mm + WAVES, nx, n2, n3, rdU);
        }
        float g[TAPS];
        load_taps<TAPS>(G1, mm, Lg, g);
#pragma unroll
        for (int i0 = 0; i0 < NSL; i0 += kSg) {
            float4 lx[kSg], l2[kSg], l3[kSg];
            if constexpr (!kPf) {
                if (i0 == 0 && have_next) {
#pragma unroll
                    for (int u = 0; u < kSg; ++u) { lx[u] = nx[u]; l2[u] = n2[u]; l3[u] = n3[u]; }
                } else {
#pragma unroll
                    for (int u = 0; u < kSg; ++u) {
                        const int s = lane + 64 * (i0 + u);
                        if (s < nslot) {
                            lx[u] = ld4(P.Xt + col, s);
                            l2[u] = ld4(P.y2 + col, s);
                            l3[u] = rdU ? make_float4(0.f, 0.f, 0.f, 0.f) : ld3(P.y3 + col3, s);
                        }
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < kSg; ++u) {
                const int i = i0 + u;
                const int s = lane + 64 * i;
                if (s < nslot) {
                    float4 xv = kPf ? xt[kPf ? i : 0] : lx[u];
                    float4 b2 = kPf ? a2[kPf ? i : 0] : l2[u];
                    float4 b3 = kPf ? a3[kPf ? i : 0] : l3[u];
                    if (rdU) csc_unpack(xv, b2, b3, lam, inv_c0);
                    float X[4];
                    recon4<TAPS>(xv, hb4, s, g, X);
                    float u2[4], y2n[4], y3n[4], v[4], xn[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const float x = X[q];
                        const float t = x + f4(b2, q);                      // shrink! (csc_kernels.jl:58-59)
                        const float at = fabsf(t) - lam;
                        const float Y2 = at > 0.f ? copysignf(at, t) : 0.f;
                        const float Y3 = (q == 0) ? x : 0.f;                  // projectionX! (:73-77), y3 = 0 there
                        u2[q] = t;
                        y2n[q] = t - Y2;                                      // train.jl:13, (y2 + X) - Y2
                        y3n[q] = (q == 0) ? 0.f : f4(b3, q) + x;              // train.jl:14
                        v[q] = (Y2 - y2n[q]) + (Y3 - y3n[q]);
                        xn[q] = __fmul_rn(v[q], inv_c0);
                    }
                    if (wrU) {
                        const float u3[4] = {X[0], y3n[1], y3n[2], y3n[3]};
                        st4(P.y2 + col, s, mk4(u2));
                        st4(P.Xt + col, s, mk4(u3));
                    } else {
                        st4(P.y2 + col, s, mk4(y2n));
                        st3(P.y3 + col3, s, y3n);
                        st4(P.Xt + col, s, mk4(xn));
                    }
                    vb4[TAPS / 4 + s] = mk4(v);
                }
            }
        }
        wave_lds_sync();
        v_head<TAPS>(vb4, nslot, lane);
        wave_lds_sync();
#pragma unroll
        for (int i = 0; i < NSL; ++i) {
            const int s = lane + 64 * i;
            if (s < nslot) conv_acc4<TAPS>(vb4, s, g, zacc4, mm == w);   // the first column starts the partial
        }
        wave_lds_sync();
        if constexpr (!kPf) have_next = false;
    }
    };
#pragma nounroll
  for (int it = 0; it < iters; ++it) {
    __syncthreads();
    if (it == 0) {
        if (iters == 1) columns(std::false_type{}, std::false_type{});
        else columns(std::false_type{}, std::true_type{});
    } else if (it + 1 < iters) {
        columns(std::true_type{}, std::true_type{});
    } else {
        columns(std::true_type{}, std::false_type{});
    }
    // the next iteration's first column does not depend on what follows: load it now (all of
    // it with prefetch, its first two slots without)
    if (it + 1 < iters && w < M) {
        if constexpr (kPf) {
            load_col(w, nx, n2, n3, true);
        } else {
            const size_t col = ((size_t)n * M + w) * Lg;
#pragma unroll
            for (int u = 0; u < kSg; ++u) {
                const int s = lane + 64 * u;
                if (s < nslot) { nx[u] = ld4(P.Xt + col, s); n2[u] = ld4(P.y2 + col, s); n3[u] = make_float4(0.f, 0.f, 0.f, 0.f); }
            }
            have_next = true;
        }
    }
    __syncthreads();

    // z = sum over waves in wave order; u = Y1 - y1' with y1' = y1 + DX - Y1 (train.jl:12)
    for (int l = tid; l < Lg; l += kThr) {
        float z = zbase[l];
        for (int ww = 1; ww < nw; ++ww) z += zbase[ww * Lg + l];
        const float Y1o = Y1s[l];
        const float y1n = y1s[l] + DXs[l] - Y1o;
        y1s[l] = y1n;
        fa[l] = make_float2(z, Y1o - y1n);
    }
    __syncthreads();
    float2 *C = team_fft<false>(fa, fb, P.plan, tw, tid, kThr, BlockSync{});
    float2 *O = (C == fa) ? fb : fa;
    const float c0 = P.sh_csc;
    for (int k = tid; k < Lg; k += kThr) {
        const float2 Ck = C[k], Cm = C[k == 0 ? 0 : Lg - k];
        const float2 Z = make_float2(0.5f * (Ck.x + Cm.x), 0.5f * (Ck.y - Cm.y));   // F(z)
        const float2 U = make_float2(0.5f * (Ck.y + Cm.y), -0.5f * (Ck.x - Cm.x));  // F(u)
        const float ss = gp(P.sS)[k];             // the spectral constants s_k, c_k (read once per
        const float den = c0 * gp(P.cS)[k];       // iteration, from the L2 cache)
        const float2 hh = make_float2((c0 * U.x - Z.x) / den, (c0 * U.y - Z.y) / den);
        const float2 dx = make_float2(Z.x * inv_c0 + ss * hh.x, Z.y * inv_c0 + ss * hh.y);
        O[k] = make_float2(hh.x - dx.y, hh.y + dx.x);   // h_hat + i DX_hat
    }
    __syncthreads();
    float2 *R = team_fft<true>(O, C, P.plan, tw, tid, kThr, BlockSync{});
    const float scale = 1.0f / (float)Lg;
    const float rho = P.rho;
    const bool last = it + 1 == iters;
    for (int l = tid; l < Lg; l += kThr) {
        const float hn = R[l].x * scale;
        const float DXn = R[l].y * scale;
        const bool win = (l >= P.gamma) && (l < P.gamma + P.L4);
        const float wl = win ? 1.f : 0.f;
        const float y1n = y1s[l];
        const float Y1n = (bs[l] + rho * (DXn + y1n)) / (wl + rho);   // update_Y1! (csc_kernels.jl:91)
        hb[l] = hn;
        for (int t = l; t < TAPS + 4; t += Lg) hb[Lg + t] = hn;   // the wrapped tail, h[t mod Lg]
        DXs[l] = DXn;
        Y1s[l] = Y1n;
        if (last) {
            gp(P.h)[ro + l] = hn;
            gp(P.DX)[ro + l] = DXn;
            gp(P.y1)[ro + l] = y1n;
            gp(P.Y1)[ro + l] = Y1n;
        }
    }
  }
}

// ---------------------------------------------------------------------------- CSC on the matrix cores
// The same iterations as k_csc (train.jl:5-43), one wave per column as there (so every
// load and store of the streamed state is one contiguous 1 KB run per wave), with the
// filter_len-tap correlation that rebuilds X and the convolution that forms z on
// v_mfma_f32_16x16x4_f32 instead of the vector ALUs.  Positions 256b .. 256b + 255 of a
// column form a 16 x 16 block D[v][u] = position 256b + 16u + v; in the MFMA result layout
// (register r of lane l is D[4(l/16) + r][l%16]) lane (a = l/16, c = l%16) holds positions
// 256b + 16c + 4a .. + 3, which is float4 slot 64b + 4c + a of the column.
//   rebuild:  D = Xt + T H with the Toeplitz T[v][k] = g[k - v] (16 x (TAPS + 16), zero off the
//             filter) and the Hankel block H[k][u] = h[256b + 16u + k];
//   convolve: z-block += T' V with T'[v][k] = g[v - k + TAPS - 1] and V[k][u] =
//             v[256b + 16u + k - TAPS + 1] (v of the column in LDS with a wrapped head), summed
//             over the wave's columns in the accumulator; lane l supplies A[l%16][l/16] and
//             B[l/16][l%16], so T and T' are KS = (TAPS + 16)/4 registers per column, read
//             from the filter in LDS with zero margins.
// The B operands of a lane step 16 positions per lane across a block, so h and v sit in LDS
// with 4 floats of padding after every 16 (word q at q + 4 floor(q/16)): the 64 lanes of one
// operand read then fall in distinct banks (two-way at most).
// The Toeplitz form spends TAPS + 16 products per output instead of TAPS; the matrix cores
// have the room, and the vector ALUs keep only the shrink / dual updates.  The wave partials of
// z are combined in a fixed order (waves 0 and 1 store, later waves add in rounds of two, then
// the two slots in order), so results are deterministic and do not depend on the launch's other problems.
typedef float floatx4 __attribute__((ext_vector_type(4)));

__host__ __device__ constexpr int csctz_ks(int taps) { return (taps + 16) / 4; }
// LDS word of position-word q in the padded h and v buffers
__host__ __device__ constexpr int csctz_pad(int q) { return q + 4 * (q >> 4); }
// words of the column v buffer (and of h with its wrapped tail), and their padded floats
__host__ __device__ constexpr int csctz_vl(int Lg, int taps) { return 256 * ((Lg + 255) / 256) + taps + 4; }
__host__ __device__ constexpr int csctz_vlp(int Lg, int taps) { return csctz_pad(csctz_vl(Lg, taps) + 15); }
__host__ __device__ constexpr int csctz_nzs(int waves) { return waves < 2 ? waves : 2; }
// Dynamic LDS of k_csc_tz in floats: twiddles, h, the waves' filters, the waves' v columns
// (also the FFT buffers), the z slots, y1, Y1, DX and the data window.
__host__ __device__ constexpr size_t csctz_lds_floats(int Lg, int taps, int waves) {
    return 2 * (size_t)Lg + (size_t)csctz_vlp(Lg, taps) + 64 * (size_t)waves + (size_t)waves * csctz_vlp(Lg, taps) +
           (size_t)csctz_nzs(waves) * Lg + 4 * (size_t)Lg;
}

// Registers: two 8-wave work-groups per CU for padded lengths up to 1024 (four 256-position
// blocks), one beyond that (the LDS of longer columns holds one), so 256 registers per lane.
template <int TAPS, int NBT, int WAVES>
__global__ __launch_bounds__(64 * WAVES, (NBT > 4 && WAVES == 8 ? 2 : 4)) void k_csc_tz(
    const CdlDev *__restrict__ tab, const int *__restrict__ starts, int np, int iters) {
    constexpr int kThr = 64 * WAVES;
    constexpr int kKS = csctz_ks(TAPS);
    extern __shared__ float4 smem4[];
    int n;
    const CdlDev &P = tab[find_problem(starts, np, blockIdx.x, n)];
    const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int a = lane >> 4, c = lane & 15;
    const int Lg = P.Lg, M = P.M, nslot = Lg >> 2;
    const int nb = (Lg + 255) >> 8;
    const int vl = csctz_vl(Lg, TAPS), vlp = csctz_vlp(Lg, TAPS);
    constexpr int nzs = csctz_nzs(WAVES);
    float *sm = reinterpret_cast<float *>(smem4);
    float2 *tw = reinterpret_cast<float2 *>(sm);
    float *hb = sm + 2 * Lg;                    // word q = h[q mod Lg], q < vl (padded)
    float *gw = hb + vlp + 64 * w;              // this wave's filter: gw[16 + j] = g[j], zero margins
    float *vbase = hb + vlp + 64 * WAVES;
    float *vb = vbase + w * vlp;                // this wave's v column: word q = v[q - TAPS] (padded)
    float *zs = vbase + WAVES * vlp;            // nzs z slots of length Lg
    float *y1s = zs + nzs * Lg;                 // y1, Y1, DX and the data window, kept in LDS
    float *Y1s = y1s + Lg;
    float *DXs = Y1s + Lg;
    float *bs = DXs + Lg;
    float2 *fa = reinterpret_cast<float2 *>(vbase);
    float2 *fb = fa + Lg;
    const size_t ro = (size_t)n * Lg;
    for (int l = tid; l < Lg; l += kThr) {
        tw[l] = gp(P.tw)[l];
        y1s[l] = gp(P.y1)[ro + l];
        Y1s[l] = gp(P.Y1)[ro + l];
        DXs[l] = gp(P.DX)[ro + l];
        const bool win = (l >= P.gamma) && (l < P.gamma + P.L4);
        bs[l] = win ? gp(P.s)[(size_t)n * P.L4 + (l - P.gamma)] : 0.f;
    }
    for (int l = tid; l < vl; l += kThr) hb[csctz_pad(l)] = gp(P.h)[ro + (l < Lg ? l : l % Lg)];
    if (n == 0 && tid == 0) P.sc->xmat = 0;   // this launch leaves h != 0
    const float *G1 = g1cur_of(P);
    const float lam = P.lambda;
    const float inv_c0 = 1.0f / P.sh_csc;
    const int nw = M < WAVES ? M : WAVES;     // waves that own columns
    // streamed state of block b of column mc (packed: (u3, u2), y3 not read); lanes past the
    // column keep what they hold (their rows of the result are never used).  Loading the next
    // block ahead of the current one measured slower (3.61 vs 3.52 ms per iteration at
    // Lg = 840, M = 64): the matrix-core chains already overlap the loads of other waves.
    float *const Xtp = P.Xt, *const y2p = P.y2, *const y3p = P.y3;
    auto load_blk = [&](int mc, int b, bool packed, float4 &x, float4 &y, float4 &z) {
        const int s = 64 * b + 4 * c + a;
        if (s < nslot) {
            const size_t col = ((size_t)n * M + mc) * Lg;
            x = ld4(Xtp + col, s);
            y = ld4(y2p + col, s);
            z = packed ? make_float4(0.f, 0.f, 0.f, 0.f) : ld3(y3p + col3_of(n, mc, M, nslot), s);
        }
    };
    // V operand words 256b + 16c + 4ks + a + 1, padded: 320b + 20c + a + 1 + 4ks + 4(ks/4), plus
    // 4 where the word crosses into the next row of 16 (a = 3 and ks = 3 mod 4)
    const int vbase0 = 20 * c + a + 1, vbase3 = vbase0 + (a == 3 ? 4 : 0);
    // The column pass of one iteration, compiled once per (rdU, wrU): (u2, u3) read / written
    // instead of (Xt, y2, y3); with the flags constant the block code has no branches on them.
    floatx4 zacc[NBT];
    auto columns = [&](auto rdU_t, auto wrU_t) {
    constexpr bool rdU = decltype(rdU_t)::value, wrU = decltype(wrU_t)::value;
    for (int m = w; m < M; m += WAVES) {
        const int mm = __builtin_amdgcn_readfirstlane(m);
        const size_t col = ((size_t)n * M + mm) * Lg, col3 = col3_of(n, mm, M, nslot);
        {
            const int j = lane - 16;
            gw[lane] = (j >= 0 && j < TAPS && j < Lg) ? gp(G1)[(size_t)mm * Lg + j] : 0.f;
        }
        wave_lds_sync();
        float tr[kKS], tc[kKS];
#pragma unroll
        for (int ks = 0; ks < kKS; ++ks) {
            tr[ks] = gw[16 + 4 * ks + a - c];
            tc[ks] = gw[16 + c - 4 * ks - a + TAPS - 1];
        }
        // X, the shrink / dual updates and v, two blocks at a time: both blocks' loads are issued
        // first, then g * h of both (it does not depend on them: two independent MFMA chains per
        // block, K-steps ks even / odd), then the elementwise updates; X = Xt + the chains' sum
        auto update = [&](int b, float4 xv, float4 b2, float4 b3, const floatx4 &c0, const floatx4 &c1) {
            const int s = 64 * b + 4 * c + a;
            if (rdU) csc_unpack(xv, b2, b3, lam, inv_c0);
            float x4[4], u2[4], y2n[4], y3n[4], v[4], xn[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float x = f4(xv, q) + (c0[q] + c1[q]);
                x4[q] = x;
                const float t = x + f4(b2, q);                      // shrink! (csc_kernels.jl:58-59)
                const float at = fabsf(t) - lam;
                const float Y2 = at > 0.f ? copysignf(at, t) : 0.f;
                const float Y3 = (q == 0) ? x : 0.f;                  // projectionX! (:73-77), y3 = 0 there
                u2[q] = t;
                y2n[q] = t - Y2;                                      // train.jl:13, (y2 + X) - Y2
                y3n[q] = (q == 0) ? 0.f : f4(b3, q) + x;              // train.jl:14
                v[q] = (Y2 - y2n[q]) + (Y3 - y3n[q]);
                xn[q] = __fmul_rn(v[q], inv_c0);
            }
            if (s < nslot) {
                if (wrU) {
                    const float u3[4] = {x4[0], y3n[1], y3n[2], y3n[3]};
                    st4(y2p + col, s, mk4(u2));
                    st4(Xtp + col, s, mk4(u3));
                } else {
                    st4(y2p + col, s, mk4(y2n));
                    st3(y3p + col3, s, y3n);
                    st4(Xtp + col, s, mk4(xn));
                }
                *reinterpret_cast<float4 *>(vb + csctz_pad(4 * s + TAPS)) = mk4(v);
            }
        };
        auto corr = [&](int b, floatx4 &c0, floatx4 &c1) {
            // word 256b + 16c + 4ks + a: (4ks + a) / 16 = ks / 4
            const float *hp = hb + 320 * b + 20 * c + a;
            c0 = c1 = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < kKS; ++ks) {
                if (ks & 1) c1 = __builtin_amdgcn_mfma_f32_16x16x4f32(tr[ks], hp[4 * ks + 4 * (ks >> 2)], c1, 0, 0, 0);
                else c0 = __builtin_amdgcn_mfma_f32_16x16x4f32(tr[ks], hp[4 * ks + 4 * (ks >> 2)], c0, 0, 0, 0);
            }
        };
#pragma unroll
        for (int b = 0; b < NBT; b += 2) {
            if (b < nb) {
                const bool two = b + 1 < nb;
                float4 xa = make_float4(0.f, 0.f, 0.f, 0.f), a2 = xa, a3 = xa, xb = xa, bb2 = xa, bb3 = xa;
                load_blk(mm, b, rdU, xa, a2, a3);
                if (two) load_blk(mm, b + 1, rdU, xb, bb2, bb3);
                floatx4 ca0, ca1, cb0, cb1;
                corr(b, ca0, ca1);
                if (two) corr(b + 1, cb0, cb1);
                update(b, xa, a2, a3, ca0, ca1);
                if (two) update(b + 1, xb, bb2, bb3, cb0, cb1);
            }
        }
        wave_lds_sync();
        if (lane < TAPS) vb[csctz_pad(lane)] = vb[csctz_pad(lane + Lg)];   // wrapped head, v[lane - TAPS]
        wave_lds_sync();
        // z-blocks += T' V
#pragma unroll
        for (int b = 0; b < NBT; ++b) {
            if (b < nb) {
                const float *vp0 = vb + 320 * b + vbase0, *vp3 = vb + 320 * b + vbase3;
#pragma unroll
                for (int ks = 0; ks < kKS; ++ks)
                    zacc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                        tc[ks], ((ks & 3) == 3 ? vp3 : vp0)[4 * ks + 4 * (ks >> 2)], zacc[b], 0, 0, 0);
            }
        }
        wave_lds_sync();   // the next column rewrites gw and vb
    }
    };
#pragma nounroll
  for (int it = 0; it < iters; ++it) {
    __syncthreads();
    // v past Lg is read only against zero Toeplitz entries; keep it finite (the region held
    // FFT data or nothing)
    for (int i = TAPS + Lg + lane; i < vl; i += 64) vb[csctz_pad(i)] = 0.f;
#pragma unroll
    for (int b = 0; b < NBT; ++b) zacc[b] = floatx4{0.f, 0.f, 0.f, 0.f};
    if (it == 0) {
        if (iters == 1) columns(std::false_type{}, std::false_type{});
        else columns(std::false_type{}, std::true_type{});
    } else if (it + 1 < iters) {
        columns(std::true_type{}, std::true_type{});
    } else {
        columns(std::true_type{}, std::false_type{});
    }
    // z slots: waves 0 .. nzs-1 store their partial, later waves add theirs in rounds of nzs
#pragma unroll
    for (int k = 0; k * nzs < WAVES; ++k) {
        if (w / nzs == k && w < nw) {
            float *zp = zs + (w - k * nzs) * Lg;
#pragma unroll
            for (int b = 0; b < NBT; ++b) {
                const int s = 64 * b + 4 * c + a;
                if (b < nb && s < nslot) {
                    float4 *q4 = reinterpret_cast<float4 *>(zp + 4 * s);
                    const float4 zv = make_float4(zacc[b][0], zacc[b][1], zacc[b][2], zacc[b][3]);
                    if (k == 0) {
                        *q4 = zv;
                    } else {
                        const float4 o = *q4;
                        *q4 = make_float4(o.x + zv.x, o.y + zv.y, o.z + zv.z, o.w + zv.w);
                    }
                }
            }
        }
        __syncthreads();
    }

    // z = the slots in order; u = Y1 - y1' with y1' = y1 + DX - Y1 (train.jl:12)
    const int nzu = nw < nzs ? nw : nzs;
    for (int l = tid; l < Lg; l += kThr) {
        float z = zs[l];
        for (int k = 1; k < nzu; ++k) z += zs[k * Lg + l];
        const float Y1o = Y1s[l];
        const float y1n = y1s[l] + DXs[l] - Y1o;
        y1s[l] = y1n;
        fa[l] = make_float2(z, Y1o - y1n);
    }
    __syncthreads();
    float2 *C = team_fft<false>(fa, fb, P.plan, tw, tid, kThr, BlockSync{});
    float2 *O = (C == fa) ? fb : fa;
    const float c0 = P.sh_csc;
    for (int k = tid; k < Lg; k += kThr) {
        const float2 Ck = C[k], Cm = C[k == 0 ? 0 : Lg - k];
        const float2 Z = make_float2(0.5f * (Ck.x + Cm.x), 0.5f * (Ck.y - Cm.y));   // F(z)
        const float2 U = make_float2(0.5f * (Ck.y + Cm.y), -0.5f * (Ck.x - Cm.x));  // F(u)
        const float ss = gp(P.sS)[k];
        const float den = c0 * gp(P.cS)[k];
        const float2 hh = make_float2((c0 * U.x - Z.x) / den, (c0 * U.y - Z.y) / den);
        const float2 dx = make_float2(Z.x * inv_c0 + ss * hh.x, Z.y * inv_c0 + ss * hh.y);
        O[k] = make_float2(hh.x - dx.y, hh.y + dx.x);   // h_hat + i DX_hat
    }
    __syncthreads();
    float2 *Rr = team_fft<true>(O, C, P.plan, tw, tid, kThr, BlockSync{});
    const float scale = 1.0f / (float)Lg;
    const float rho = P.rho;
    const bool last = it + 1 == iters;
    for (int l = tid; l < Lg; l += kThr) {
        const float hn = Rr[l].x * scale;
        const float DXn = Rr[l].y * scale;
        const bool win = (l >= P.gamma) && (l < P.gamma + P.L4);
        const float wl = win ? 1.f : 0.f;
        const float y1n = y1s[l];
        const float Y1n = (bs[l] + rho * (DXn + y1n)) / (wl + rho);   // update_Y1! (csc_kernels.jl:91)
        for (int t = l; t < vl; t += Lg) hb[csctz_pad(t)] = hn;      // h and its wrapped tail
        DXs[l] = DXn;
        Y1s[l] = Y1n;
        if (last) {
            gp(P.h)[ro + l] = hn;
            gp(P.DX)[ro + l] = DXn;
            gp(P.y1)[ro + l] = y1n;
            gp(P.Y1)[ro + l] = Y1n;
        }
    }
  }
}

// ---------------------------------------------------------------------------- best / objective
// Per-sequence partials of objective() (CDL_setup.jl:286): sum (DX_window - s)^2, sum |X|,
// and the positive-X statistics of update_best!/wrap.jl:12; writes the candidate
// X_best slice X[gamma : 4 : gamma+4L-1] (store.jl:15) and optionally X itself.  X is
// written back materialised (Xt = X, h = 0) unless it already is.
template <int TAPS>
__global__ __launch_bounds__(kThreads) void k_best(const CdlDev *__restrict__ tab, const int *__restrict__ starts, int np,
                                                   float *Xfull) {
    extern __shared__ float4 smem4[];
    __shared__ double red[kThreads];
    int n;
    const CdlDev &P = tab[find_problem(starts, np, blockIdx.x, n)];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int Lg = P.Lg, M = P.M, nslot = Lg >> 2, L = P.L;
    float *hb = reinterpret_cast<float *>(smem4);
    const float4 *hb4 = smem4;
    const size_t ro = (size_t)n * Lg;
    load_h<TAPS>(P, ro, hb, tid, kThreads);
    __syncthreads();
    const float *G1 = g1cur_of(P);
    float *xc = P.xs[(P.sc->best_idx & 1) ^ 1];
    const bool mat = !P.sc->xmat;   // read before any block of this launch sets it
    const int s0 = P.gamma >> 2;
    double ab = 0.0, cnt = 0.0, sx = 0.0, sxx = 0.0;
    for (int m = w; m < M; m += kWaves) {
        const int mm = __builtin_amdgcn_readfirstlane(m);
        const size_t col = ((size_t)n * M + mm) * Lg;
        float g[TAPS];
        load_taps<TAPS>(G1, mm, Lg, g);
        float abf = 0.f;
        for (int sb = lane; sb < nslot; sb += 4 * 64) {   // four slots' loads in flight, then use them in order
            float4 xv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (sb + 64 * i < nslot) xv[i] = ld4(P.Xt + col, sb + 64 * i);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int s = sb + 64 * i;
                if (s >= nslot) break;
                float X[4];
                recon4<TAPS>(xv[i], hb4, s, g, X);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    abf += fabsf(X[q]);
                    if (X[q] > 0.f) { cnt += 1.0; sx += X[q]; sxx += (double)X[q] * X[q]; }
                }
                if (Xfull) st4(Xfull + col, s, mk4(X));
                if (mat) st4(P.Xt + col, s, mk4(X));
                if (s >= s0 && s < s0 + L) gp(xc)[((size_t)n * M + mm) * L + (s - s0)] = X[0];
            }
        }
        ab += abf;
    }
    if (mat)
        for (int l = tid; l < Lg; l += kThreads) gp(P.h)[ro + l] = 0.f;
    double fid = 0.0;
    for (int i = tid; i < P.L4; i += kThreads) {
        const float r = gp(P.DX)[ro + P.gamma + i] - gp(P.s)[(size_t)n * P.L4 + i];
        fid += (double)r * r;
    }
    double vals[5] = {fid, ab, cnt, sx, sxx};
    for (int f = 0; f < 5; ++f) {
        red[tid] = vals[f];
        __syncthreads();
        for (int st = kThreads / 2; st > 0; st >>= 1) {
            if (tid < st) red[tid] += red[tid + st];
            __syncthreads();
        }
        if (tid == 0) gp(P.partials)[(size_t)(P.n_off + n) * kPartStride + f] = red[0];
        __syncthreads();
    }
}

// X is held materialised (what k_reduce_partials records after k_best), for a chunk of a
// streamed problem whose k_best ran but whose partials are not all in yet; one thread per problem.
__global__ void k_mark_materialised(const CdlDev *__restrict__ tab, int np) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < np) tab[p].sc->xmat = 1;
}

// Fixed-order sum of the per-sequence partials; one block per problem.
__global__ __launch_bounds__(kThreads) void k_reduce_partials(const CdlDev *__restrict__ tab) {
    __shared__ double red[kThreads];
    const CdlDev &P = tab[blockIdx.x];
    const int tid = threadIdx.x;
    for (int f = 0; f < 5; ++f) {
        double a = 0.0;
        for (int n = tid; n < P.Nt; n += kThreads) a += gp(P.partials)[(size_t)n * kPartStride + f];
        red[tid] = a;
        __syncthreads();
        for (int st = kThreads / 2; st > 0; st >>= 1) {
            if (tid < st) red[tid] += red[tid + st];
            __syncthreads();
        }
        if (tid == 0) P.sc->red[f] = red[0];
        __syncthreads();
    }
    if (tid == 0) P.sc->xmat = 1;   // k_best left X materialised
}

// update_best! (train.jl:91-98): obj = 0.5*norm(W DX - s)^2 + lambda*sum|X| with the
// reference's Float32 intermediates (norm and the |X| sum in Float32, 0.5 in Float64);
// obj_best is a Float32 field.  force=1 accepts the candidate (constructor, :233-236);
// decide=0 only records the objective.  One block per problem.
__global__ __launch_bounds__(kThreads) void k_decide(const CdlDev *__restrict__ tab, int force, int decide, int slot) {
    __shared__ int imp;
    const CdlDev &P = tab[blockIdx.x];
    CdlScal *sc = P.sc;
    if (threadIdx.x == 0) {
        const float nrm = (float)sqrt(sc->red[0]);
        const float nrm2 = nrm * nrm;
        const float reg = P.lambda * (float)sc->red[1];
        const double obj = 0.5 * (double)nrm2 + (double)reg;
        int better = decide && (force || (obj < (double)sc->obj_best));
        if (slot >= 0 && slot < P.trace_cap) gp(P.trace)[slot] = obj;
        sc->red[5] = obj;
        if (decide) sc->improved = better;
        if (better) {
            sc->obj_best = (float)obj;
            sc->best_idx ^= 1;
            sc->best_stats[0] = sc->red[2];
            sc->best_stats[1] = sc->red[3];
            sc->best_stats[2] = sc->red[4];
        }
        imp = better;
    }
    __syncthreads();
    if (imp)
        for (int i = threadIdx.x; i < P.Lg * P.M; i += kThreads) gp(P.Dbest)[i] = gp(P.meand)[i];
}

// ---------------------------------------------------------------------------- dict
// One iteration of dict_update! (train.jl:50-72) for one sequence; the means over n
// (:76,:83) follow in k_dict_reduce / k_dict_finalize.
// State.  d, g1 and g3 are held in the transform domain (half spectra, bins 0..Lg/2, as the
// reference keeps dh, g1h, g3h), and after step t they are
//   dh_t  = (G1h_t - A) + (F_t - A) + conj(Y3h_t) kappa_t     (sherman_CDL_combine, :60)
//   g1h_t = (A + dh_t) - G1h_t                                  (:70)
//   g3h_t = g1h_t - d13h_t                                      (g1 - g3 is one array per filter)
// with A = g1h_{t-1}, step t's Y3h and kappa, its filter spectra G1h_t and
// F_t = G3h_t + d13h_{t-1}.  So the state kept is A (the g1h buffer), the half spectra of Y3h
// (y3s), kappa per sequence (kap) and the two filter spectra (G1hp, Fp): dh and g1h_t are never
// written.  The next step rebuilds them per bin while it streams the column anyway, and the
// means over n follow from sum_n A and sum_n conj(Y3h) kappa (k_dict_reduce), so a step reads
// and writes each spectral column once.  Before the first step A = Y3h = kappa = Fp = 0 and
// G1hp = F(D): dh = F(D), g1h = 0.
// Step t, columns m = w, w+W, ... of wave w (W = blockDim/64):
//   * Y3 = Pi4(X + y3) = Pi4(X) (y3 = 0 on the support) has one sample per slot, so Y3h is a
//     length-Lg/4 DFT repeated 4 times: one wave-level complex transform per pair of columns
//     (m, m + W), whose real and imaginary inputs are the two columns' real Y3;
//   * per bin: dh_{t-1} and g1h_{t-1} from (A, Y3h_{t-1}, kappa_{t-1}); Qh = (G1h - g1h_{t-1})
//     + (G3h - g3h_{t-1}) (cdl_kernels.jl:8); |Y3h|^2, Y3h dh_{t-1} and Y3h Qh add into the
//     wave's registers in ascending m (:52-59); g1h_{t-1} replaces A, Y3h_t replaces Y3h_{t-1};
//   * the wave partials are summed in wave order; a block-wide spectral step gives
//     Td = F^-1(sum_m Y3h dh_{t-1}) (:56-57, quirk: the previous d), G2, g2 (update_g2!,
//     :68-71) and kappa_t = (c1 F(G2 - g2) - sum_m Y3h Qh) / (|Y3h|^2 + c1), stored.
// Unless X is held materialised already it is written back materialised (Xt = X, h = 0),
// since G1 changes after this step.  The partial-sum order depends on W, which belongs to the
// problem's launch choices (a problem runs with the same W alone and in a batch).  A lane owns
// the float4 slots t = lane + 64 i of a spectral column (bins 2t, 2t+1) and s = lane + 64 i
// of a spatial one.
// LDS: twiddles | h with wrapped tail | accS, accT, accP | kappa_{t-1} | fa, fb (Lg complex,
// the block transforms) | per wave: ya, yb (LgMax/4 + 2 complex each: after the transform they
// hold both columns' Y3h_t and the second column's Y3h_{t-1}), yo (the first column's Y3h_{t-1}).
// Registers: at most 168 (3 waves per SIMD) for padded lengths up to 1020 and filters up to
// 24; 256 otherwise.
template <int TAPS, int NSL>
__global__ __launch_bounds__(512, (NSL <= 4 && TAPS <= 24 ? 3 : 2)) void k_dict(
    const CdlDev *__restrict__ tab, const int *__restrict__ starts, int np, int LgMax) {
    constexpr int NQ = NSL <= 4 ? 4 : 9;        // spectral float4 slots per lane: Hs/2 <= 64 NQ
    constexpr int NY = NSL <= 4 ? 2 : 5;        // stored Y3h bins per lane: Lg/8 + 1 <= 64 NY
    extern __shared__ float4 smem4[];
    int n;
    const CdlDev &P = tab[find_problem(starts, np, blockIdx.x, n)];
    float *sm = reinterpret_cast<float *>(smem4);
    const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int THR = blockDim.x, W = THR >> 6;
    const int Lg = P.Lg, M = P.M, Pq = Lg >> 2, Ph = Pq / 2 + 1, H = Lg / 2 + 1, Hs = half_bins(Lg), nq = Hs >> 1;
    const int HsX = ((LgMax / 2 + 1) + 3) & ~3, PqX = LgMax >> 2, PhX = dict_phx(LgMax);
    float2 *tw = reinterpret_cast<float2 *>(sm);
    float *hb = sm + 2 * LgMax;
    float *accS = hb + ((LgMax + TAPS + 4 + 3) & ~3);
    float2 *accT = reinterpret_cast<float2 *>(accS + HsX);
    float2 *accP = accT + HsX;
    float2 *kold = accP + HsX;
    float2 *fa = kold + HsX, *fb = fa + LgMax;
    float2 *ya = fb + LgMax + w * (2 * (PqX + 2) + PhX), *yb = ya + PqX + 2, *yo = yb + PqX + 2;
    const size_t ro = (size_t)n * Lg;
    float2 *y3s = P.y3s + (size_t)n * M * Ph;
    float2 *kap = P.kap + (size_t)n * Hs;
    const float *G1 = g1cur_of(P);
    const float *G1h = reinterpret_cast<const float *>(P.G1hb[P.sc->g1cur & 1]);
    const float *G3h = reinterpret_cast<const float *>(P.G3h);
    const float *D13h = reinterpret_cast<const float *>(P.d13h);
    const float *G1hp = reinterpret_cast<const float *>(P.G1hp);
    const float *Fp = reinterpret_cast<const float *>(P.Fp);
    const bool mat = !P.sc->xmat;   // set again only by k_cs after this step
    const float4 *hb4 = reinterpret_cast<const float4 *>(hb);

    for (int l = tid; l < Lg; l += THR) tw[l] = gp(P.tw)[l];
    for (int k = tid; k < H; k += THR) kold[k] = gp(kap)[k];
    if (mat) load_h<TAPS>(P, ro, hb, tid, THR);
    __syncthreads();

    // Y3h of bin k from the half spectrum in buffer Y
    auto y3h = [&](const float2 *Y, int k) {
        int kq = k;
        while (kq >= Pq) kq -= Pq;
        return kq < Ph ? Y[kq] : cconj(Y[Pq - kq]);
    };

    float aS[2 * NQ];
    float2 aT[2 * NQ], aP[2 * NQ];
#pragma unroll
    for (int i = 0; i < 2 * NQ; ++i) { aS[i] = 0.f; aT[i] = make_float2(0.f, 0.f); aP[i] = make_float2(0.f, 0.f); }
    // the support samples X[4s] of a column (all the transform needs once X is materialised)
    float xa[NSL], xb[NSL];
    auto loadx0 = [&](int m, float *x) {
        const size_t col = ((size_t)n * M + m) * Lg;
#pragma unroll
        for (int i = 0; i < NSL; ++i) {
            const int s = lane + 64 * i;
            x[i] = (m < M && s < Pq) ? gp(P.Xt)[col + 4 * (size_t)s] : 0.f;
        }
    };
    // X of column m rebuilt from (Xt, h) and written back materialised; its support samples go to
    // part c (0: real, 1: imaginary) of the transform input
    auto rebuild = [&](int m, int c) {
        const size_t col = ((size_t)n * M + m) * Lg;
        float g[TAPS];
        load_taps<TAPS>(G1, m, Lg, g);
        float *yf = reinterpret_cast<float *>(ya) + c;
#pragma unroll
        for (int i = 0; i < NSL; ++i) {
            const int s = lane + 64 * i;
            if (s < Pq) {
                float X[4];
                recon4<TAPS>(ld4(P.Xt + col, s), hb4, s, g, X);
                st4(P.Xt + col, s, mk4(X));
                yf[2 * s] = X[0];
            }
        }
    };
    // The per-bin step of column m: its A (in av, or loaded here when av is null), the stored
    // Y3h_{t-1} (Yo) and the new Y3h (Yn)
    auto column_bins = [&](int m, const float4 *av, const float2 *Yo, const float2 *Yn) {
        const size_t colh = ((size_t)n * M + m) * Hs, fcol = (size_t)m * Hs;
        float *Acol = reinterpret_cast<float *>(P.g1h + colh);
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            const int t = lane + 64 * i;
            if (t < nq) {
                const float4 A = av ? av[i] : ld4(Acol, t);
                const float4 P1 = ld4(G1hp + 2 * fcol, t), PF = ld4(Fp + 2 * fcol, t);
                const float4 F1 = ld4(G1h + 2 * fcol, t), F3 = ld4(G3h + 2 * fcol, t), DL = ld4(D13h + 2 * fcol, t);
                float an[4];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int k = 2 * t + h, j = 2 * i + h;
                    float2 Ac = make_float2(0.f, 0.f);
                    if (k < H) {
                        const float2 a = make_float2(f4(A, 2 * h), f4(A, 2 * h + 1));
                        const float2 p1 = make_float2(f4(P1, 2 * h), f4(P1, 2 * h + 1));
                        const float2 pf = make_float2(f4(PF, 2 * h), f4(PF, 2 * h + 1));
                        const float2 co = cmul(cconj(y3h(Yo, k)), kold[k]);
                        const float2 dho = make_float2(((p1.x - a.x) + (pf.x - a.x)) + co.x,
                                                       ((p1.y - a.y) + (pf.y - a.y)) + co.y);   // dh_{t-1}
                        Ac = make_float2((a.x + dho.x) - p1.x, (a.y + dho.y) - p1.y);            // g1h_{t-1}
                        const float2 a3 = make_float2(Ac.x - f4(DL, 2 * h), Ac.y - f4(DL, 2 * h + 1));   // g3h_{t-1}
                        const float2 Q = make_float2((f4(F1, 2 * h) - Ac.x) + (f4(F3, 2 * h) - a3.x),
                                                     (f4(F1, 2 * h + 1) - Ac.y) + (f4(F3, 2 * h + 1) - a3.y));
                        const float2 Yk = y3h(Yn, k);
                        aS[j] += Yk.x * Yk.x + Yk.y * Yk.y;
                        aT[j] = cadd(aT[j], cmul(Yk, dho));
                        aP[j] = cadd(aP[j], cmul(Yk, Q));
                    }
                    an[2 * h] = Ac.x;
                    an[2 * h + 1] = Ac.y;
                }
                st4(Acol, t, mk4(an));
            }
        }
    };
    auto loadA = [&](int m, float4 *av) {
        const float *Acol = reinterpret_cast<const float *>(P.g1h + ((size_t)n * M + m) * Hs);
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            const int t = lane + 64 * i;
            if (t < nq) av[i] = ld4(Acol, t);
        }
    };
    // Columns in pairs (m, m + W) of the wave's columns w, w + W, ... (one at a time at NSL = 8):
    // one length-Lg/4 complex transform of Y3_m + i Y3_m2 gives both columns' spectra (both
    // inputs are real: Y3h_m(k) = (Z(k) + conj Z(-k)) / 2, Y3h_m2(k) = (Z(k) - conj Z(-k)) / 2i);
    // the per-bin steps then run for m and for m2, so the sums over m keep their ascending order.
    if constexpr (NSL <= 4) {
    if (!mat && w < M) { loadx0(w, xa); loadx0(w + W, xb); }
    for (int m = w; m < M; m += 2 * W) {
        const int mm = __builtin_amdgcn_readfirstlane(m);
        const int m2 = mm + W;
        const bool two = m2 < M;
        if (mat) {
            rebuild(mm, 0);
            if (two) {
                rebuild(m2, 1);
            } else {
                for (int s = lane; s < Pq; s += 64) ya[s].y = 0.f;
            }
        } else {
#pragma unroll
            for (int i = 0; i < NSL; ++i) {
                const int s = lane + 64 * i;
                if (s < Pq) ya[s] = make_float2(xa[i], xb[i]);
            }
        }
        // the stored Y3h_{t-1} of both columns and the first column's A, in flight during the transform
        float4 av[NQ];
        float2 yv[NY], yv2[NY];
        loadA(mm, av);
#pragma unroll
        for (int i = 0; i < NY; ++i) {
            const int k = lane + 64 * i;
            if (k < Ph) {
                yv[i] = gp(y3s)[(size_t)mm * Ph + k];
                yv2[i] = two ? gp(y3s)[(size_t)m2 * Ph + k] : make_float2(0.f, 0.f);
            }
        }
        wave_lds_sync();
        const float2 *Z = team_fft<false>(ya, yb, P.plan4, tw, lane, 64, WaveSync{}, 1);
        float2 *Zw = (Z == ya) ? ya : yb;
        float2 *O = (Z == ya) ? yb : ya;
        float2 y1n[NY], y2n[NY];
#pragma unroll
        for (int i = 0; i < NY; ++i) {
            const int k = lane + 64 * i;
            if (k < Ph) {
                const float2 a = Z[k], b = cconj(Z[k == 0 ? 0 : Pq - k]);
                y1n[i] = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y + b.y));
                y2n[i] = make_float2(0.5f * (a.y - b.y), -0.5f * (a.x - b.x));
            }
        }
        wave_lds_sync();   // every lane has read Z before it is overwritten
        // LDS: Y3h_t of m in O[0, Ph), of m2 in Zw[0, Ph); Y3h_{t-1} of m in yo, of m2 in O[Ph, 2 Ph)
#pragma unroll
        for (int i = 0; i < NY; ++i) {
            const int k = lane + 64 * i;
            if (k < Ph) {
                O[k] = y1n[i];
                Zw[k] = y2n[i];
                yo[k] = yv[i];
                O[Ph + k] = yv2[i];
                gp(y3s)[(size_t)mm * Ph + k] = y1n[i];
                if (two) gp(y3s)[(size_t)m2 * Ph + k] = y2n[i];
            }
        }
        wave_lds_sync();
        // the next pair's support samples, in flight during this pair's per-bin steps
        if (!mat && mm + 2 * W < M) { loadx0(mm + 2 * W, xa); loadx0(mm + 3 * W, xb); }
        column_bins(mm, av, yo, O);
        if (two) {
            loadA(m2, av);
            column_bins(m2, av, O + Ph, Zw);
        }
        wave_lds_sync();
    }
    } else {
    // NSL = 8: one column per transform (its per-bin step's registers leave no room for a second)
    float4 nx[NSL];
    auto loadx = [&](int m) {
        const size_t col = ((size_t)n * M + m) * Lg;
#pragma unroll
        for (int i = 0; i < NSL; ++i) {
            const int s = lane + 64 * i;
            if (s < Pq) nx[i] = ld4(P.Xt + col, s);
        }
    };
    if (w < M) loadx(w);
    for (int m = w; m < M; m += W) {
        const int mm = __builtin_amdgcn_readfirstlane(m);
        const size_t col = ((size_t)n * M + mm) * Lg;
        float2 yv[NY];
#pragma unroll
        for (int i = 0; i < NY; ++i) {
            const int k = lane + 64 * i;
            if (k < Ph) yv[i] = gp(y3s)[(size_t)mm * Ph + k];
        }
        float g[TAPS];
        if (mat) load_taps<TAPS>(G1, mm, Lg, g);
#pragma unroll
        for (int i = 0; i < NSL; ++i) {
            const int s = lane + 64 * i;
            if (s < Pq) {
                float x0 = nx[i].x;
                if (mat) {
                    float X[4];
                    recon4<TAPS>(nx[i], hb4, s, g, X);
                    st4(P.Xt + col, s, mk4(X));
                    x0 = X[0];
                }
                ya[s] = make_float2(x0, 0.f);
            }
        }
        if (mm + W < M) loadx(mm + W);
#pragma unroll
        for (int i = 0; i < NY; ++i) {
            const int k = lane + 64 * i;
            if (k < Ph) yo[k] = yv[i];
        }
        wave_lds_sync();
        const float2 *Y = team_fft<false>(ya, yb, P.plan4, tw, lane, 64, WaveSync{}, 1);
        for (int k = lane; k < Ph; k += 64) gp(y3s)[(size_t)mm * Ph + k] = Y[k];   // the old values are in yo
        column_bins(mm, nullptr, yo, Y);
        wave_lds_sync();
    }
    }
    if (mat)
        for (int l = tid; l < Lg; l += THR) gp(P.h)[ro + l] = 0.f;
    // wave partials summed in wave order
    const int nw = M < W ? M : W;
    for (int ww = 0; ww < nw; ++ww) {
        if (w == ww) {
#pragma unroll
            for (int i = 0; i < NQ; ++i) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int k = 2 * (lane + 64 * i) + h, j = 2 * i + h;
                    if (k < H) {
                        if (ww == 0) { accS[k] = aS[j]; accT[k] = aT[j]; accP[k] = aP[j]; }
                        else { accS[k] += aS[j]; accT[k] = cadd(accT[k], aT[j]); accP[k] = cadd(accP[k], aP[j]); }
                    }
                }
            }
        }
        __syncthreads();
    }
    if (nw == 0) {   // no columns: sums of nothing
        for (int k = tid; k < H; k += THR) { accS[k] = 0.f; accT[k] = make_float2(0.f, 0.f); accP[k] = make_float2(0.f, 0.f); }
        __syncthreads();
    }

    // ---- per-sequence spectral step: Td, G2/g2 (update_g2!, train.jl:68-71), kappa
    float *tdb = reinterpret_cast<float *>(fb + LgMax);    // Lg floats over wave 0's ya, yb
    const float scale = 1.0f / (float)Lg;
    for (int k = tid; k < Lg; k += THR) fa[k] = (k < H) ? accT[k] : cconj(accT[Lg - k]);
    __syncthreads();
    {
        const float2 *R2 = team_fft<true>(fa, fb, P.plan, tw, tid, THR, BlockSync{}, 1);
        for (int l = tid; l < Lg; l += THR) tdb[l] = R2[l].x * scale;
    }
    __syncthreads();
    {
        // F(G2 - g2) of the previous dictionary iteration (sherman_CDL_RHS!, cdl_kernels.jl:8)
        for (int l = tid; l < Lg; l += THR) fa[l] = make_float2(gp(P.G2)[ro + l] - gp(P.g2)[ro + l], 0.f);
        __syncthreads();
        const float2 *R = team_fft<false>(fa, fb, P.plan, tw, tid, THR, BlockSync{}, 1);
        const float c1 = P.sh_cdl;
        for (int k = tid; k < Hs; k += THR) {
            float2 kp = make_float2(0.f, 0.f);
            if (k < H) {
                const float2 Gk = R[k];
                const float den = accS[k] + c1;
                kp = make_float2((c1 * Gk.x - accP[k].x) / den, (c1 * Gk.y - accP[k].y) / den);
            }
            gp(kap)[k] = kp;
        }
    }
    for (int l = tid; l < Lg; l += THR) {
        const float Td = tdb[l];
        const bool win = (l >= P.gamma) && (l < P.gamma + P.L4);
        const float b = win ? gp(P.s)[(size_t)n * P.L4 + (l - P.gamma)] : 0.f;
        const float wl = win ? 1.f : 0.f;
        const float g2o = gp(P.g2)[ro + l];
        const float G2n = (b + P.sigma * (Td + g2o)) / (wl + P.sigma);
        gp(P.G2)[ro + l] = G2n;
        gp(P.g2)[ro + l] = g2o + Td - G2n;
        gp(P.Td)[ro + l] = Td;
    }
}

// Partial sums over a chunk of sequences, per (bin, m), of A = g1h_{t-1} and of
// c = conj(Y3h_t) kappa_t (the same float product the next step forms), in double, real and
// imaginary parts: the means over n of d, d + g1 and d + g3 (train.jl:76,83,95) follow from
// them (k_dict_finalize).  Blocks of a problem: (tile, chunk), mapped so that all tiles of a
// chunk share the block index modulo 8, i.e. one XCD: every tile of a chunk reads the same
// Y3 half spectra and kappa of its sequences, which then come from HBM once into that XCD's L2
// (the chunk count is a multiple of 8).
__global__ __launch_bounds__(kThreads) void k_dict_reduce(const CdlDev *__restrict__ tab, const int *__restrict__ starts,
                                                          int np) {
    int b;
    const CdlDev &P = tab[find_problem(starts, np, blockIdx.x, b)];
    const int Lg = P.Lg, Hs = half_bins(Lg), H = Lg / 2 + 1, M = P.M, HM = Hs * M;
    const int Pq = Lg >> 2, Ph = Pq / 2 + 1;
    const int tiles = (HM + kThreads - 1) / kThreads;
    int c, tile;
    if ((P.red_cn & 7) == 0) {
        const int g = b & 7, j = b >> 3;                  // XCD group, index within it
        c = g + 8 * (j / tiles);
        tile = j % tiles;
    } else {
        c = b / tiles;
        tile = b - c * tiles;
    }
    c += P.red_c0;
    const int i = tile * kThreads + threadIdx.x;
    if (i >= HM) return;
    const int m = i / Hs, k = i - m * Hs;
    int kq = k;
    while (kq >= Pq) kq -= Pq;
    const bool cj = kq >= Ph;
    const int yi = cj ? Pq - kq : kq;
    const int Nt = P.Nt, nch = P.nch;
    // the chunk's sequences, as indices into the sequences this launch sees
    const int n0 = (int)((long)c * Nt / nch) - P.n_off, n1 = (int)((long)(c + 1) * Nt / nch) - P.n_off;
    double ax = 0.0, ay = 0.0, cx = 0.0, cy = 0.0;
    if (k < H) {
        for (int n = n0; n < n1; ++n) {
            const float2 a = gp(P.g1h)[(size_t)n * HM + i];
            float2 y = gp(P.y3s)[((size_t)n * M + m) * Ph + yi];
            if (cj) y.y = -y.y;
            const float2 cc = cmul(cconj(y), gp(P.kap)[(size_t)n * Hs + k]);
            ax += (double)a.x;
            ay += (double)a.y;
            cx += (double)cc.x;
            cy += (double)cc.y;
        }
    }
    double *o = gp(P.dpart) + (size_t)c * kDpart * HM + i;
    o[0 * (size_t)HM] = ax;
    o[1 * (size_t)HM] = ay;
    o[2 * (size_t)HM] = cx;
    o[3 * (size_t)HM] = cy;
}

// Column m: the mean spectra of (d+g1), (d+g3) and d back to the spatial domain (one packed
// inverse transform for the first two, one for the third); v1 = mean(d+g1), v2 = mean(d+g3);
// G3 = projection2box!(v2) (cdl_kernels.jl:37-51); G1new = prox_simplex_zeros!(v1) (:53-66,
// FP64 combination, quirk group 4*floor(l/4)) into the non-current G1 buffer; meand =
// mean(d) (D_best candidate, train.jl:95); gsq = |F(G1new)|^2 and the half spectra of G1new
// and G3.  This step's filter spectra are kept as G1hp and Fp = G3h + d13h (the dictionary
// state refers to them) before they are replaced.
__global__ __launch_bounds__(kThreads) void k_dict_finalize(const CdlDev *__restrict__ tab, const int *__restrict__ starts,
                                                            int np, int LgMax) {
    extern __shared__ float4 smem4[];
    int m;
    const CdlDev &P = tab[find_problem(starts, np, blockIdx.x, m)];
    float *v1 = reinterpret_cast<float *>(smem4);
    float2 *tw = reinterpret_cast<float2 *>(v1 + ((LgMax + 3) & ~3));
    float2 *fa = tw + LgMax;
    float2 *fb = fa + LgMax;
    const int tid = threadIdx.x, Lg = P.Lg, H = Lg / 2 + 1, Hs = half_bins(Lg);
    const int HM = Hs * P.M;
    const int flm = gp(P.fl)[m];
    const double invN = 1.0 / (double)P.Nt;
    const float scale = 1.0f / (float)Lg;
    const int nxt = (P.sc->g1cur & 1) ^ 1;
    float *G1new = P.G1b[nxt];
    float2 *G1hnew = P.G1hb[nxt];
    // means over n at bin k < H of d (j = 1) or d + g1 (j = 0), with this step's G1h = f1 and
    // F = G3h + d13h (old): from S_A = mean(A), S_c = mean(conj(Y3h) kappa),
    //   mean(dh) = G1h + F - 2 S_A + S_c,   mean(g1h_new) = F - S_A + S_c
    auto mean = [&](int j, int k, float2 f1, float2 F, float2 sub) {
        const size_t i = (size_t)m * Hs + k;
        double s[4] = {0.0, 0.0, 0.0, 0.0};
        for (int c = 0; c < P.nch; ++c)
            for (int q = 0; q < 4; ++q) s[q] += gp(P.dpart)[((size_t)c * kDpart + q) * HM + i];
        const double ax = s[0] * invN, ay = s[1] * invN, cx = s[2] * invN, cy = s[3] * invN;
        double dx = (double)f1.x + (double)F.x - 2.0 * ax + cx, dy = (double)f1.y + (double)F.y - 2.0 * ay + cy;
        if (j == 0) {
            dx += (double)F.x - ax + cx;
            dy += (double)F.y - ay + cy;
        }
        return make_float2((float)(dx - (double)sub.x), (float)(dy - (double)sub.y));
    };
    for (int l = tid; l < Lg; l += kThreads) tw[l] = gp(P.tw)[l];
    // v1 + i v2 = F^-1(V1 + i V2) for Hermitian V1, V2: the means of d + g1 and of
    // d + g3 = (d + g1) - (g1 - g3), with g1, g3 after this step's updates (train.jl:70,72
    // precede the means at :76,:83): g1 - g3 first advances by (dh - G1h) - (dh - G3h) =
    // G3h - G1h with this step's filters, read before G3h is replaced below.
    const float2 *G1hcur = P.G1hb[nxt ^ 1];
    const float2 zero2 = make_float2(0.f, 0.f);
    for (int k = tid; k < Hs; k += kThreads) {
        const size_t i = (size_t)m * Hs + k;
        if (k >= H) {
            gp(P.G1hp)[i] = zero2;
            gp(P.Fp)[i] = zero2;
            continue;
        }
        const float2 dl = gp(P.d13h)[i], f1 = gp(G1hcur)[i], f3 = gp(P.G3h)[i];
        const float2 F = make_float2(f3.x + dl.x, f3.y + dl.y);
        gp(P.G1hp)[i] = f1;
        gp(P.Fp)[i] = F;
        const float2 dn = make_float2(dl.x - (f1.x - f3.x), dl.y - (f1.y - f3.y));
        gp(P.d13h)[i] = dn;
        const float2 a = mean(0, k, f1, F, zero2), b = mean(0, k, f1, F, dn);
        fa[k] = make_float2(a.x - b.y, a.y + b.x);
        if (k > 0 && Lg - k >= H) fa[Lg - k] = make_float2(a.x + b.y, b.x - a.y);
    }
    __syncthreads();
    {
        const float2 *R = team_fft<true>(fa, fb, P.plan, tw, tid, kThreads, BlockSync{});
        for (int l = tid; l < Lg; l += kThreads) {
            v1[l] = R[l].x * scale;
            const float v2 = R[l].y * scale;
            const bool sup = l < flm;
            float g3;
            if (v2 >= 1.f) g3 = sup ? 1.f : 0.f;
            else if (v2 <= 0.f) g3 = 0.f;
            else g3 = sup ? v2 : 0.f;
            gp(P.G3)[(size_t)m * Lg + l] = g3;
        }
    }
    __syncthreads();
    for (int k = tid; k < H; k += kThreads) {
        const size_t i = (size_t)m * Hs + k;
        const float2 a = mean(1, k, gp(P.G1hp)[i], gp(P.Fp)[i], zero2);
        fa[k] = a;
        if (k > 0 && Lg - k >= H) fa[Lg - k] = cconj(a);
    }
    __syncthreads();
    {
        const float2 *R = team_fft<true>(fa, fb, P.plan, tw, tid, kThreads, BlockSync{});
        for (int l = tid; l < Lg; l += kThreads) gp(P.meand)[(size_t)m * Lg + l] = R[l].x * scale;
    }
    __syncthreads();
    for (int l = tid; l < Lg; l += kThreads) {
        float g = 0.f;
        if (l < flm) {
            const int gi = 4 * ((l + 1) / 4);
            const float s4 = v1[gi] + v1[gi + 1] + v1[gi + 2] + v1[gi + 3] - 1.0f;
            g = (float)((double)v1[l] - 0.25 * (double)s4);
        }
        gp(G1new)[(size_t)m * Lg + l] = g;
        fa[l] = make_float2(g, 0.f);
    }
    __syncthreads();
    {
        const float2 *C = team_fft<false>(fa, fb, P.plan, tw, tid, kThreads, BlockSync{});
        for (int k = tid; k < Lg; k += kThreads) gp(P.gsq)[(size_t)m * Lg + k] = C[k].x * C[k].x + C[k].y * C[k].y;
        for (int k = tid; k < Hs; k += kThreads) gp(G1hnew)[(size_t)m * Hs + k] = k < H ? C[k] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    for (int l = tid; l < Lg; l += kThreads) fa[l] = make_float2(gp(P.G3)[(size_t)m * Lg + l], 0.f);
    __syncthreads();
    {
        const float2 *C = team_fft<false>(fa, fb, P.plan, tw, tid, kThreads, BlockSync{});
        for (int k = tid; k < Hs; k += kThreads) gp(P.G3h)[(size_t)m * Hs + k] = k < H ? C[k] : make_float2(0.f, 0.f);
    }
}

// |F(gp(G1)[:, m])|^2 and the half spectra of the current G1 and of G3 (constructor and
// reset).
__global__ __launch_bounds__(kThreads) void k_g1_spectrum(const CdlDev *__restrict__ tab, const int *__restrict__ starts,
                                                          int np, int LgMax) {
    extern __shared__ float4 smem4[];
    int m;
    const CdlDev &P = tab[find_problem(starts, np, blockIdx.x, m)];
    float2 *tw = reinterpret_cast<float2 *>(smem4);
    float2 *fa = tw + LgMax;
    float2 *fb = fa + LgMax;
    const int tid = threadIdx.x, Lg = P.Lg, H = Lg / 2 + 1, Hs = half_bins(Lg);
    const int cur = P.sc->g1cur & 1;
    const float *G1 = P.G1b[cur];
    for (int l = tid; l < Lg; l += kThreads) {
        tw[l] = gp(P.tw)[l];
        fa[l] = make_float2(gp(G1)[(size_t)m * Lg + l], 0.f);
    }
    __syncthreads();
    {
        const float2 *C = team_fft<false>(fa, fb, P.plan, tw, tid, kThreads, BlockSync{});
        for (int k = tid; k < Lg; k += kThreads) gp(P.gsq)[(size_t)m * Lg + k] = C[k].x * C[k].x + C[k].y * C[k].y;
        for (int k = tid; k < Hs; k += kThreads) gp(P.G1hb[cur])[(size_t)m * Hs + k] = k < H ? C[k] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    for (int l = tid; l < Lg; l += kThreads) fa[l] = make_float2(gp(P.G3)[(size_t)m * Lg + l], 0.f);
    __syncthreads();
    {
        const float2 *C = team_fft<false>(fa, fb, P.plan, tw, tid, kThreads, BlockSync{});
        for (int k = tid; k < Hs; k += kThreads) gp(P.G3h)[(size_t)m * Hs + k] = k < H ? C[k] : make_float2(0.f, 0.f);
    }
}

// One column (n, m) of d (what = 0), g1 (1) or g3 (2) in the spatial domain (field reads):
// their half spectra from the dictionary state (dh_t = (G1hp - A) + (Fp - A) + conj(Y3h) kappa,
// g1h_t = (A + dh_t) - G1hp, g3h_t = g1h_t - d13h, as k_dict rebuilds them), then the inverse
// transform of the Hermitian extension.  out[(n M + m) Lg + l].  LDS: twiddles | a | b.
__global__ __launch_bounds__(kThreads) void k_dict_to_spatial(const CdlDev *__restrict__ tab, int what,
                                                              float *__restrict__ out) {
    extern __shared__ float4 smem4[];
    const CdlDev &P = tab[0];
    const int Lg = P.Lg, H = Lg / 2 + 1, Hs = half_bins(Lg), tid = threadIdx.x;
    const int Pq = Lg >> 2, Ph = Pq / 2 + 1;
    const size_t c = blockIdx.x, n = c / (size_t)P.M, m = c % (size_t)P.M;
    float2 *tw = reinterpret_cast<float2 *>(smem4);
    float2 *fa = tw + Lg;
    float2 *fb = fa + Lg;
    for (int l = tid; l < Lg; l += kThreads) tw[l] = gp(P.tw)[l];
    for (int k = tid; k < H; k += kThreads) {
        int kq = k;
        while (kq >= Pq) kq -= Pq;
        float2 y = gp(P.y3s)[c * Ph + (kq < Ph ? kq : Pq - kq)];
        if (kq >= Ph) y.y = -y.y;
        const float2 co = cmul(cconj(y), gp(P.kap)[n * Hs + k]);
        const float2 a = gp(P.g1h)[c * Hs + k], p1 = gp(P.G1hp)[m * Hs + k], pf = gp(P.Fp)[m * Hs + k];
        const float2 dh = make_float2(((p1.x - a.x) + (pf.x - a.x)) + co.x, ((p1.y - a.y) + (pf.y - a.y)) + co.y);
        float2 v = dh;
        if (what > 0) {
            v = make_float2((a.x + dh.x) - p1.x, (a.y + dh.y) - p1.y);
            if (what == 2) {
                const float2 u = gp(P.d13h)[m * Hs + k];
                v = make_float2(v.x - u.x, v.y - u.y);
            }
        }
        fa[k] = v;
        if (k > 0 && Lg - k >= H) fa[Lg - k] = cconj(v);
    }
    __syncthreads();
    const float2 *R = team_fft<true>(fa, fb, P.plan, tw, tid, kThreads, BlockSync{});
    const float scale = 1.0f / (float)Lg;
    for (int l = tid; l < Lg; l += kThreads) gp(out)[c * Lg + l] = R[l].x * scale;
}

// s_k = sum_m |G1h|^2 (FP32, m ascending; train.jl:18-19), c_k = s_k + sherman_const_csc.
// flip != 0 (after k_dict_finalize): the new filters become current.
__global__ __launch_bounds__(kThreads) void k_cs(const CdlDev *__restrict__ tab, const int *__restrict__ starts, int np,
                                                 int flip) {
    int b;
    const CdlDev &P = tab[find_problem(starts, np, blockIdx.x, b)];
    const int k = b * kThreads + threadIdx.x;
    if (flip && b == 0 && threadIdx.x == 0) {
        P.sc->g1cur ^= 1;
        P.sc->xmat = 1;   // the dictionary step left X materialised
    }
    if (k >= P.Lg) return;
    float a = 0.f;
    for (int m = 0; m < P.M; ++m) a += gp(P.gsq)[(size_t)m * P.Lg + k];
    gp(P.sS)[k] = a;
    gp(P.cS)[k] = a + P.sh_csc;
}

// nb transforms of length plan.n (table length plan.n * plan.twstride) by the team FFT, for
// the FFT self-test.  LDS: twiddles | a | b.
__global__ __launch_bounds__(kThreads) void k_fft_selftest(const float2 *__restrict__ in, float2 *__restrict__ out,
                                                           const float2 *__restrict__ twg, FftPlan plan, int nb, int inv) {
    extern __shared__ float4 smem4[];
    const int n = plan.n, N = n * plan.twstride, tid = threadIdx.x;
    float2 *tw = reinterpret_cast<float2 *>(smem4);
    float2 *a = tw + N;
    float2 *b = a + (size_t)nb * n;
    for (int i = tid; i < N; i += kThreads) tw[i] = twg[i];
    for (int i = tid; i < nb * n; i += kThreads) a[i] = in[i];
    __syncthreads();
    float2 *R = inv ? team_fft<true>(a, b, plan, tw, tid, kThreads, BlockSync{}, nb)
                    : team_fft<false>(a, b, plan, tw, tid, kThreads, BlockSync{}, nb);
    for (int i = tid; i < nb * n; i += kThreads) out[i] = R[i];
}

#define CDLM_TAPS_SWITCH(taps, CALL)                                                       \
    switch (taps) {                                                                        \
    case 4: CALL(4); break;                                                                \
    case 8: CALL(8); break;                                                                \
    case 12: CALL(12); break;                                                              \
    case 16: CALL(16); break;                                                              \
    case 20: CALL(20); break;                                                              \
    case 24: CALL(24); break;                                                              \
    case 28: CALL(28); break;                                                              \
    case 32: CALL(32); break;                                                              \
    default: break;                                                                        \
    }

}  // namespace
