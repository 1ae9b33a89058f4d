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
//   * the dictionary step keeps d spatial; its solve is
//       d_m = q_m + F^-1[conj(Y3h_m) kappa],   kappa = (c1 F(G2-g2) - sum_m Y3h_m Q_m)/(|Y3h|^2 + c1)
//     with q_m = (G1_m - g1_m) + (G3_m - g3_m) (c1 = sherman_const_cdl; no 1/c1, as
//     sherman_CDL_combine).  Y3 = Pi4(X + y3) is nonzero on every 4th sample only, so its
//     spectrum is a length Lg/4 DFT repeated 4 times.
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
// one sequence n and streams its contiguous (Lg x M) slab.  y3 is stored without its
// stride-4 support samples: there y3 + X - Pi4(X + y3) = (y3 + X) - (X + y3) = 0 exactly
// (IEEE addition commutes) from the first iteration on, and y3 starts at 0, so the
// support samples are identically zero and only l % 4 != 0 is kept, 3 floats per slot.
#pragma once

#include "cdlm_internal.h"

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kMaxTaps = 32;
constexpr int kPartStride = 8;  // doubles per sequence in the objective partials

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
    float rho, sigma, lambda, sh_csc, sh_cdl;
    const float *s;
    const int *fl;
    float *Xt, *y2, *d, *g1, *g3;                 // (Lg, M, N)
    float *y3;                                     // (3 Lg/4, M, N): y3 off the stride-4 support
    float *h, *y1, *Y1, *DX, *G2, *g2, *Td;       // (Lg, N)
    float *G1b[2], *G3, *meand, *Dbest, *gsq;     // (Lg, M)
    float *cS, *sS;                                // (Lg)
    float *xs[2];                                  // (L, M, N) X_best slices
    float2 *y3s;                                   // (Lg/8+1, M, N) dictionary scratch
    const float2 *tw;                              // (Lg) twiddles
    double *partials, *dpart, *trace;
    CdlScal *sc;
    const float *X0;       // (Lg, M, N) initial codes, or null: N(0,1) from seed
    const float *D0;       // (Lg, M) masked initial filters
    unsigned long long seed;
    FftPlan plan, plan4;
};

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
#if defined(__HIP_DEVICE_COMPILE__)
#define CDLM_GLOBAL __attribute__((address_space(1)))
#else
#define CDLM_GLOBAL   // the host pass only parses device code
#endif
template <typename T>
__device__ __forceinline__ CDLM_GLOBAL T *gp(T *p) { return (CDLM_GLOBAL T *)p; }
__device__ __forceinline__ const float *g1cur_of(const CdlDev &P) { return P.G1b[P.sc->g1cur & 1]; }

__device__ __forceinline__ float4 ld4(const float *p, int s) { return ((const CDLM_GLOBAL float4 *)p)[s]; }
__device__ __forceinline__ void st4(float *p, int s, float4 v) { ((CDLM_GLOBAL float4 *)p)[s] = v; }
__device__ __forceinline__ float f4(const float4 &v, int q) { return q == 0 ? v.x : q == 1 ? v.y : q == 2 ? v.z : v.w; }
__device__ __forceinline__ float4 mk4(const float *a) { return make_float4(a[0], a[1], a[2], a[3]); }
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
#pragma unroll
    for (int j = 0; j < TAPS; ++j) g[j] = (j < Lg) ? gp(G)[(size_t)m * Lg + j] : 0.f;
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
// spectral product of :201-203 as a circular convolution); d = D; Y1 = y = g = 0; G2 = DX.
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
            const float x = P.X0 ? gp(P.X0)[col + 4 * s] : gauss_of(P.seed, col + 4 * (size_t)s);
            const float4 xp = make_float4(x, 0.f, 0.f, 0.f);
            st4(P.Xt + col, s, xp);
            st4(P.y2 + col, s, zero4);
            st3(P.y3 + col3_of(n, mm, M, nslot), s, reinterpret_cast<const float *>(&zero4));
            st4(P.g1 + col, s, zero4);
            st4(P.g3 + col, s, zero4);
            st4(P.d + col, s, ld4(G1 + (size_t)mm * Lg, s));
            vb4[TAPS / 4 + s] = xp;
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
}

// ---------------------------------------------------------------------------- CSC
// `iters` iterations of code_update! (train.jl:5-43) for one sequence.  The iterations of
// different sequences are independent (the dictionary spectrum is fixed during code_update!),
// so a work-group runs all of its sequence's iterations back to back and h stays in LDS
// between them (written to HBM after the last).  WAVES waves per work-group; wave w owns the
// columns m = w, w + WAVES, ...; a lane owns the float4 slots s = lane + 64 i (i < NSL) of a
// column, the same in every iteration.  With WAVES = 4 (three work-groups per CU) column
// m + WAVES's loads are issued before column m is processed (two register sets), and so
// with WAVES = 8 (one work-group per CU by registers); with WAVES = 16 (one work-group per
// CU, at most 128 VGPRs) there is no prefetch and the 16 waves hide the latency instead.
// With one sequence per CU resident, a sequence's state (Xt, y2, y3: 0.6 MB at Lg = 840,
// M = 64) is largely re-read from the last-level cache in the next iteration rather than
// from HBM.
// The (Lg) vectors y1, Y1, DX and the data window stay in LDS for the whole launch, and the
// next iteration's first column is loaded before the per-sequence transforms.
// LDS (float4 units): twiddles (Lg/2) | h with wrapped tail | WAVES x v column with wrapped
// head | WAVES x z partial | y1, Y1, DX, data window, s_k, c_k (Lg each); the FFT ping-pong buffers
// reuse the v columns after the column loop.
template <int WAVES, int NSL, int TAPS>
struct CscShape {
    static constexpr int kThr = 64 * WAVES;
    static constexpr bool kPrefetch = WAVES <= 8;
    static constexpr int kWavesPerSimd = WAVES <= 4 ? (NSL <= 4 && TAPS <= 20 ? 3 : 2) : WAVES / 4;
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
    float *sSs = bs + Lg;                           // the spectral constants s_k, c_k
    float *cSs = sSs + Lg;
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
        sSs[l] = gp(P.sS)[l];
        cSs[l] = gp(P.cS)[l];
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
    auto load_col = [&](int m, float4 *a, float4 *b, float4 *c) {
        const size_t col = ((size_t)n * M + m) * Lg, col3 = col3_of(n, m, M, nslot);
#pragma unroll
        for (int i = 0; i < NSL; ++i) {
            const int s = lane + 64 * i;
            if (s < nslot) { a[i] = ld4(P.Xt + col, s); b[i] = ld4(P.y2 + col, s); c[i] = ld3(P.y3 + col3, s); }
        }
    };
    if (kPf && w < M) load_col(w, nx, n2, n3);
    const int nw = M < WAVES ? M : WAVES;       // waves that own columns (and z partials)
#pragma nounroll
  for (int it = 0; it < iters; ++it) {
    __syncthreads();
    for (int m = w; m < M; m += WAVES) {
        const int mm = __builtin_amdgcn_readfirstlane(m);
        const size_t col = ((size_t)n * M + mm) * Lg, col3 = col3_of(n, mm, M, nslot);
        float4 xt[kPf ? NSL : 1], a2[kPf ? NSL : 1], a3[kPf ? NSL : 1];
        if constexpr (kPf) {
#pragma unroll
            for (int i = 0; i < NSL; ++i) { xt[i] = nx[i]; a2[i] = n2[i]; a3[i] = n3[i]; }
            if (mm + WAVES < M) load_col(mm + WAVES, nx, n2, n3);
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
                        if (s < nslot) { lx[u] = ld4(P.Xt + col, s); l2[u] = ld4(P.y2 + col, s); l3[u] = ld3(P.y3 + col3, s); }
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < kSg; ++u) {
                const int i = i0 + u;
                const int s = lane + 64 * i;
                if (s < nslot) {
                    const float4 xv = kPf ? xt[kPf ? i : 0] : lx[u];
                    const float4 b2 = kPf ? a2[kPf ? i : 0] : l2[u];
                    const float4 b3 = kPf ? a3[kPf ? i : 0] : l3[u];
                    float X[4];
                    recon4<TAPS>(xv, hb4, s, g, X);
                    float y2n[4], y3n[4], v[4], xn[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const float x = X[q];
                        const float t = x + f4(b2, q);                      // shrink! (csc_kernels.jl:58-59)
                        const float at = fabsf(t) - lam;
                        const float Y2 = at > 0.f ? copysignf(at, t) : 0.f;
                        const float Y3 = (q == 0) ? x : 0.f;                  // projectionX! (:73-77), y3 = 0 there
                        y2n[q] = f4(b2, q) + x - Y2;                          // train.jl:13
                        y3n[q] = (q == 0) ? 0.f : f4(b3, q) + x;              // train.jl:14
                        v[q] = (Y2 - y2n[q]) + (Y3 - y3n[q]);
                        xn[q] = v[q] * inv_c0;
                    }
                    st4(P.y2 + col, s, mk4(y2n));
                    st3(P.y3 + col3, s, y3n);
                    st4(P.Xt + col, s, mk4(xn));
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
    // the next iteration's first column does not depend on what follows: load it now (all of
    // it with prefetch, its first two slots without)
    if (it + 1 < iters && w < M) {
        if constexpr (kPf) {
            load_col(w, nx, n2, n3);
        } else {
            const size_t col = ((size_t)n * M + w) * Lg, col3 = col3_of(n, w, M, nslot);
#pragma unroll
            for (int u = 0; u < kSg; ++u) {
                const int s = lane + 64 * u;
                if (s < nslot) { nx[u] = ld4(P.Xt + col, s); n2[u] = ld4(P.y2 + col, s); n3[u] = ld3(P.y3 + col3, s); }
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
        const float ss = sSs[k];
        const float den = c0 * cSs[k];
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
        for (int s = lane; s < nslot; s += 64) {
            float X[4];
            recon4<TAPS>(ld4(P.Xt + col, s), hb4, s, g, X);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                abf += fabsf(X[q]);
                if (X[q] > 0.f) { cnt += 1.0; sx += X[q]; sxx += (double)X[q] * X[q]; }
            }
            if (Xfull) st4(Xfull + col, s, mk4(X));
            if (mat) st4(P.Xt + col, s, mk4(X));
            if (s >= s0 && s < s0 + L) gp(xc)[((size_t)n * M + mm) * L + (s - s0)] = X[0];
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
        if (tid == 0) gp(P.partials)[(size_t)n * kPartStride + f] = red[0];
        __syncthreads();
    }
}

// Fixed-order sum of the per-sequence partials; one block per problem.
__global__ __launch_bounds__(kThreads) void k_reduce_partials(const CdlDev *__restrict__ tab) {
    __shared__ double red[kThreads];
    const CdlDev &P = tab[blockIdx.x];
    const int tid = threadIdx.x;
    for (int f = 0; f < 5; ++f) {
        double a = 0.0;
        for (int n = tid; n < P.N; n += kThreads) a += gp(P.partials)[(size_t)n * kPartStride + f];
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

// X on the first sample of slot s only (the stride-4 support of Y3).
template <int TAPS>
__device__ __forceinline__ float recon1(float xt0, const float *hb, int s, const float *g) {
    float acc = xt0;
#pragma unroll
    for (int j = 0; j < TAPS; ++j) acc = fmaf(g[j], hb[4 * s + j], acc);
    return acc;
}

// ---------------------------------------------------------------------------- dict
// One iteration of dict_update! (train.jl:50-72) for one sequence (THR threads: 256 with 2 or
// 4 columns per pass and several sequences per CU, or 1024 with 8 columns per pass and one
// sequence per CU, which needs a quarter of the passes and so of the barrier phases); the means over n
// (:76,:83) follow in k_dict_reduce / k_dict_finalize.  Columns go through the transforms B
// at a time (B even): pass 1 packs two columns' Y3 into one complex length-Lg/4 FFT, stores
// each column's half spectrum in y3s and, unless X is held materialised already, writes it
// back materialised (Xt = X, h = 0), since G1 changes after this step; pass 2
// packs two columns' real inverse transforms into one complex FFT.  Accumulations over m
// stay in ascending m.
// LDS: twiddles | h with wrapped tail | taps of the batch | fa, fb (B x Lg) | ya, yb
// (B/2 x P) | ys (B x P) | accS, accT, accP (Lg/2+1 bins), with P = Lg/4.
template <int TAPS, int THR>
__global__ __launch_bounds__(THR) void k_dict(const CdlDev *__restrict__ tab, const int *__restrict__ starts, int np,
                                                   int B, int LgMax) {
    extern __shared__ float4 smem4[];
    int n;
    const CdlDev &P = tab[find_problem(starts, np, blockIdx.x, n)];
    float *sm = reinterpret_cast<float *>(smem4);
    const int tid = threadIdx.x;
    const int Lg = P.Lg, M = P.M, nslot = Lg >> 2, Pq = Lg >> 2, Ph = Pq / 2 + 1;
    const int H = Lg / 2 + 1;
    // layout sized for the largest problem of the launch
    const int HpX = ((LgMax / 2 + 1) + 3) & ~3, PqX = LgMax >> 2;
    float2 *tw = reinterpret_cast<float2 *>(sm);
    float *hb = sm + 2 * LgMax;
    float *gtap = hb + LgMax + TAPS + 4;                  // B * TAPS (rounded to 4)
    float2 *fa = reinterpret_cast<float2 *>(gtap + ((B * TAPS + 3) & ~3));
    float2 *fb = fa + B * LgMax;
    float2 *ya = fb + B * LgMax;
    float2 *yb = ya + (B / 2) * PqX;
    float2 *ys = yb + (B / 2) * PqX;
    float *accS = reinterpret_cast<float *>(ys + B * PqX);
    float2 *accT = reinterpret_cast<float2 *>(accS + HpX);
    float2 *accP = accT + HpX;
    const size_t ro = (size_t)n * Lg;
    const float rslot = 1.0f / (float)nslot, rLg = 1.0f / (float)Lg, rPq = 1.0f / (float)Pq;
    float2 *y3s = P.y3s + (size_t)n * M * Ph;
    const float *G1 = g1cur_of(P);
    const bool mat = !P.sc->xmat;   // set again only by k_cs after this step

    for (int l = tid; l < Lg; l += THR) tw[l] = gp(P.tw)[l];
    if (mat) load_h<TAPS>(P, ro, hb, tid, THR);
    for (int k = tid; k < HpX; k += THR) { accS[k] = 0.f; accT[k] = make_float2(0.f, 0.f); accP[k] = make_float2(0.f, 0.f); }

    // ---- pass 1: spectra sums over m; X materialised
    for (int m0 = 0; m0 < M; m0 += B) {
        const int nb = (M - m0) < B ? (M - m0) : B;
        const int npair = (nb + 1) / 2;
        if (mat)
            for (int i = tid; i < nb * TAPS; i += THR) {
                const int j = i % TAPS;
                gtap[i] = (j < Lg) ? gp(G1)[(size_t)(m0 + i / TAPS) * Lg + j] : 0.f;
            }
        if (nb & 1)
            for (int s = tid; s < Pq; s += THR) ya[(npair - 1) * Pq + s].y = 0.f;
        __syncthreads();
        for (int idx = tid; idx < nb * nslot; idx += THR) {
            int s;
            const int bb = fdivmod(idx, nslot, rslot, s), m = m0 + bb;
            const size_t col = ((size_t)n * M + m) * Lg;
            const float4 xt = ld4(P.Xt + col, s);
            const float4 dv = ld4(P.d + col, s);
            const float4 a1 = ld4(P.g1 + col, s);
            const float4 c3 = ld4(P.g3 + col, s);
            const float4 G1v = ld4(G1 + (size_t)m * Lg, s);
            const float4 G3v = ld4(P.G3 + (size_t)m * Lg, s);
            float x0 = xt.x;
            if (mat) {
                float X[4];
                recon4<TAPS>(xt, reinterpret_cast<const float4 *>(hb), s, gtap + bb * TAPS, X);
                st4(P.Xt + col, s, mk4(X));
                x0 = X[0];
            }
            float *yaf = reinterpret_cast<float *>(ya);
            yaf[2 * ((bb >> 1) * Pq + s) + (bb & 1)] = x0;    // Y3 = Pi4(X + y3) = Pi4(X): y3 = 0 on the support
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float qq = (f4(G1v, q) - f4(a1, q)) + (f4(G3v, q) - f4(c3, q));
                fa[bb * Lg + 4 * s + q] = make_float2(f4(dv, q), qq);
            }
        }
        __syncthreads();
        float2 *Y = team_fft<false>(ya, yb, P.plan4, tw, tid, THR, BlockSync{}, npair);
        float2 *C = team_fft<false>(fa, fb, P.plan, tw, tid, THR, BlockSync{}, nb);
        // unpack the column pairs: Y3h of column 2p is the Hermitian part, of 2p+1 the other
        for (int idx = tid; idx < nb * Pq; idx += THR) {
            int k;
            const int bb = fdivmod(idx, Pq, rPq, k);
            const float2 Yk = Y[(bb >> 1) * Pq + k], Ym = Y[(bb >> 1) * Pq + (k == 0 ? 0 : Pq - k)];
            const float2 v = (bb & 1) ? make_float2(0.5f * (Yk.y + Ym.y), -0.5f * (Yk.x - Ym.x))
                                      : make_float2(0.5f * (Yk.x + Ym.x), 0.5f * (Yk.y - Ym.y));
            ys[bb * Pq + k] = v;
            if (k < Ph) gp(y3s)[(size_t)(m0 + bb) * Ph + k] = v;
        }
        __syncthreads();
        for (int k = tid; k < H; k += THR) {
            int kq;
            (void)fdivmod(k, Pq, rPq, kq);
            float aS = accS[k];
            float2 aT = accT[k], aP = accP[k];
            for (int bb = 0; bb < nb; ++bb) {
                const float2 Ck = C[bb * Lg + k], Cm = C[bb * Lg + (k == 0 ? 0 : Lg - k)];
                const float2 Dh = make_float2(0.5f * (Ck.x + Cm.x), 0.5f * (Ck.y - Cm.y));
                const float2 Qh = make_float2(0.5f * (Ck.y + Cm.y), -0.5f * (Ck.x - Cm.x));
                const float2 Yk = ys[bb * Pq + kq];
                aS += Yk.x * Yk.x + Yk.y * Yk.y;
                aT = cadd(aT, cmul(Yk, Dh));
                aP = cadd(aP, cmul(Yk, Qh));
            }
            accS[k] = aS; accT[k] = aT; accP[k] = aP;
        }
        __syncthreads();
    }
    if (mat)
        for (int l = tid; l < Lg; l += THR) gp(P.h)[ro + l] = 0.f;

    // ---- per-sequence spectral step: Td, G2/g2 (update_g2!, train.jl:68-71), kappa
    for (int k = tid; k < Lg; k += THR) fa[k] = (k < H) ? accT[k] : cconj(accT[Lg - k]);
    __syncthreads();
    float2 *R2 = team_fft<true>(fa, fb, P.plan, tw, tid, THR, BlockSync{}, 1);
    float *tdb = reinterpret_cast<float *>(ys);      // Lg floats fit in ys (B * Lg/4 complex, B >= 2)
    {
        const float scale = 1.0f / (float)Lg;
        for (int l = tid; l < Lg; l += THR) tdb[l] = R2[l].x * scale;
    }
    __syncthreads();
    {
        // F(G2 - g2) of the previous dictionary iteration (sherman_CDL_RHS!, cdl_kernels.jl:8)
        for (int l = tid; l < Lg; l += THR) fa[l] = make_float2(gp(P.G2)[ro + l] - gp(P.g2)[ro + l], 0.f);
        __syncthreads();
        float2 *R = team_fft<false>(fa, fb, P.plan, tw, tid, THR, BlockSync{}, 1);
        const float c1 = P.sh_cdl;
        for (int k = tid; k < H; k += THR) {
            const float2 Gk = R[k];
            const float den = accS[k] + c1;
            accP[k] = make_float2((c1 * Gk.x - accP[k].x) / den, (c1 * Gk.y - accP[k].y) / den);
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
    __syncthreads();

    // ---- pass 2: d_m = q_m + F^-1[conj(Y3h_m) kappa]; g1 += d - G1; g3 += d - G3
    const float scale = 1.0f / (float)Lg;
    for (int m0 = 0; m0 < M; m0 += B) {
        const int nb = (M - m0) < B ? (M - m0) : B;
        const int npair = (nb + 1) / 2;
        for (int idx = tid; idx < nb * Ph; idx += THR) {
            const int bb = idx / Ph;
            const int k = idx - bb * Ph;
            const float2 v = gp(y3s)[(size_t)(m0 + bb) * Ph + k];
            ys[bb * Pq + k] = v;
            if (k > 0 && Pq - k >= Ph) ys[bb * Pq + Pq - k] = cconj(v);
        }
        __syncthreads();
        for (int idx = tid; idx < npair * Lg; idx += THR) {
            int k, kq;
            const int pp = fdivmod(idx, Lg, rLg, k);
            (void)fdivmod(k, Pq, rPq, kq);
            const float2 kap = (k < H) ? accP[k] : cconj(accP[Lg - k]);
            const float2 CA = cmul(cconj(ys[(2 * pp) * Pq + kq]), kap);
            const float2 CB = (2 * pp + 1 < nb) ? cmul(cconj(ys[(2 * pp + 1) * Pq + kq]), kap) : make_float2(0.f, 0.f);
            fa[idx] = make_float2(CA.x - CB.y, CA.y + CB.x);
        }
        __syncthreads();
        float2 *R = team_fft<true>(fa, fb, P.plan, tw, tid, THR, BlockSync{}, npair);
        for (int idx = tid; idx < nb * nslot; idx += THR) {
            int s;
            const int bb = fdivmod(idx, nslot, rslot, s), m = m0 + bb;
            const size_t col = ((size_t)n * M + m) * Lg;
            const float4 G1v = ld4(G1 + (size_t)m * Lg, s);
            const float4 G3v = ld4(P.G3 + (size_t)m * Lg, s);
            const float4 a1 = ld4(P.g1 + col, s);
            const float4 c3 = ld4(P.g3 + col, s);
            const float2 *Rp = R + (bb >> 1) * Lg + 4 * s;
            float dn[4], g1n[4], g3n[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float qv = (f4(G1v, q) - f4(a1, q)) + (f4(G3v, q) - f4(c3, q));
                const float cv = (bb & 1) ? Rp[q].y : Rp[q].x;
                dn[q] = qv + cv * scale;
                g1n[q] = f4(a1, q) + dn[q] - f4(G1v, q);       // train.jl:70
                g3n[q] = f4(c3, q) + dn[q] - f4(G3v, q);       // train.jl:72
            }
            st4(P.d + col, s, mk4(dn));
            st4(P.g1 + col, s, mk4(g1n));
            st4(P.g3 + col, s, mk4(g3n));
        }
        __syncthreads();
    }
}

// Partial sums over a chunk of sequences of (d + g1), (d + g3) and d, per (l, m)
// (train.jl:76,83, train.jl:95), in double.  Blocks of a problem: (tile, chunk).
__global__ __launch_bounds__(kThreads) void k_dict_reduce(const CdlDev *__restrict__ tab, const int *__restrict__ starts,
                                                          int np) {
    int b;
    const CdlDev &P = tab[find_problem(starts, np, blockIdx.x, b)];
    const int LM = P.Lg * P.M;
    const int tiles = (LM + kThreads - 1) / kThreads;
    const int c = b / tiles, tile = b - c * tiles;
    const int i = tile * kThreads + threadIdx.x;
    if (i >= LM) return;
    const int N = P.N, nch = P.nch;
    const int n0 = (int)((long)c * N / nch), n1 = (int)((long)(c + 1) * N / nch);
    double a1 = 0.0, a3 = 0.0, ad = 0.0;
    for (int n = n0; n < n1; ++n) {
        const size_t e = (size_t)n * LM + i;
        const float dv = gp(P.d)[e];
        a1 += (double)(dv + gp(P.g1)[e]);
        a3 += (double)(dv + gp(P.g3)[e]);
        ad += (double)dv;
    }
    gp(P.dpart)[((size_t)c * 3 + 0) * LM + i] = a1;
    gp(P.dpart)[((size_t)c * 3 + 1) * LM + i] = a3;
    gp(P.dpart)[((size_t)c * 3 + 2) * LM + i] = ad;
}

// Column m: v1 = mean(d+g1), v2 = mean(d+g3); G3 = projection2box!(v2) (cdl_kernels.jl:37-51);
// G1new = prox_simplex_zeros!(v1) (:53-66, FP64 combination, quirk group 4*floor(l/4)) into
// the non-current G1 buffer; meand = mean(d) (D_best candidate, train.jl:95);
// gsq = |F(G1new)|^2.
__global__ __launch_bounds__(kThreads) void k_dict_finalize(const CdlDev *__restrict__ tab, const int *__restrict__ starts,
                                                            int np, int LgMax) {
    extern __shared__ float4 smem4[];
    int m;
    const CdlDev &P = tab[find_problem(starts, np, blockIdx.x, m)];
    float *v1 = reinterpret_cast<float *>(smem4);
    float2 *tw = reinterpret_cast<float2 *>(v1 + ((LgMax + 3) & ~3));
    float2 *fa = tw + LgMax;
    float2 *fb = fa + LgMax;
    const int tid = threadIdx.x, Lg = P.Lg;
    const int LM = Lg * P.M;
    const int flm = gp(P.fl)[m];
    const double invN = 1.0 / (double)P.N;
    float *G1new = P.G1b[(P.sc->g1cur & 1) ^ 1];
    for (int l = tid; l < Lg; l += kThreads) tw[l] = gp(P.tw)[l];
    for (int l = tid; l < Lg; l += kThreads) {
        const int i = m * Lg + l;
        double s1 = 0.0, s3 = 0.0, sd = 0.0;
        for (int c = 0; c < P.nch; ++c) {
            s1 += gp(P.dpart)[((size_t)c * 3 + 0) * LM + i];
            s3 += gp(P.dpart)[((size_t)c * 3 + 1) * LM + i];
            sd += gp(P.dpart)[((size_t)c * 3 + 2) * LM + i];
        }
        v1[l] = (float)(s1 * invN);
        const float v2 = (float)(s3 * invN);
        const bool sup = l < flm;
        float g3;
        if (v2 >= 1.f) g3 = sup ? 1.f : 0.f;
        else if (v2 <= 0.f) g3 = 0.f;
        else g3 = sup ? v2 : 0.f;
        gp(P.G3)[i] = g3;
        gp(P.meand)[i] = (float)(sd * invN);
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
    float2 *C = team_fft<false>(fa, fb, P.plan, tw, tid, kThreads, BlockSync{});
    for (int k = tid; k < Lg; k += kThreads) gp(P.gsq)[(size_t)m * Lg + k] = C[k].x * C[k].x + C[k].y * C[k].y;
}

// |F(gp(G1)[:, m])|^2 for the current filters (constructor and reset).
__global__ __launch_bounds__(kThreads) void k_g1_spectrum(const CdlDev *__restrict__ tab, const int *__restrict__ starts,
                                                          int np, int LgMax) {
    extern __shared__ float4 smem4[];
    int m;
    const CdlDev &P = tab[find_problem(starts, np, blockIdx.x, m)];
    float2 *tw = reinterpret_cast<float2 *>(smem4);
    float2 *fa = tw + LgMax;
    float2 *fb = fa + LgMax;
    const int tid = threadIdx.x, Lg = P.Lg;
    const float *G1 = g1cur_of(P);
    for (int l = tid; l < Lg; l += kThreads) {
        tw[l] = gp(P.tw)[l];
        fa[l] = make_float2(gp(G1)[(size_t)m * Lg + l], 0.f);
    }
    __syncthreads();
    float2 *C = team_fft<false>(fa, fb, P.plan, tw, tid, kThreads, BlockSync{});
    for (int k = tid; k < Lg; k += kThreads) gp(P.gsq)[(size_t)m * Lg + k] = C[k].x * C[k].x + C[k].y * C[k].y;
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
