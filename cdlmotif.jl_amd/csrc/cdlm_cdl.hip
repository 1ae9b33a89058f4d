// cdlm_cdl.hip -- Fourier-domain CDL ADMM of CDLmotif.jl for gfx950.
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
//     where (*) is circular cross-correlation with the filter g_m = G1[:, m], which is zero
//     past filter_len (prox_simplex_zeros!), so it costs filter_len taps per sample;
//     S_v = F(sum_m g_m (x) v_m) is a filter_len-tap circular convolution summed over m.
//   * X is held as the pair (Xt = v/c0, h): X = Xt + g_rec (*) h is rebuilt on the fly
//     by the next kernel that reads X, with g_rec the filter h was computed with.
//   * the dictionary step keeps d spatial; its solve is
//       d_m = q_m + F^-1[conj(Y3h_m) kappa],   kappa = (c1 F(G2-g2) - sum_m Y3h_m Q_m)/(|Y3h|^2 + c1)
//     with q_m = (G1_m - g1_m) + (G3_m - g3_m) (c1 = sherman_const_cdl; no 1/c1, as
//     sherman_CDL_combine).  Y3 = Pi4(X + y3) is nonzero on every 4th sample only, so its
//     spectrum is a length Lg/4 DFT repeated 4 times.
// All quantities are exact rewrites of the reference iteration; only rounding differs.
//
// HBM layout: the reference's column-major (Lg, M, N) with l fastest; one work-group owns
// one sequence n and streams its contiguous (Lg x M) slab.

#include <cmath>
#include <cstdlib>
#include <cstdarg>
#include <cstring>
#include <memory>

#include "cdlm_internal.h"

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kMaxTaps = 32;
constexpr int kPartStride = 8;  // doubles per sequence in the objective partials

// Device-side scalar state of one problem.
struct CdlScal {
    float obj_best;
    int best_idx;
    int improved;
    int pad;
    double best_stats[3];  // count, sum, sum of squares of positive X_best entries
    double red[kPartStride];
};

struct CdlDev {
    int L4, N, M, Lg, gamma, L;
    float rho, sigma, lambda, sh_csc, sh_cdl;
    const float *s;
    float *Xt, *y2, *y3, *d, *g1, *g3;
    float *h, *y1, *Y1, *DX, *G2, *g2, *Td;
    const float *G1cur, *G1rec;
    float *G3;
    const int *fl;
    const float *cS, *sS;
    const float2 *tw;
    float2 *y3s;          // per (m, n): Y3 half spectrum, bins 0..Lg/8 (dictionary scratch)
    FftPlan plan, plan4;
};

__device__ __forceinline__ float4 ld4(const float *p, int s) { return reinterpret_cast<const float4 *>(p)[s]; }
__device__ __forceinline__ void st4(float *p, int s, float4 v) { reinterpret_cast<float4 *>(p)[s] = v; }
__device__ __forceinline__ float f4(const float4 &v, int q) { return q == 0 ? v.x : q == 1 ? v.y : q == 2 ? v.z : v.w; }
__device__ __forceinline__ float4 mk4(const float *a) { return make_float4(a[0], a[1], a[2], a[3]); }

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

// X on the first sample of slot s only (the stride-4 support of Y3).
template <int TAPS>
__device__ __forceinline__ float recon1(float xt0, const float *hb, int s, const float *g) {
    float acc = xt0;
#pragma unroll
    for (int j = 0; j < TAPS; ++j) acc = fmaf(g[j], hb[4 * s + j], acc);
    return acc;
}

// z[4s..4s+3] += sum_j g[j] * v[(l - j) mod Lg]; vb4 holds v at float4 offset TAPS/4 with
// the wrapped head of TAPS samples in front.
template <int TAPS>
__device__ __forceinline__ void conv_acc4(const float4 *vb4, int s, const float *g, float4 *zacc4) {
    float vw[TAPS + 4];
#pragma unroll
    for (int u = 0; u <= TAPS / 4; ++u) {
        float4 t = vb4[s + u];
        vw[4 * u] = t.x; vw[4 * u + 1] = t.y; vw[4 * u + 2] = t.z; vw[4 * u + 3] = t.w;
    }
    float4 z = zacc4[s];
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

// sum_j g[j] * v[(4s + q - j) mod Lg] for q = 0..3, returned in registers.
template <int TAPS>
__device__ __forceinline__ float4 conv4(const float4 *vb4, int s, const float *g) {
    float vw[TAPS + 4];
#pragma unroll
    for (int u = 0; u <= TAPS / 4; ++u) {
        float4 t = vb4[s + u];
        vw[4 * u] = t.x; vw[4 * u + 1] = t.y; vw[4 * u + 2] = t.z; vw[4 * u + 3] = t.w;
    }
    float o[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        float acc = 0.f;
#pragma unroll
        for (int j = 0; j < TAPS; ++j) acc = fmaf(g[j], vw[TAPS + q - j], acc);
        o[q] = acc;
    }
    return mk4(o);
}

template <int TAPS>
__device__ __forceinline__ void load_taps(const float *G, int m, int Lg, float *g) {
#pragma unroll
    for (int j = 0; j < TAPS; ++j) g[j] = G[(size_t)m * Lg + j];
}

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

// ---------------------------------------------------------------------------- init
// CDL_setup.jl:183-230 for one sequence: X = Pi4(X0); DX = sum_m D_m (x) X_m (the
// spectral product of :201-203 as a circular convolution); d = D; Y1 = y = g = 0; G2 = DX.
template <int TAPS>
__global__ __launch_bounds__(kThreads) void k_init(CdlDev P, const float *__restrict__ X0, uint64_t seed) {
    extern __shared__ float4 smem4[];
    float *sm = reinterpret_cast<float *>(smem4);
    const int n = blockIdx.x, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int Lg = P.Lg, M = P.M, nslot = Lg >> 2;
    const int vbLen = Lg + TAPS;
    float *vbase = sm;
    float *zbase = vbase + kWaves * vbLen;
    for (int i = tid; i < kWaves * Lg; i += kThreads) zbase[i] = 0.f;
    __syncthreads();
    float4 *vb4 = smem4 + (w * vbLen) / 4;
    float4 *zacc4 = smem4 + (kWaves * vbLen + w * Lg) / 4;
    const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int m = w; m < M; m += kWaves) {
        const int mm = __builtin_amdgcn_readfirstlane(m);
        const size_t col = ((size_t)n * M + mm) * Lg;
        float g[TAPS];
        load_taps<TAPS>(P.G1cur, mm, Lg, g);
        for (int s = lane; s < nslot; s += 64) {
            float x;
            if (X0) x = X0[col + 4 * s];
            else x = gauss_of(seed, col + 4 * (size_t)s);
            float4 xp = make_float4(x, 0.f, 0.f, 0.f);
            st4(P.Xt + col, s, xp);
            st4(P.y2 + col, s, zero4);
            st4(P.y3 + col, s, zero4);
            st4(P.g1 + col, s, zero4);
            st4(P.g3 + col, s, zero4);
            st4(P.d + col, s, ld4(P.G1cur + (size_t)mm * Lg, s));
            vb4[TAPS / 4 + s] = xp;
        }
        wave_lds_sync();
        if (lane < TAPS / 4) vb4[TAPS / 4 - 1 - lane] = vb4[TAPS / 4 + nslot - 1 - lane];
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
        P.DX[ro + l] = z;
        P.G2[ro + l] = z;
        P.h[ro + l] = 0.f;
        P.y1[ro + l] = 0.f;
        P.Y1[ro + l] = 0.f;
        P.g2[ro + l] = 0.f;
        P.Td[ro + l] = 0.f;
    }
}

// ---------------------------------------------------------------------------- CSC
// One iteration of code_update! (train.jl:5-43) for sequence n = blockIdx.x.  Each wave owns
// the columns m = w, w+4, ...; a lane owns the float4 slots s = lane + 64 i (i < NSL) of a
// column.  Column m+4's loads are issued before column m is processed (two register sets),
// so a wave always has a column of loads in flight while it computes.
// LDS (float4 units): twiddles (Lg/2) | h with wrapped tail | 4 x v column with wrapped
// head | 4 x z partial; the FFT ping-pong buffers reuse the v columns after the loop.
template <int TAPS, int NSL>
__global__ __launch_bounds__(kThreads, (NSL <= 4 ? 3 : 2)) void k_csc(CdlDev P) {
    extern __shared__ float4 smem4[];
    const int n = blockIdx.x, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int Lg = P.Lg, M = P.M, nslot = Lg >> 2;
    const int hb4Len = (Lg + TAPS + 4) / 4, vb4Len = (Lg + TAPS) / 4;
    float2 *tw = reinterpret_cast<float2 *>(smem4);
    float4 *hb4 = smem4 + Lg / 2;
    float4 *vbase4 = hb4 + hb4Len;
    float4 *zbase4 = vbase4 + kWaves * vb4Len;
    float *hb = reinterpret_cast<float *>(hb4);
    float *zbase = reinterpret_cast<float *>(zbase4);
    float2 *fa = reinterpret_cast<float2 *>(vbase4);
    float2 *fb = fa + Lg;
    const size_t ro = (size_t)n * Lg;

    for (int l = tid; l < Lg; l += kThreads) { tw[l] = P.tw[l]; hb[l] = P.h[ro + l]; }
    for (int t = tid; t < TAPS + 4; t += kThreads) hb[Lg + t] = P.h[ro + t];
    for (int i = tid; i < kWaves * Lg; i += kThreads) zbase[i] = 0.f;
    __syncthreads();

    const float lam = P.lambda;
    const float inv_c0 = 1.0f / P.sh_csc;
    float4 *vb4 = vbase4 + w * vb4Len;
    float4 *zacc4 = zbase4 + w * (Lg / 4);
    float4 nx[NSL], n2[NSL], n3[NSL];
    if (w < M) {
        const size_t col = ((size_t)n * M + w) * Lg;
#pragma unroll
        for (int i = 0; i < NSL; ++i) {
            const int s = lane + 64 * i;
            if (s < nslot) { nx[i] = ld4(P.Xt + col, s); n2[i] = ld4(P.y2 + col, s); n3[i] = ld4(P.y3 + col, s); }
        }
    }
    for (int m = w; m < M; m += kWaves) {
        const int mm = __builtin_amdgcn_readfirstlane(m);
        const size_t col = ((size_t)n * M + mm) * Lg;
        float4 xt[NSL], a2[NSL], a3[NSL];
#pragma unroll
        for (int i = 0; i < NSL; ++i) { xt[i] = nx[i]; a2[i] = n2[i]; a3[i] = n3[i]; }
        if (mm + kWaves < M) {
            const size_t coln = ((size_t)n * M + mm + kWaves) * Lg;
#pragma unroll
            for (int i = 0; i < NSL; ++i) {
                const int s = lane + 64 * i;
                if (s < nslot) { nx[i] = ld4(P.Xt + coln, s); n2[i] = ld4(P.y2 + coln, s); n3[i] = ld4(P.y3 + coln, s); }
            }
        }
        float gr[TAPS], gc[TAPS];
        load_taps<TAPS>(P.G1rec, mm, Lg, gr);
        load_taps<TAPS>(P.G1cur, mm, Lg, gc);
#pragma unroll
        for (int i = 0; i < NSL; ++i) {
            const int s = lane + 64 * i;
            if (s < nslot) {
                float X[4];
                recon4<TAPS>(xt[i], hb4, s, gr, X);
                float y2n[4], y3n[4], v[4], xn[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float x = X[q];
                    const float t = x + f4(a2[i], q);                   // shrink! (csc_kernels.jl:58-59)
                    const float at = fabsf(t) - lam;
                    const float Y2 = at > 0.f ? copysignf(at, t) : 0.f;
                    const float Y3 = (q == 0) ? x + f4(a3[i], q) : 0.f;  // projectionX! (:73-77)
                    y2n[q] = f4(a2[i], q) + x - Y2;                       // train.jl:13
                    y3n[q] = f4(a3[i], q) + x - Y3;                       // train.jl:14
                    v[q] = (Y2 - y2n[q]) + (Y3 - y3n[q]);
                    xn[q] = v[q] * inv_c0;
                }
                st4(P.y2 + col, s, mk4(y2n));
                st4(P.y3 + col, s, mk4(y3n));
                st4(P.Xt + col, s, mk4(xn));
                vb4[TAPS / 4 + s] = mk4(v);
            }
        }
        wave_lds_sync();
        if (lane < TAPS / 4) vb4[TAPS / 4 - 1 - lane] = vb4[TAPS / 4 + nslot - 1 - lane];
        wave_lds_sync();
#pragma unroll
        for (int i = 0; i < NSL; ++i) {
            const int s = lane + 64 * i;
            if (s < nslot) conv_acc4<TAPS>(vb4, s, gc, zacc4);
        }
        wave_lds_sync();
    }
    __syncthreads();

    // z = sum over waves; u = Y1 - y1' with y1' = y1 + DX - Y1 (train.jl:12)
    for (int l = tid; l < Lg; l += kThreads) {
        float z = zbase[l];
#pragma unroll
        for (int ww = 1; ww < kWaves; ++ww) z += zbase[ww * Lg + l];
        const float Y1o = P.Y1[ro + l];
        const float y1n = P.y1[ro + l] + P.DX[ro + l] - Y1o;
        zbase[l] = y1n;
        fa[l] = make_float2(z, Y1o - y1n);
    }
    __syncthreads();
    float2 *C = team_fft<false>(fa, fb, P.plan, tw, tid, kThreads, BlockSync{});
    float2 *O = (C == fa) ? fb : fa;
    const float c0 = P.sh_csc;
    for (int k = tid; k < Lg; k += kThreads) {
        const float2 Ck = C[k], Cm = C[k == 0 ? 0 : Lg - k];
        const float2 Z = make_float2(0.5f * (Ck.x + Cm.x), 0.5f * (Ck.y - Cm.y));   // F(z)
        const float2 U = make_float2(0.5f * (Ck.y + Cm.y), -0.5f * (Ck.x - Cm.x));  // F(u)
        const float ss = P.sS[k];
        const float den = c0 * P.cS[k];
        const float2 hh = make_float2((c0 * U.x - Z.x) / den, (c0 * U.y - Z.y) / den);
        const float2 dx = make_float2(Z.x * inv_c0 + ss * hh.x, Z.y * inv_c0 + ss * hh.y);
        O[k] = make_float2(hh.x - dx.y, hh.y + dx.x);   // h_hat + i DX_hat
    }
    __syncthreads();
    float2 *R = team_fft<true>(O, C, P.plan, tw, tid, kThreads, BlockSync{});
    const float scale = 1.0f / (float)Lg;
    const float rho = P.rho;
    for (int l = tid; l < Lg; l += kThreads) {
        const float hn = R[l].x * scale;
        const float DXn = R[l].y * scale;
        const bool win = (l >= P.gamma) && (l < P.gamma + P.L4);
        const float b = win ? P.s[(size_t)n * P.L4 + (l - P.gamma)] : 0.f;
        const float wl = win ? 1.f : 0.f;
        const float y1n = zbase[l];
        P.h[ro + l] = hn;
        P.DX[ro + l] = DXn;
        P.y1[ro + l] = y1n;
        P.Y1[ro + l] = (b + rho * (DXn + y1n)) / (wl + rho);   // update_Y1! (csc_kernels.jl:91)
    }
}

// ---------------------------------------------------------------------------- resident CSC
// K consecutive CSC iterations with the per-sequence state held on chip.  Between
// dictionary updates G1 is fixed and a CSC iteration of sequence n touches only that
// sequence, so the state (Xt, y2, y3: one column per wave in registers; h, y1, Y1, DX in LDS)
// stays resident for all K iterations and goes through HBM once per launch.  A sequence is
// owned by a group of G work-groups (one per CU, co-resident by cooperative launch); member j
// holds columns [j*Mg, (j+1)*Mg).  Per iteration the members exchange their z partial sums
// through L2 (agent-scope release/acquire on a per-member counter, bounded spin) and each
// computes the identical per-sequence spectral step (same summation order in every member).
// Group g's members are blocks g, g+NG, ... (NG a multiple of 8: one XCD).
constexpr int kResThreads = 1024;
constexpr int kResWaves = kResThreads / 64;
constexpr long long kSpinLimit = 1LL << 22;

struct ResArgs {
    int K, G, NG, Mg;
    unsigned long long epoch;
    unsigned long long *flags;   // [NG * G] progress counters
    float *xbuf;                 // [NG][G][2][Lg] z partials (double buffered by iteration parity)
    int *err;                    // set when an exchange wait times out
};

__device__ __forceinline__ void st_agent(float *p, float v) {
    __hip_atomic_store(reinterpret_cast<unsigned int *>(p), __float_as_uint(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float ld_agent(const float *p) {
    return __uint_as_float(__hip_atomic_load(reinterpret_cast<const unsigned int *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

template <int TAPS>
__global__ __launch_bounds__(kResThreads) void k_csc_res(CdlDev P, ResArgs A) {
    extern __shared__ float4 smem4[];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int Lg = P.Lg, M = P.M, nslot = Lg >> 2;
    const int group = blockIdx.x % A.NG, member = blockIdx.x / A.NG;
    const int mw = member * A.Mg + w;
    const int mm = __builtin_amdgcn_readfirstlane(mw);
    const bool own = (w < A.Mg) && (mm < M);
    int nw = M - member * A.Mg;                               // waves holding a column (the lowest ones)
    nw = nw < A.Mg ? nw : A.Mg;
    nw = nw < 0 ? 0 : nw;
    const int hb4Len = (Lg + TAPS + 4) / 4, vb4Len = (Lg + TAPS) / 4;
    float2 *tw = reinterpret_cast<float2 *>(smem4);
    float4 *hb4 = smem4 + Lg / 2;
    float *hb = reinterpret_cast<float *>(hb4);
    float4 *vbase4 = hb4 + hb4Len;
    float *y1v = reinterpret_cast<float *>(vbase4 + kResWaves * vb4Len);
    float *Y1v = y1v + Lg;
    float *DXv = Y1v + Lg;
    float *zt = DXv + Lg;
    float2 *fa = reinterpret_cast<float2 *>(zt + Lg);
    float2 *fb = fa + Lg;
    float *taps = reinterpret_cast<float *>(fb + Lg);         // [wave][2][TAPS]: rebuild filter, current filter
    float4 *vb4 = vbase4 + w * vb4Len;
    float *gr = taps + w * 2 * TAPS;
    float *gc = gr + TAPS;
    const float lam = P.lambda, c0 = P.sh_csc, inv_c0 = 1.0f / P.sh_csc, rho = P.rho;
    const float scale = 1.0f / (float)Lg;

    for (int l = tid; l < Lg; l += kResThreads) tw[l] = P.tw[l];
    int round = 0;
    for (int n = group; n < P.N; n += A.NG, ++round) {
        const size_t ro = (size_t)n * Lg;
        const size_t col = ((size_t)n * M + mm) * Lg;
        if (own && lane < TAPS) {
            gr[lane] = P.G1rec[(size_t)mm * Lg + lane];     // the filter this sequence's h was made with
            gc[lane] = P.G1cur[(size_t)mm * Lg + lane];
        }
        for (int l = tid; l < Lg; l += kResThreads) {
            hb[l] = P.h[ro + l];
            y1v[l] = P.y1[ro + l];
            Y1v[l] = P.Y1[ro + l];
            DXv[l] = P.DX[ro + l];
        }
        for (int t = tid; t < TAPS + 4; t += kResThreads) hb[Lg + t] = P.h[ro + t];
        float4 xt[4], a2[4], a3[4];
        if (own) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int s = lane + 64 * i;
                if (s < nslot) { xt[i] = ld4(P.Xt + col, s); a2[i] = ld4(P.y2 + col, s); a3[i] = ld4(P.y3 + col, s); }
            }
        }
        __syncthreads();
        for (int it = 0; it < A.K; ++it) {
            if (own) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int s = lane + 64 * i;
                    if (s < nslot) {
                        float X[4];
                        recon4<TAPS>(xt[i], hb4, s, gr, X);
                        float y2n[4], y3n[4], v[4];
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const float x = X[q];
                            const float t = x + f4(a2[i], q);                   // shrink! (csc_kernels.jl:58-59)
                            const float at = fabsf(t) - lam;
                            const float Y2 = at > 0.f ? copysignf(at, t) : 0.f;
                            const float Y3 = (q == 0) ? x + f4(a3[i], q) : 0.f;  // projectionX! (:73-77)
                            y2n[q] = f4(a2[i], q) + x - Y2;                       // train.jl:13
                            y3n[q] = f4(a3[i], q) + x - Y3;                       // train.jl:14
                            v[q] = (Y2 - y2n[q]) + (Y3 - y3n[q]);
                        }
                        a2[i] = mk4(y2n);
                        a3[i] = mk4(y3n);
                        xt[i] = make_float4(v[0] * inv_c0, v[1] * inv_c0, v[2] * inv_c0, v[3] * inv_c0);
                        vb4[TAPS / 4 + s] = mk4(v);
                    }
                }
                wave_lds_sync();
                if (lane < TAPS / 4) vb4[TAPS / 4 - 1 - lane] = vb4[TAPS / 4 + nslot - 1 - lane];
                if (it == 0 && lane < TAPS) gr[lane] = gc[lane];   // later rebuilds use this filter
                wave_lds_sync();
                // z partial of this column, written in place over the v buffer: slot s reads
                // v slots s..s+TAPS/4 before writing slot s, and slot groups of consecutive
                // unrolled steps are disjoint
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int s = lane + 64 * i;
                    if (s < nslot) {
                        const float4 z = conv4<TAPS>(vb4, s, gc);
                        vb4[s] = z;
                    }
                }
            }
            __syncthreads();
            // z partial of this member: waves in ascending order
            const int parity = it & 1;
            float *xme = A.xbuf + (((size_t)group * A.G + member) * 2 + parity) * Lg;
            for (int l = tid; l < Lg; l += kResThreads) {
                float acc = 0.f;
                for (int ww = 0; ww < nw; ++ww) acc += reinterpret_cast<const float *>(vbase4 + ww * vb4Len)[l];
                if (A.G == 1) zt[l] = acc;
                else st_agent(xme + l, acc);
            }
            if (A.G > 1) {
                const unsigned long long target = (A.epoch << 32) | (unsigned long long)(round * A.K + it + 1);
                __syncthreads();
                if (tid == 0)
                    __hip_atomic_store(A.flags + (size_t)group * A.G + member, target, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                if (tid < A.G && tid != member) {
                    const unsigned long long *f = A.flags + (size_t)group * A.G + tid;
                    long long spins = 0;
                    while (__hip_atomic_load(f, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
                        if (++spins > kSpinLimit) { atomicExch(A.err, 1); break; }
                        if ((spins & 255) == 0 && __hip_atomic_load(A.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
                        __builtin_amdgcn_s_sleep(2);
                    }
                }
                __syncthreads();
                for (int l = tid; l < Lg; l += kResThreads) {
                    float acc = 0.f;
                    for (int j = 0; j < A.G; ++j) acc += ld_agent(A.xbuf + (((size_t)group * A.G + j) * 2 + parity) * Lg + l);
                    zt[l] = acc;
                }
            }
            __syncthreads();
            // u = Y1 - y1' with y1' = y1 + DX - Y1 (train.jl:12)
            for (int l = tid; l < Lg; l += kResThreads) {
                const float Y1o = Y1v[l];
                const float y1n = y1v[l] + DXv[l] - Y1o;
                y1v[l] = y1n;
                fa[l] = make_float2(zt[l], Y1o - y1n);
            }
            __syncthreads();
            float2 *C = team_fft<false>(fa, fb, P.plan, tw, tid, kResThreads, BlockSync{});
            float2 *O = (C == fa) ? fb : fa;
            for (int k = tid; k < Lg; k += kResThreads) {
                const float2 Ck = C[k], Cm = C[k == 0 ? 0 : Lg - k];
                const float2 Z = make_float2(0.5f * (Ck.x + Cm.x), 0.5f * (Ck.y - Cm.y));
                const float2 U = make_float2(0.5f * (Ck.y + Cm.y), -0.5f * (Ck.x - Cm.x));
                const float ss = P.sS[k];
                const float den = c0 * P.cS[k];
                const float2 hh = make_float2((c0 * U.x - Z.x) / den, (c0 * U.y - Z.y) / den);
                const float2 dx = make_float2(Z.x * inv_c0 + ss * hh.x, Z.y * inv_c0 + ss * hh.y);
                O[k] = make_float2(hh.x - dx.y, hh.y + dx.x);
            }
            __syncthreads();
            float2 *R = team_fft<true>(O, C, P.plan, tw, tid, kResThreads, BlockSync{});
            for (int l = tid; l < Lg; l += kResThreads) {
                const float hn = R[l].x * scale;
                const float DXn = R[l].y * scale;
                const bool win = (l >= P.gamma) && (l < P.gamma + P.L4);
                const float b = win ? P.s[(size_t)n * P.L4 + (l - P.gamma)] : 0.f;
                const float wl = win ? 1.f : 0.f;
                hb[l] = hn;
                if (l < TAPS + 4) hb[Lg + l] = hn;
                DXv[l] = DXn;
                Y1v[l] = (b + rho * (DXn + y1v[l])) / (wl + rho);   // update_Y1! (csc_kernels.jl:91)
            }
            __syncthreads();
        }
        // write the state back
        if (own) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int s = lane + 64 * i;
                if (s < nslot) { st4(P.Xt + col, s, xt[i]); st4(P.y2 + col, s, a2[i]); st4(P.y3 + col, s, a3[i]); }
            }
        }
        if (member == 0)
            for (int l = tid; l < Lg; l += kResThreads) {
                P.h[ro + l] = hb[l];
                P.y1[ro + l] = y1v[l];
                P.Y1[ro + l] = Y1v[l];
                P.DX[ro + l] = DXv[l];
            }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------- best / objective
// Per-sequence partials of objective() (CDL_setup.jl:286): sum (DX_window - s)^2, sum |X|,
// and the positive-X statistics of update_best!/wrap.jl:12; writes the candidate
// X_best slice X[gamma : 4 : gamma+4L-1] (store.jl:15) and optionally X itself.
template <int TAPS>
__global__ __launch_bounds__(kThreads) void k_best(CdlDev P, float *xs0, float *xs1, const CdlScal *sc,
                                                   double *partials, float *Xfull) {
    extern __shared__ float4 smem4[];
    float *sm = reinterpret_cast<float *>(smem4);
    __shared__ double red[kThreads];
    const int n = blockIdx.x, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int Lg = P.Lg, M = P.M, nslot = Lg >> 2, L = P.L;
    float *hb = sm;
    const size_t ro = (size_t)n * Lg;
    for (int l = tid; l < Lg; l += kThreads) hb[l] = P.h[ro + l];
    for (int t = tid; t < TAPS + 4; t += kThreads) hb[Lg + t] = P.h[ro + t];
    __syncthreads();
    float *xc = (sc->best_idx == 0) ? xs1 : xs0;
    const int s0 = P.gamma >> 2;
    double ab = 0.0, cnt = 0.0, sx = 0.0, sxx = 0.0;
    for (int m = w; m < M; m += kWaves) {
        const int mm = __builtin_amdgcn_readfirstlane(m);
        const size_t col = ((size_t)n * M + mm) * Lg;
        float gr[TAPS];
        load_taps<TAPS>(P.G1rec, mm, Lg, gr);
        float abf = 0.f;
        for (int s = lane; s < nslot; s += 64) {
            float X[4];
            recon4<TAPS>(ld4(P.Xt + col, s), reinterpret_cast<const float4 *>(hb), s, gr, X);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                abf += fabsf(X[q]);
                if (X[q] > 0.f) { cnt += 1.0; sx += X[q]; sxx += (double)X[q] * X[q]; }
            }
            if (Xfull) st4(Xfull + col, s, mk4(X));
            if (s >= s0 && s < s0 + L) xc[((size_t)n * M + mm) * L + (s - s0)] = X[0];
        }
        ab += abf;
    }
    double fid = 0.0;
    for (int i = tid; i < P.L4; i += kThreads) {
        const float r = P.DX[ro + P.gamma + i] - P.s[(size_t)n * P.L4 + i];
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
        if (tid == 0) partials[(size_t)n * kPartStride + f] = red[0];
        __syncthreads();
    }
}

// Fixed-order sum of the per-sequence partials.
__global__ __launch_bounds__(kThreads) void k_reduce_partials(const double *partials, int N, CdlScal *sc) {
    __shared__ double red[kThreads];
    const int tid = threadIdx.x;
    for (int f = 0; f < 5; ++f) {
        double a = 0.0;
        for (int n = tid; n < N; n += kThreads) a += partials[(size_t)n * kPartStride + f];
        red[tid] = a;
        __syncthreads();
        for (int st = kThreads / 2; st > 0; st >>= 1) {
            if (tid < st) red[tid] += red[tid + st];
            __syncthreads();
        }
        if (tid == 0) sc->red[f] = red[0];
        __syncthreads();
    }
}

// update_best! (train.jl:91-98): obj = 0.5*norm(W DX - s)^2 + lambda*sum|X| with the
// reference's Float32 intermediates (norm and the |X| sum in Float32, 0.5 in Float64);
// obj_best is a Float32 field.  force=1 accepts the candidate (constructor, :233-236).
__global__ __launch_bounds__(kThreads) void k_decide(CdlScal *sc, float lambda, int force, double *trace, int slot,
                                                     const float *meand, float *Dbest, int LM) {
    __shared__ int imp;
    if (threadIdx.x == 0) {
        const float nrm = (float)sqrt(sc->red[0]);
        const float nrm2 = nrm * nrm;
        const float reg = lambda * (float)sc->red[1];
        const double obj = 0.5 * (double)nrm2 + (double)reg;
        int better = force || (obj < (double)sc->obj_best);
        if (trace) trace[slot] = obj;
        sc->red[5] = obj;
        sc->improved = better;
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
    if (imp && meand)
        for (int i = threadIdx.x; i < LM; i += kThreads) Dbest[i] = meand[i];
}

// ---------------------------------------------------------------------------- dict
// One iteration of dict_update! (train.jl:50-72) for sequence n; the means over n
// (:76,:83) follow in k_dict_reduce / k_dict_finalize.  Columns go through the transforms B
// at a time (B even): pass 1 packs two columns' Y3 into one complex length-Lg/4 FFT and stores
// each column's half spectrum in y3s; pass 2 packs two columns' real inverse transforms
// into one complex FFT.  Accumulations over m stay in ascending m.
// LDS: twiddles | h with wrapped tail | taps of the batch | fa, fb (B x Lg) | ya, yb
// (B/2 x P) | ys (B x P) | accS, accT, accP (Lg/2+1 bins), with P = Lg/4.
template <int TAPS>
__global__ __launch_bounds__(kThreads) void k_dict(CdlDev P, int B) {
    extern __shared__ float4 smem4[];
    float *sm = reinterpret_cast<float *>(smem4);
    const int n = blockIdx.x, tid = threadIdx.x;
    const int Lg = P.Lg, M = P.M, nslot = Lg >> 2, Pq = Lg >> 2, Ph = Pq / 2 + 1;
    const int H = Lg / 2 + 1, Hp = (H + 3) & ~3;
    const int hbLen = Lg + TAPS + 4;
    float2 *tw = reinterpret_cast<float2 *>(sm);
    float *hb = sm + 2 * Lg;
    float *gtap = hb + hbLen;                       // B * TAPS (rounded to 4)
    float2 *fa = reinterpret_cast<float2 *>(gtap + ((B * TAPS + 3) & ~3));
    float2 *fb = fa + B * Lg;
    float2 *ya = fb + B * Lg;
    float2 *yb = ya + (B / 2) * Pq;
    float2 *ys = yb + (B / 2) * Pq;
    float *accS = reinterpret_cast<float *>(ys + B * Pq);
    float2 *accT = reinterpret_cast<float2 *>(accS + Hp);
    float2 *accP = accT + Hp;
    const size_t ro = (size_t)n * Lg;
    const float rslot = 1.0f / (float)nslot, rLg = 1.0f / (float)Lg, rPq = 1.0f / (float)Pq;
    float2 *y3s = P.y3s + (size_t)n * M * Ph;

    for (int l = tid; l < Lg; l += kThreads) { tw[l] = P.tw[l]; hb[l] = P.h[ro + l]; }
    for (int t = tid; t < TAPS + 4; t += kThreads) hb[Lg + t] = P.h[ro + t];
    for (int k = tid; k < Hp; k += kThreads) { accS[k] = 0.f; accT[k] = make_float2(0.f, 0.f); accP[k] = make_float2(0.f, 0.f); }

    // ---- pass 1: spectra sums over m
    for (int m0 = 0; m0 < M; m0 += B) {
        const int nb = (M - m0) < B ? (M - m0) : B;
        const int np = (nb + 1) / 2;
        for (int i = tid; i < nb * TAPS; i += kThreads) gtap[i] = P.G1rec[(size_t)(m0 + i / TAPS) * Lg + i % TAPS];
        if (nb & 1)
            for (int s = tid; s < Pq; s += kThreads) ya[(np - 1) * Pq + s].y = 0.f;
        __syncthreads();
        for (int idx = tid; idx < nb * nslot; idx += kThreads) {
            int s;
            const int bb = fdivmod(idx, nslot, rslot, s), m = m0 + bb;
            const size_t col = ((size_t)n * M + m) * Lg;
            const float4 xt = ld4(P.Xt + col, s);
            const float4 a3 = ld4(P.y3 + col, s);
            const float4 dv = ld4(P.d + col, s);
            const float4 a1 = ld4(P.g1 + col, s);
            const float4 c3 = ld4(P.g3 + col, s);
            const float4 G1v = ld4(P.G1cur + (size_t)m * Lg, s);
            const float4 G3v = ld4(P.G3 + (size_t)m * Lg, s);
            const float X0 = recon1<TAPS>(xt.x, hb, s, gtap + bb * TAPS);
            float *yaf = reinterpret_cast<float *>(ya);
            yaf[2 * ((bb >> 1) * Pq + s) + (bb & 1)] = X0 + a3.x;    // Y3 on the stride-4 support
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float qq = (f4(G1v, q) - f4(a1, q)) + (f4(G3v, q) - f4(c3, q));
                fa[bb * Lg + 4 * s + q] = make_float2(f4(dv, q), qq);
            }
        }
        __syncthreads();
        float2 *Y = team_fft<false>(ya, yb, P.plan4, tw, tid, kThreads, BlockSync{}, np);
        float2 *C = team_fft<false>(fa, fb, P.plan, tw, tid, kThreads, BlockSync{}, nb);
        // unpack the column pairs: Y3h of column 2p is the Hermitian part, of 2p+1 the other
        for (int idx = tid; idx < nb * Pq; idx += kThreads) {
            int k;
            const int bb = fdivmod(idx, Pq, rPq, k);
            const float2 Yk = Y[(bb >> 1) * Pq + k], Ym = Y[(bb >> 1) * Pq + (k == 0 ? 0 : Pq - k)];
            const float2 v = (bb & 1) ? make_float2(0.5f * (Yk.y + Ym.y), -0.5f * (Yk.x - Ym.x))
                                      : make_float2(0.5f * (Yk.x + Ym.x), 0.5f * (Yk.y - Ym.y));
            ys[bb * Pq + k] = v;
            if (k < Ph) y3s[(size_t)(m0 + bb) * Ph + k] = v;
        }
        __syncthreads();
        for (int k = tid; k < H; k += kThreads) {
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

    // ---- per-sequence spectral step: Td, G2/g2 (update_g2!, train.jl:68-71), kappa
    for (int k = tid; k < Lg; k += kThreads) fa[k] = (k < H) ? accT[k] : cconj(accT[Lg - k]);
    __syncthreads();
    float2 *R2 = team_fft<true>(fa, fb, P.plan, tw, tid, kThreads, BlockSync{}, 1);       // Td (fa[0:Lg])
    float *tdb = reinterpret_cast<float *>(ys);      // Lg floats fit in ys (B * Lg/4 complex, B >= 2)
    {
        const float scale = 1.0f / (float)Lg;
        for (int l = tid; l < Lg; l += kThreads) tdb[l] = R2[l].x * scale;
    }
    __syncthreads();
    {
        // F(G2 - g2) of the previous dictionary iteration (sherman_CDL_RHS!, cdl_kernels.jl:8)
        for (int l = tid; l < Lg; l += kThreads) fa[l] = make_float2(P.G2[ro + l] - P.g2[ro + l], 0.f);
        __syncthreads();
        float2 *R = team_fft<false>(fa, fb, P.plan, tw, tid, kThreads, BlockSync{}, 1);
        const float c1 = P.sh_cdl;
        for (int k = tid; k < H; k += kThreads) {
            const float2 Gk = R[k];
            const float den = accS[k] + c1;
            accP[k] = make_float2((c1 * Gk.x - accP[k].x) / den, (c1 * Gk.y - accP[k].y) / den);
        }
    }
    for (int l = tid; l < Lg; l += kThreads) {
        const float Td = tdb[l];
        const bool win = (l >= P.gamma) && (l < P.gamma + P.L4);
        const float b = win ? P.s[(size_t)n * P.L4 + (l - P.gamma)] : 0.f;
        const float wl = win ? 1.f : 0.f;
        const float g2o = P.g2[ro + l];
        const float G2n = (b + P.sigma * (Td + g2o)) / (wl + P.sigma);
        P.G2[ro + l] = G2n;
        P.g2[ro + l] = g2o + Td - G2n;
        P.Td[ro + l] = Td;
    }
    __syncthreads();

    // ---- pass 2: d_m = q_m + F^-1[conj(Y3h_m) kappa]; g1 += d - G1; g3 += d - G3
    const float scale = 1.0f / (float)Lg;
    for (int m0 = 0; m0 < M; m0 += B) {
        const int nb = (M - m0) < B ? (M - m0) : B;
        const int np = (nb + 1) / 2;
        for (int idx = tid; idx < nb * Ph; idx += kThreads) {
            int k;
            const int bb = idx / Ph;
            k = idx - bb * Ph;
            const float2 v = y3s[(size_t)(m0 + bb) * Ph + k];
            ys[bb * Pq + k] = v;
            if (k > 0 && Pq - k >= Ph) ys[bb * Pq + Pq - k] = cconj(v);
        }
        __syncthreads();
        for (int idx = tid; idx < np * Lg; idx += kThreads) {
            int k, kq;
            const int pp = fdivmod(idx, Lg, rLg, k);
            (void)fdivmod(k, Pq, rPq, kq);
            const float2 kap = (k < H) ? accP[k] : cconj(accP[Lg - k]);
            const float2 CA = cmul(cconj(ys[(2 * pp) * Pq + kq]), kap);
            const float2 CB = (2 * pp + 1 < nb) ? cmul(cconj(ys[(2 * pp + 1) * Pq + kq]), kap) : make_float2(0.f, 0.f);
            fa[idx] = make_float2(CA.x - CB.y, CA.y + CB.x);
        }
        __syncthreads();
        float2 *R = team_fft<true>(fa, fb, P.plan, tw, tid, kThreads, BlockSync{}, np);
        for (int idx = tid; idx < nb * nslot; idx += kThreads) {
            int s;
            const int bb = fdivmod(idx, nslot, rslot, s), m = m0 + bb;
            const size_t col = ((size_t)n * M + m) * Lg;
            const float4 G1v = ld4(P.G1cur + (size_t)m * Lg, s);
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
// (train.jl:76,83, train.jl:95), in double.
__global__ __launch_bounds__(kThreads) void k_dict_reduce(const float *__restrict__ d, const float *__restrict__ g1,
                                                          const float *__restrict__ g3, int LM, int N, int nch,
                                                          double *part) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= LM) return;
    const int n0 = (int)((long)c * N / nch), n1 = (int)((long)(c + 1) * N / nch);
    double a1 = 0.0, a3 = 0.0, ad = 0.0;
    for (int n = n0; n < n1; ++n) {
        const size_t e = (size_t)n * LM + i;
        const float dv = d[e];
        a1 += (double)(dv + g1[e]);
        a3 += (double)(dv + g3[e]);
        ad += (double)dv;
    }
    part[((size_t)c * 3 + 0) * LM + i] = a1;
    part[((size_t)c * 3 + 1) * LM + i] = a3;
    part[((size_t)c * 3 + 2) * LM + i] = ad;
}

// Column m: v1 = mean(d+g1), v2 = mean(d+g3); G3 = projection2box!(v2) (cdl_kernels.jl:37-51);
// G1new = prox_simplex_zeros!(v1) (:53-66, FP64 combination, quirk group 4*floor(l/4));
// meand = mean(d) (D_best candidate, train.jl:95); gsq = |F(G1new)|^2.
__global__ __launch_bounds__(kThreads) void k_dict_finalize(CdlDev P, const double *part, int nch, float *G1new,
                                                            float *meand, float *gsq) {
    extern __shared__ float4 smem4[];
    float *v1 = reinterpret_cast<float *>(smem4);
    float2 *fa = reinterpret_cast<float2 *>(v1 + ((P.Lg + 3) & ~3));
    float2 *fb = fa + P.Lg;
    const int m = blockIdx.x, tid = threadIdx.x, Lg = P.Lg;
    const int LM = Lg * P.M;
    const int flm = P.fl[m];
    const double invN = 1.0 / (double)P.N;
    for (int l = tid; l < Lg; l += kThreads) {
        const int i = m * Lg + l;
        double s1 = 0.0, s3 = 0.0, sd = 0.0;
        for (int c = 0; c < nch; ++c) {
            s1 += part[((size_t)c * 3 + 0) * LM + i];
            s3 += part[((size_t)c * 3 + 1) * LM + i];
            sd += part[((size_t)c * 3 + 2) * LM + i];
        }
        v1[l] = (float)(s1 * invN);
        const float v2 = (float)(s3 * invN);
        const bool sup = l < flm;
        float g3;
        if (v2 >= 1.f) g3 = sup ? 1.f : 0.f;
        else if (v2 <= 0.f) g3 = 0.f;
        else g3 = sup ? v2 : 0.f;
        P.G3[i] = g3;
        meand[i] = (float)(sd * invN);
    }
    __syncthreads();
    for (int l = tid; l < Lg; l += kThreads) {
        float g = 0.f;
        if (l < flm) {
            const int gi = 4 * ((l + 1) / 4);
            const float s4 = v1[gi] + v1[gi + 1] + v1[gi + 2] + v1[gi + 3] - 1.0f;
            g = (float)((double)v1[l] - 0.25 * (double)s4);
        }
        G1new[(size_t)m * Lg + l] = g;
        fa[l] = make_float2(g, 0.f);
    }
    __syncthreads();
    float2 *C = team_fft<false>(fa, fb, P.plan, P.tw, tid, kThreads, BlockSync{});
    for (int k = tid; k < Lg; k += kThreads) gsq[(size_t)m * Lg + k] = C[k].x * C[k].x + C[k].y * C[k].y;
}

// |F(G1[:, m])|^2 for the filters as given (constructor and reset).
__global__ __launch_bounds__(kThreads) void k_g1_spectrum(CdlDev P, float *gsq) {
    extern __shared__ float4 smem4[];
    float2 *fa = reinterpret_cast<float2 *>(smem4);
    float2 *fb = fa + P.Lg;
    const int m = blockIdx.x, tid = threadIdx.x, Lg = P.Lg;
    for (int l = tid; l < Lg; l += kThreads) fa[l] = make_float2(P.G1cur[(size_t)m * Lg + l], 0.f);
    __syncthreads();
    float2 *C = team_fft<false>(fa, fb, P.plan, P.tw, tid, kThreads, BlockSync{});
    for (int k = tid; k < Lg; k += kThreads) gsq[(size_t)m * Lg + k] = C[k].x * C[k].x + C[k].y * C[k].y;
}

// s_k = sum_m |G1h|^2 (FP32, m ascending; train.jl:18-19), c_k = s_k + sherman_const_csc.
__global__ __launch_bounds__(kThreads) void k_cs(const float *gsq, int Lg, int M, float c0, float *sS, float *cS) {
    const int k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= Lg) return;
    float a = 0.f;
    for (int m = 0; m < M; ++m) a += gsq[(size_t)m * Lg + k];
    sS[k] = a;
    cS[k] = a + c0;
}

// ---------------------------------------------------------------------------- dispatch helpers
template <template <int> class K>
struct TapsDispatch;

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

// ============================================================================ host side
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

struct cdlm_problem {
    cdlm_ctx *ctx = nullptr;
    int L4 = 0, N = 0, M = 0, Lg = 0, gamma = 0, L = 0, taps = 0;
    float rho = 0, sigma = 0, lambda = 0, sh_csc = 2, sh_cdl = 2;
    uint64_t seed = 0;
    std::vector<int> fl;
    std::vector<float> D0m;  // masked initial filters (Lg, M)
    std::vector<DevBuf> bufs;
    size_t total_bytes = 0;
    float *s = nullptr, *X0 = nullptr;
    float *Xt = nullptr, *y2 = nullptr, *y3 = nullptr, *d = nullptr, *g1 = nullptr, *g3 = nullptr;
    float *h = nullptr, *y1 = nullptr, *Y1 = nullptr, *DX = nullptr, *G2 = nullptr, *g2 = nullptr, *Td = nullptr;
    float *G1buf[3] = {nullptr, nullptr, nullptr};
    int g1cur = 0, g1rec = 0;
    float *G3 = nullptr, *meand = nullptr, *Dbest = nullptr, *gsq = nullptr, *cS = nullptr, *sS = nullptr;
    float *xs[2] = {nullptr, nullptr};
    float2 *y3s = nullptr;
    int *fl_dev = nullptr;
    float2 *tw = nullptr;
    double *partials = nullptr, *dpart = nullptr, *trace = nullptr;
    int nch = 1;
    int trace_cap = 0;
    CdlScal *sc = nullptr;
    FftPlan plan, plan4;
    size_t lds_init = 0, lds_csc = 0, lds_best = 0, lds_dict = 0, lds_fin = 0, lds_g1 = 0;
    int dict_batch = 1;
    int nsl = 4;
    bool timing = false;
    struct Ev { hipEvent_t a, b; int count; };
    std::vector<Ev> ev[4];
    // resident CSC (k_csc_res): groups of G co-resident work-groups per sequence
    bool res = false;
    int resG = 1, resNG = 0, resMg = 0;
    size_t lds_res = 0;
    unsigned long long epoch = 0;
    unsigned long long *rflags = nullptr;
    float *rxbuf = nullptr;
    int *rerr = nullptr;
};

namespace {

cdlm_status palloc(cdlm_problem *p, void **out, size_t bytes) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, bytes ? bytes : 4);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return cdlm_fail(p->ctx, CDLM_ERR_OOM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    }
    p->bufs.push_back({q, bytes});
    p->total_bytes += bytes;
    *out = q;
    return CDLM_OK;
}

template <typename T>
cdlm_status palloc_t(cdlm_problem *p, T **out, size_t count) {
    void *q = nullptr;
    cdlm_status st = palloc(p, &q, count * sizeof(T));
    *out = static_cast<T *>(q);
    return st;
}

CdlDev dev_view(const cdlm_problem *p) {
    CdlDev P;
    P.L4 = p->L4; P.N = p->N; P.M = p->M; P.Lg = p->Lg; P.gamma = p->gamma; P.L = p->L;
    P.rho = p->rho; P.sigma = p->sigma; P.lambda = p->lambda; P.sh_csc = p->sh_csc; P.sh_cdl = p->sh_cdl;
    P.s = p->s;
    P.Xt = p->Xt; P.y2 = p->y2; P.y3 = p->y3; P.d = p->d; P.g1 = p->g1; P.g3 = p->g3;
    P.h = p->h; P.y1 = p->y1; P.Y1 = p->Y1; P.DX = p->DX; P.G2 = p->G2; P.g2 = p->g2; P.Td = p->Td;
    P.G1cur = p->G1buf[p->g1cur]; P.G1rec = p->G1buf[p->g1rec]; P.G3 = p->G3;
    P.fl = p->fl_dev; P.cS = p->cS; P.sS = p->sS; P.tw = p->tw; P.y3s = p->y3s;
    P.plan = p->plan; P.plan4 = p->plan4;
    return P;
}

size_t lds_bytes_csc(int Lg, int taps) {
    return sizeof(float) * (2 * (size_t)Lg + (size_t)(Lg + taps + 4) + kWaves * (size_t)(Lg + taps) + kWaves * (size_t)Lg);
}
size_t lds_bytes_init(int Lg, int taps) { return sizeof(float) * (kWaves * (size_t)(Lg + taps) + kWaves * (size_t)Lg); }
size_t lds_bytes_best(int Lg, int taps) { return sizeof(float) * (size_t)(Lg + taps + 4); }
size_t lds_bytes_dict(int Lg, int taps, int B) {
    const int H = Lg / 2 + 1, Hp = (H + 3) & ~3;
    const size_t Pq = (size_t)Lg / 4;
    return sizeof(float) * (2 * (size_t)Lg + (size_t)(Lg + taps + 4) + (size_t)((B * taps + 3) & ~3) + 4 * (size_t)B * Lg +
                            2 * 2 * (size_t)(B / 2) * Pq + 2 * (size_t)B * Pq + 5 * (size_t)Hp);
}
size_t lds_bytes_fin(int Lg) { return sizeof(float) * (((size_t)Lg + 3) / 4 * 4 + 4 * (size_t)Lg); }
constexpr size_t kLdsLimit = 160 * 1024;

cdlm_status set_lds_limits(cdlm_problem *p) {
    // Dynamic LDS above the default needs hipFuncSetAttribute; each kernel gets exactly the
    // size it is launched with (the attribute also counts against the kernel's static LDS).
#define SETA(T)                                                                                            \
    do {                                                                                                   \
        CDLM_HIP(p->ctx, hipFuncSetAttribute((const void *)k_init<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds_init)); \
        CDLM_HIP(p->ctx, hipFuncSetAttribute((const void *)k_csc<T, 4>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds_csc));  \
        CDLM_HIP(p->ctx, hipFuncSetAttribute((const void *)k_csc<T, 8>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds_csc));  \
        CDLM_HIP(p->ctx, hipFuncSetAttribute((const void *)k_best<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds_best)); \
        CDLM_HIP(p->ctx, hipFuncSetAttribute((const void *)k_dict<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds_dict)); \
        if (p->res) CDLM_HIP(p->ctx, hipFuncSetAttribute((const void *)k_csc_res<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds_res)); \
    } while (0)
    switch (p->taps) {
    case 4: SETA(4); break;
    case 8: SETA(8); break;
    case 12: SETA(12); break;
    case 16: SETA(16); break;
    case 20: SETA(20); break;
    case 24: SETA(24); break;
    case 28: SETA(28); break;
    case 32: SETA(32); break;
    default: return cdlm_fail(p->ctx, CDLM_ERR_UNSUPPORTED, "filter length %d unsupported", p->taps);
    }
#undef SETA
    CDLM_HIP(p->ctx, hipFuncSetAttribute((const void *)k_dict_finalize, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds_fin));
    CDLM_HIP(p->ctx, hipFuncSetAttribute((const void *)k_g1_spectrum, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds_g1));
    return CDLM_OK;
}

// Event pair around one launch when timing is on.
struct TimedLaunch {
    cdlm_problem *p;
    int kid;
    int count;
    hipEvent_t a = nullptr, b = nullptr;
    TimedLaunch(cdlm_problem *p_, int kid_, int count_ = 1) : p(p_), kid(kid_), count(count_) {
        if (p->timing && hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess)
            (void)hipEventRecord(a, p->ctx->stream);
        else if (p->timing) { (void)hipGetLastError(); a = b = nullptr; }
    }
    ~TimedLaunch() {
        if (a && b) {
            (void)hipEventRecord(b, p->ctx->stream);
            p->ev[kid].push_back({a, b, count});
        }
    }
};

void clear_events(cdlm_problem *p) {
    for (auto &v : p->ev) {
        for (auto &e : v) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
        v.clear();
    }
}

cdlm_status launch_check(cdlm_problem *p, const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return cdlm_fail(p->ctx, CDLM_ERR_HIP, "launch of %s failed: %s", what, hipGetErrorString(e));
    return CDLM_OK;
}

cdlm_status run_cs(cdlm_problem *p) {
    CdlDev P = dev_view(p);
    hipStream_t st = p->ctx->stream;
    hipLaunchKernelGGL(k_g1_spectrum, dim3(p->M), dim3(kThreads), p->lds_g1, st, P, p->gsq);
    hipLaunchKernelGGL(k_cs, dim3((p->Lg + kThreads - 1) / kThreads), dim3(kThreads), 0, st, p->gsq, p->Lg, p->M,
                       p->sh_csc, p->sS, p->cS);
    return launch_check(p, "k_g1_spectrum/k_cs");
}

// k_best + reduction + decision; slot < 0: no trace entry.
cdlm_status run_best(cdlm_problem *p, int force, int slot, float *Xfull) {
    CdlDev P = dev_view(p);
    hipStream_t st = p->ctx->stream;
    {
        TimedLaunch tl(p, 2);
#define L_BEST(T) hipLaunchKernelGGL(k_best<T>, dim3(p->N), dim3(kThreads), p->lds_best, st, P, p->xs[0], p->xs[1], p->sc, p->partials, Xfull)
        CDLM_TAPS_SWITCH(p->taps, L_BEST);
#undef L_BEST
    }
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kThreads), 0, st, p->partials, p->N, p->sc);
    hipLaunchKernelGGL(k_decide, dim3(1), dim3(kThreads), 0, st, p->sc, p->lambda, force,
                       slot >= 0 ? p->trace : nullptr, slot, p->meand, p->Dbest, p->Lg * p->M);
    return launch_check(p, "k_best");
}

cdlm_status run_csc(cdlm_problem *p) {
    CdlDev P = dev_view(p);
    {
        TimedLaunch tl(p, 0);
#define L_CSC(T)                                                                                                   \
    if (p->nsl == 4) hipLaunchKernelGGL((k_csc<T, 4>), dim3(p->N), dim3(kThreads), p->lds_csc, p->ctx->stream, P); \
    else hipLaunchKernelGGL((k_csc<T, 8>), dim3(p->N), dim3(kThreads), p->lds_csc, p->ctx->stream, P)
        CDLM_TAPS_SWITCH(p->taps, L_CSC);
#undef L_CSC
    }
    p->g1rec = p->g1cur;
    return launch_check(p, "k_csc");
}

// K CSC iterations in one cooperative launch of the resident kernel.
cdlm_status run_csc_res(cdlm_problem *p, int K) {
    if (K <= 0) return CDLM_OK;
    CdlDev P = dev_view(p);
    ResArgs A;
    A.K = K; A.G = p->resG; A.NG = p->resNG; A.Mg = p->resMg;
    A.epoch = ++p->epoch;
    A.flags = p->rflags; A.xbuf = p->rxbuf; A.err = p->rerr;
    void *args[] = {(void *)&P, (void *)&A};
    hipError_t e = hipSuccess;
    {
        TimedLaunch tl(p, 0, K);
#define L_RES(T) e = hipLaunchCooperativeKernel((const void *)k_csc_res<T>, dim3(p->resG * p->resNG), dim3(kResThreads), args, (unsigned)p->lds_res, p->ctx->stream)
        CDLM_TAPS_SWITCH(p->taps, L_RES);
#undef L_RES
    }
    if (e != hipSuccess) return cdlm_fail(p->ctx, CDLM_ERR_HIP, "cooperative launch of k_csc_res: %s", hipGetErrorString(e));
    p->g1rec = p->g1cur;
    return launch_check(p, "k_csc_res");
}

// CSC iterations through whichever kernel the shape uses.
cdlm_status run_csc_n(cdlm_problem *p, int K) {
    if (p->res) return run_csc_res(p, K);
    for (int i = 0; i < K; ++i) {
        cdlm_status st = run_csc(p);
        if (st != CDLM_OK) return st;
    }
    return CDLM_OK;
}

// Reports a timed-out exchange of the resident kernel (set on the device).
cdlm_status check_res_err(cdlm_problem *p) {
    if (!p->res) return CDLM_OK;
    int err = 0;
    CDLM_HIP(p->ctx, hipMemcpy(&err, p->rerr, sizeof(int), hipMemcpyDeviceToHost));
    if (err) return cdlm_fail(p->ctx, CDLM_ERR_STATE, "resident CSC exchange timed out (co-residency lost)");
    return CDLM_OK;
}

cdlm_status run_dict(cdlm_problem *p) {
    CdlDev P = dev_view(p);
    hipStream_t st = p->ctx->stream;
    {
        TimedLaunch tl(p, 1);
#define L_DICT(T) hipLaunchKernelGGL(k_dict<T>, dim3(p->N), dim3(kThreads), p->lds_dict, st, P, p->dict_batch)
        CDLM_TAPS_SWITCH(p->taps, L_DICT);
#undef L_DICT
    }
    const int LM = p->Lg * p->M;
    {
        TimedLaunch tl(p, 3);
        hipLaunchKernelGGL(k_dict_reduce, dim3((LM + kThreads - 1) / kThreads, p->nch), dim3(kThreads), 0, st, p->d,
                           p->g1, p->g3, LM, p->N, p->nch, p->dpart);
    }
    int nxt = 0;
    while (nxt == p->g1cur || nxt == p->g1rec) ++nxt;
    hipLaunchKernelGGL(k_dict_finalize, dim3(p->M), dim3(kThreads), p->lds_fin, st, P, p->dpart, p->nch,
                       p->G1buf[nxt], p->meand, p->gsq);
    hipLaunchKernelGGL(k_cs, dim3((p->Lg + kThreads - 1) / kThreads), dim3(kThreads), 0, st, p->gsq, p->Lg, p->M,
                       p->sh_csc, p->sS, p->cS);
    p->g1cur = nxt;
    return launch_check(p, "k_dict");
}

cdlm_status do_reset(cdlm_problem *p) {
    hipStream_t st = p->ctx->stream;
    const size_t LM = (size_t)p->Lg * p->M;
    p->g1cur = p->g1rec = 0;
    CDLM_HIP(p->ctx, hipMemcpyAsync(p->G1buf[0], p->D0m.data(), LM * sizeof(float), hipMemcpyHostToDevice, st));
    CDLM_HIP(p->ctx, hipMemcpyAsync(p->G3, p->D0m.data(), LM * sizeof(float), hipMemcpyHostToDevice, st));
    CDLM_HIP(p->ctx, hipMemcpyAsync(p->meand, p->D0m.data(), LM * sizeof(float), hipMemcpyHostToDevice, st));
    CDLM_HIP(p->ctx, hipMemcpyAsync(p->Dbest, p->D0m.data(), LM * sizeof(float), hipMemcpyHostToDevice, st));
    CdlScal sc0;
    memset(&sc0, 0, sizeof(sc0));
    sc0.obj_best = INFINITY;
    CDLM_HIP(p->ctx, hipMemcpyAsync(p->sc, &sc0, sizeof(sc0), hipMemcpyHostToDevice, st));
    cdlm_status s = run_cs(p);
    if (s != CDLM_OK) return s;
    CdlDev P = dev_view(p);
#define L_INIT(T) hipLaunchKernelGGL(k_init<T>, dim3(p->N), dim3(kThreads), p->lds_init, st, P, p->X0, p->seed)
    CDLM_TAPS_SWITCH(p->taps, L_INIT);
#undef L_INIT
    s = launch_check(p, "k_init");
    if (s != CDLM_OK) return s;
    return run_best(p, 1, -1, nullptr);
}

// rand_probvec groups (CDL_setup.jl:59-68,183) from a 64-bit seed, on the host.
void host_rand_probvec(std::vector<float> &D, int Lg, int M, uint64_t seed) {
    uint64_t x = seed * 0x9e3779b97f4a7c15ULL + 12345;
    auto next = [&]() {
        x += 0x9e3779b97f4a7c15ULL;
        uint64_t z = x;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    for (int m = 0; m < M; ++m)
        for (int q = 0; q < Lg / 4; ++q) {
            double u[3] = {next(), next(), next()};
            if (u[0] > u[1]) std::swap(u[0], u[1]);
            if (u[1] > u[2]) std::swap(u[1], u[2]);
            if (u[0] > u[1]) std::swap(u[0], u[1]);
            float *g = &D[(size_t)m * Lg + 4 * q];
            g[0] = (float)u[0]; g[1] = (float)(u[1] - u[0]); g[2] = (float)(u[2] - u[1]); g[3] = (float)(1.0 - u[2]);
        }
}

cdlm_status sync_stream(cdlm_problem *p) {
    CDLM_HIP(p->ctx, hipStreamSynchronize(p->ctx->stream));
    return CDLM_OK;
}

}  // namespace

// ============================================================================ C ABI
extern "C" size_t cdlm_footprint_bytes(int32_t L4, int32_t N, int32_t M, int32_t gamma, int32_t keep_x0) {
    const size_t Lg = (size_t)L4 + 2 * (size_t)gamma;
    const size_t E = Lg * M * (size_t)N, LN = Lg * N, LM = Lg * M;
    const size_t L = (size_t)L4 / 4;
    size_t b = 0;
    b += (size_t)L4 * N * 4;                  // s
    b += 6 * E * 4;                           // Xt, y2, y3, d, g1, g3
    if (keep_x0) b += E * 4;
    b += 7 * LN * 4;                          // h, y1, Y1, DX, G2, g2, Td
    b += 2 * L * M * N * 4;                   // X_best slices
    b += (size_t)M * N * (Lg / 8 + 1) * 8;    // dictionary Y3 half spectra
    b += (3 + 4) * LM * 4 + 2 * Lg * 4;       // G1 x3, G3, meand, Dbest, gsq, cS, sS
    b += (size_t)N * kPartStride * 8;
    const size_t blocks = (LM + kThreads - 1) / kThreads;
    size_t nch = (2048 + blocks - 1) / blocks;
    if (nch < 1) nch = 1;
    if (nch > (size_t)N) nch = N;
    b += nch * 3 * LM * 8;
    return b;
}

extern "C" cdlm_status cdlm_problem_create(cdlm_ctx *ctx, const cdlm_params *prm, const float *s, int32_t L4, int32_t N,
                                           const float *D0, const float *X0, uint64_t seed, cdlm_problem **out) {
    if (!ctx) return CDLM_ERR_ARG;
    if (!prm || !s || !out || !prm->filter_len) return cdlm_fail(ctx, CDLM_ERR_ARG, "null argument");
    *out = nullptr;
    if (L4 <= 0 || L4 % 4 != 0) return cdlm_fail(ctx, CDLM_ERR_ARG, "first dimension of dat must be a positive multiple of 4 (got %d)", L4);
    if (N <= 0) return cdlm_fail(ctx, CDLM_ERR_ARG, "N must be positive (got %d)", N);
    if (prm->M <= 0) return cdlm_fail(ctx, CDLM_ERR_ARG, "M must be positive");
    std::unique_ptr<cdlm_problem> p(new cdlm_problem());
    p->ctx = ctx;
    p->L4 = L4; p->N = N; p->M = prm->M; p->L = L4 / 4;
    p->fl.assign(prm->filter_len, prm->filter_len + prm->M);
    int gamma = 0;
    for (int f : p->fl) {
        if (f <= 0 || f % 4 != 0) return cdlm_fail(ctx, CDLM_ERR_ARG, "filter lengths must be positive multiples of 4 (got %d)", f);
        gamma = f > gamma ? f : gamma;
    }
    if (gamma > kMaxTaps) return cdlm_fail(ctx, CDLM_ERR_UNSUPPORTED, "max filter length %d > %d", gamma, kMaxTaps);
    p->gamma = gamma;
    p->taps = gamma;
    p->Lg = L4 + 2 * gamma;
    p->rho = prm->rho; p->sigma = prm->sigma; p->lambda = prm->lambda;
    p->sh_csc = prm->sherman_csc; p->sh_cdl = prm->sherman_cdl;
    if (!(p->sh_csc > 0.f) || !(p->sh_cdl > 0.f)) return cdlm_fail(ctx, CDLM_ERR_ARG, "sherman constants must be positive");
    p->seed = seed;
    const int Lg = p->Lg, M = p->M;
    if (!make_fft_plan(Lg, 1, &p->plan) || !make_fft_plan(Lg / 4, 4, &p->plan4))
        return cdlm_fail(ctx, CDLM_ERR_UNSUPPORTED, "no FFT plan for length %d", Lg);
    p->lds_csc = lds_bytes_csc(Lg, gamma);
    p->lds_init = lds_bytes_init(Lg, gamma);
    p->lds_best = lds_bytes_best(Lg, gamma);
    if (Lg > 2048) return cdlm_fail(ctx, CDLM_ERR_UNSUPPORTED, "padded length %d > 2048", Lg);
    p->nsl = (Lg / 4 <= 256) ? 4 : 8;
    {
        // columns per batched FFT pass: most columns in flight per CU
        int best = 1;
        long bestv = 0;
        for (int B : {2, 4}) {
            const size_t b = lds_bytes_dict(Lg, gamma, B);
            if (b > kLdsLimit) continue;
            const long v = (long)B * (long)(kLdsLimit / b);
            if (v > bestv) { bestv = v; best = B; }
        }
        p->dict_batch = best;
    }
    p->lds_dict = lds_bytes_dict(Lg, gamma, p->dict_batch);
    p->lds_fin = lds_bytes_fin(Lg);
    p->lds_g1 = sizeof(float2) * 2 * (size_t)Lg;
    {
        // resident CSC: one column per wave, 16 waves per work-group, ceil(M/16) work-groups per sequence
        const char *env = getenv("CDLM_CSC_RESIDENT");
        int coop = 0;
        (void)hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, ctx->device);
        // opt-in: measured slower than the streaming kernel at N=20000, M=64, Lg=840 (7.55 vs
        // 5.44 ms per iteration) because one sequence per CU exposes the per-iteration FFT and
        // exchange latency
        p->res = coop && (Lg / 4 <= 256) && (env && env[0] == '1');
        if (p->res) {
            p->resG = (M + kResWaves - 1) / kResWaves;
            p->resMg = (M + p->resG - 1) / p->resG;
            p->lds_res = sizeof(float) * (27 * (size_t)Lg + 17 * (size_t)gamma + 4 + kResWaves * 2 * (size_t)gamma);
            if (p->lds_res > kLdsLimit) p->res = false;
        }
    }
    size_t mx = p->lds_csc;
    for (size_t v : {p->lds_init, p->lds_best, p->lds_dict, p->lds_fin, p->lds_g1}) mx = v > mx ? v : mx;
    if (mx > kLdsLimit - kThreads * sizeof(double)) return cdlm_fail(ctx, CDLM_ERR_UNSUPPORTED, "padded length %d needs %zu B of LDS", Lg, mx);
    cdlm_status st = set_lds_limits(p.get());
    if (st != CDLM_OK) return st;
    if (p->res) {
        int nb = 0;
        hipError_t oe = hipErrorUnknown;
#define OCC(T) oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_csc_res<T>, kResThreads, p->lds_res)
        CDLM_TAPS_SWITCH(p->taps, OCC);
#undef OCC
        if (oe != hipSuccess || nb < 1) {
            (void)hipGetLastError();
            p->res = false;
        } else {
            int ng = (nb * ctx->num_cu) / p->resG;
            if (ng >= 8) ng -= ng % 8;              // members of a group on one XCD
            const int cap = ((N + 7) / 8) * 8;
            if (ng > cap) ng = cap;
            p->resNG = ng;
            if (ng < 1) p->res = false;
        }
    }

    // host-side masked D0 (P_D .* D, CDL_setup.jl:183)
    const size_t LM = (size_t)Lg * M, E = LM * N, LN = (size_t)Lg * N;
    p->D0m.assign(LM, 0.f);
    if (D0) std::memcpy(p->D0m.data(), D0, LM * sizeof(float));
    else host_rand_probvec(p->D0m, Lg, M, seed);
    for (int m = 0; m < M; ++m)
        for (int l = p->fl[m]; l < Lg; ++l) p->D0m[(size_t)m * Lg + l] = 0.f;

    const size_t blocks = (LM + kThreads - 1) / kThreads;
    int nch = (int)((2048 + blocks - 1) / blocks);
    if (nch < 1) nch = 1;
    if (nch > N) nch = N;
    p->nch = nch;
    p->trace_cap = 4096;

#define A(ptr, cnt) do { st = palloc_t(p.get(), &(ptr), (cnt)); if (st != CDLM_OK) return st; } while (0)
    A(p->s, (size_t)L4 * N);
    A(p->Xt, E); A(p->y2, E); A(p->y3, E); A(p->d, E); A(p->g1, E); A(p->g3, E);
    if (X0) A(p->X0, E);
    A(p->h, LN); A(p->y1, LN); A(p->Y1, LN); A(p->DX, LN); A(p->G2, LN); A(p->g2, LN); A(p->Td, LN);
    A(p->G1buf[0], LM); A(p->G1buf[1], LM); A(p->G1buf[2], LM);
    A(p->G3, LM); A(p->meand, LM); A(p->Dbest, LM); A(p->gsq, LM);
    A(p->cS, (size_t)Lg); A(p->sS, (size_t)Lg);
    A(p->xs[0], (size_t)p->L * M * N); A(p->xs[1], (size_t)p->L * M * N);
    A(p->y3s, (size_t)M * N * (Lg / 8 + 1));
    if (p->res) {
        A(p->rflags, (size_t)p->resNG * p->resG);
        A(p->rxbuf, (size_t)p->resNG * p->resG * 2 * Lg);
        A(p->rerr, 1);
    }
    A(p->fl_dev, (size_t)M);
    A(p->tw, (size_t)Lg);
    A(p->partials, (size_t)N * kPartStride);
    A(p->dpart, (size_t)nch * 3 * LM);
    A(p->trace, (size_t)p->trace_cap);
    A(p->sc, 1);
#undef A
    hipStream_t hs = ctx->stream;
    CDLM_HIP(ctx, hipMemcpyAsync(p->s, s, (size_t)L4 * N * sizeof(float), hipMemcpyHostToDevice, hs));
    if (X0) CDLM_HIP(ctx, hipMemcpyAsync(p->X0, X0, E * sizeof(float), hipMemcpyHostToDevice, hs));
    CDLM_HIP(ctx, hipMemcpyAsync(p->fl_dev, p->fl.data(), M * sizeof(int), hipMemcpyHostToDevice, hs));
    if (p->res) {
        CDLM_HIP(ctx, hipMemsetAsync(p->rflags, 0, sizeof(unsigned long long) * p->resNG * p->resG, hs));
        CDLM_HIP(ctx, hipMemsetAsync(p->rerr, 0, sizeof(int), hs));
    }
    std::vector<float2> tw(Lg);
    for (int t = 0; t < Lg; ++t) {
        const double a = -2.0 * M_PI * (double)t / (double)Lg;
        tw[t] = make_float2((float)cos(a), (float)sin(a));
    }
    CDLM_HIP(ctx, hipMemcpyAsync(p->tw, tw.data(), Lg * sizeof(float2), hipMemcpyHostToDevice, hs));
    st = do_reset(p.get());
    if (st != CDLM_OK) return st;
    st = sync_stream(p.get());
    if (st != CDLM_OK) return st;
    *out = p.release();
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_problem_destroy(cdlm_problem *p) {
    if (!p) return CDLM_OK;
    (void)hipStreamSynchronize(p->ctx->stream);
    clear_events(p);
    for (auto &b : p->bufs) (void)hipFree(b.p);
    delete p;
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_problem_reset(cdlm_problem *p) {
    if (!p) return CDLM_ERR_ARG;
    cdlm_status st = do_reset(p);
    if (st != CDLM_OK) return st;
    return sync_stream(p);
}

extern "C" cdlm_status cdlm_problem_dims(const cdlm_problem *p, int32_t *Lg, int32_t *gamma, int32_t *L, int32_t *N,
                                         int32_t *M) {
    if (!p) return CDLM_ERR_ARG;
    if (Lg) *Lg = p->Lg;
    if (gamma) *gamma = p->gamma;
    if (L) *L = p->L;
    if (N) *N = p->N;
    if (M) *M = p->M;
    return CDLM_OK;
}

extern "C" size_t cdlm_problem_bytes(const cdlm_problem *p) { return p ? p->total_bytes : 0; }

extern "C" cdlm_status cdlm_code_update(cdlm_problem *p, int32_t num_iter) {
    if (!p) return CDLM_ERR_ARG;
    if (num_iter < 0) return cdlm_fail(p->ctx, CDLM_ERR_ARG, "num_iter < 0");
    cdlm_status st = run_csc_n(p, num_iter);
    if (st != CDLM_OK) return st;
    st = sync_stream(p);
    if (st != CDLM_OK) return st;
    return check_res_err(p);
}

extern "C" cdlm_status cdlm_dict_update(cdlm_problem *p, int32_t num_iter) {
    if (!p) return CDLM_ERR_ARG;
    if (num_iter < 0) return cdlm_fail(p->ctx, CDLM_ERR_ARG, "num_iter < 0");
    for (int i = 0; i < num_iter; ++i) {
        cdlm_status st = run_dict(p);
        if (st != CDLM_OK) return st;
    }
    return sync_stream(p);
}

extern "C" cdlm_status cdlm_update_best(cdlm_problem *p, double *obj, int32_t *improved) {
    if (!p) return CDLM_ERR_ARG;
    cdlm_status st = run_best(p, 0, -1, nullptr);
    if (st != CDLM_OK) return st;
    CdlScal sc;
    CDLM_HIP(p->ctx, hipMemcpyAsync(&sc, p->sc, sizeof(sc), hipMemcpyDeviceToHost, p->ctx->stream));
    st = sync_stream(p);
    if (st != CDLM_OK) return st;
    if (obj) *obj = sc.red[5];
    if (improved) *improved = sc.improved;
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_objective(cdlm_problem *p, double *obj) {
    if (!p || !obj) return CDLM_ERR_ARG;
    CdlDev P = dev_view(p);
    hipStream_t st = p->ctx->stream;
    // candidate slice buffer is scratch here; the decision kernel is not run
#define L_BEST(T) hipLaunchKernelGGL(k_best<T>, dim3(p->N), dim3(kThreads), p->lds_best, st, P, p->xs[0], p->xs[1], p->sc, p->partials, (float *)nullptr)
    CDLM_TAPS_SWITCH(p->taps, L_BEST);
#undef L_BEST
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kThreads), 0, st, p->partials, p->N, p->sc);
    cdlm_status s = launch_check(p, "objective");
    if (s != CDLM_OK) return s;
    CdlScal sc;
    CDLM_HIP(p->ctx, hipMemcpyAsync(&sc, p->sc, sizeof(sc), hipMemcpyDeviceToHost, st));
    s = sync_stream(p);
    if (s != CDLM_OK) return s;
    const float nrm = (float)sqrt(sc.red[0]);
    const float nrm2 = nrm * nrm;
    const float reg = p->lambda * (float)sc.red[1];
    *obj = 0.5 * (double)nrm2 + (double)reg;
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_optimize(cdlm_problem *p, int32_t iter_num, int32_t csc_iter, int32_t cdl_iter,
                                     double *obj_trace) {
    if (!p) return CDLM_ERR_ARG;
    if (iter_num < 0 || csc_iter < 0 || cdl_iter < 0) return cdlm_fail(p->ctx, CDLM_ERR_ARG, "negative iteration count");
    if (iter_num > p->trace_cap) return cdlm_fail(p->ctx, CDLM_ERR_ARG, "iter_num > %d", p->trace_cap);
    cdlm_status st;
    for (int it = 0; it < iter_num; ++it) {
        if ((st = run_csc_n(p, csc_iter)) != CDLM_OK) return st;
        if ((st = run_best(p, 0, it, nullptr)) != CDLM_OK) return st;
        for (int i = 0; i < cdl_iter; ++i) if ((st = run_dict(p)) != CDLM_OK) return st;
    }
    if (obj_trace && iter_num > 0)
        CDLM_HIP(p->ctx, hipMemcpyAsync(obj_trace, p->trace, iter_num * sizeof(double), hipMemcpyDeviceToHost, p->ctx->stream));
    st = sync_stream(p);
    if (st != CDLM_OK) return st;
    return check_res_err(p);
}

extern "C" cdlm_status cdlm_code_thresh(cdlm_problem *p, double mult, double *thresh, int64_t *count) {
    if (!p) return CDLM_ERR_ARG;
    CdlScal sc;
    CDLM_HIP(p->ctx, hipMemcpyAsync(&sc, p->sc, sizeof(sc), hipMemcpyDeviceToHost, p->ctx->stream));
    cdlm_status st = sync_stream(p);
    if (st != CDLM_OK) return st;
    const double cnt = sc.best_stats[0], s1 = sc.best_stats[1], s2 = sc.best_stats[2];
    if (count) *count = (int64_t)cnt;
    if (thresh) {
        if (cnt < 2) *thresh = NAN;
        else {
            double var = (s2 - s1 * s1 / cnt) / (cnt - 1.0);
            if (var < 0) var = 0;
            *thresh = mult * (double)(float)sqrt(var);
        }
    }
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_obj_best(cdlm_problem *p, float *obj_best) {
    if (!p || !obj_best) return CDLM_ERR_ARG;
    CdlScal sc;
    CDLM_HIP(p->ctx, hipMemcpyAsync(&sc, p->sc, sizeof(sc), hipMemcpyDeviceToHost, p->ctx->stream));
    cdlm_status st = sync_stream(p);
    if (st != CDLM_OK) return st;
    *obj_best = sc.obj_best;
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_get(cdlm_problem *p, int32_t field, void *dst, size_t bytes) {
    if (!p || !dst) return CDLM_ERR_ARG;
    const size_t E = (size_t)p->Lg * p->M * p->N, LN = (size_t)p->Lg * p->N, LM = (size_t)p->Lg * p->M;
    const void *src = nullptr;
    size_t cnt = 0;
    hipStream_t st = p->ctx->stream;
    switch (field) {
    case CDLM_F_X: {
        if (bytes != E * sizeof(float)) return cdlm_fail(p->ctx, CDLM_ERR_ARG, "field X needs %zu bytes", E * sizeof(float));
        float *tmp = nullptr;
        hipError_t e = hipMalloc(&tmp, E * sizeof(float));
        if (e != hipSuccess) { (void)hipGetLastError(); return cdlm_fail(p->ctx, CDLM_ERR_OOM, "no memory for X"); }
        CdlDev P = dev_view(p);
        // the candidate slice is rewritten with the same values; partials are scratch
#define L_BEST(T) hipLaunchKernelGGL(k_best<T>, dim3(p->N), dim3(kThreads), p->lds_best, st, P, p->xs[0], p->xs[1], p->sc, p->partials, tmp)
        CDLM_TAPS_SWITCH(p->taps, L_BEST);
#undef L_BEST
        hipError_t e2 = hipMemcpyAsync(dst, tmp, bytes, hipMemcpyDeviceToHost, st);
        hipError_t e3 = hipStreamSynchronize(st);
        (void)hipFree(tmp);
        if (e2 != hipSuccess || e3 != hipSuccess) return cdlm_fail(p->ctx, CDLM_ERR_HIP, "copy of X failed");
        return CDLM_OK;
    }
    case CDLM_F_DX: src = p->DX; cnt = LN; break;
    case CDLM_F_D: src = p->d; cnt = E; break;
    case CDLM_F_G1: src = p->G1buf[p->g1cur]; cnt = LM; break;
    case CDLM_F_G2: src = p->G2; cnt = LN; break;
    case CDLM_F_G3: src = p->G3; cnt = LM; break;
    case CDLM_F_XBEST_SLICE: {
        CdlScal sc;
        CDLM_HIP(p->ctx, hipMemcpyAsync(&sc, p->sc, sizeof(sc), hipMemcpyDeviceToHost, st));
        CDLM_HIP(p->ctx, hipStreamSynchronize(st));
        src = p->xs[sc.best_idx];
        cnt = (size_t)p->L * p->M * p->N;
        break;
    }
    case CDLM_F_DBEST: src = p->Dbest; cnt = LM; break;
    case CDLM_F_Y1: src = p->Y1; cnt = LN; break;
    case CDLM_F_y1: src = p->y1; cnt = LN; break;
    case CDLM_F_y2: src = p->y2; cnt = E; break;
    case CDLM_F_y3: src = p->y3; cnt = E; break;
    case CDLM_F_g1: src = p->g1; cnt = E; break;
    case CDLM_F_g2: src = p->g2; cnt = LN; break;
    case CDLM_F_g3: src = p->g3; cnt = E; break;
    case CDLM_F_TD: src = p->Td; cnt = LN; break;
    default: return cdlm_fail(p->ctx, CDLM_ERR_ARG, "unknown field %d", field);
    }
    if (bytes != cnt * sizeof(float)) return cdlm_fail(p->ctx, CDLM_ERR_ARG, "field %d needs %zu bytes, got %zu", field, cnt * sizeof(float), bytes);
    CDLM_HIP(p->ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
    return sync_stream(p);
}

extern "C" cdlm_status cdlm_timing_enable(cdlm_problem *p, int32_t enable) {
    if (!p) return CDLM_ERR_ARG;
    (void)hipStreamSynchronize(p->ctx->stream);
    clear_events(p);
    p->timing = enable != 0;
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_timing_read(cdlm_problem *p, int32_t kernel, double *total_ms, int64_t *launches) {
    if (!p || kernel < 0 || kernel > 3) return CDLM_ERR_ARG;
    CDLM_HIP(p->ctx, hipStreamSynchronize(p->ctx->stream));
    double t = 0.0;
    int64_t cnt = 0;
    for (auto &e : p->ev[kernel]) {
        float ms = 0.f;
        CDLM_HIP(p->ctx, hipEventElapsedTime(&ms, e.a, e.b));
        t += ms;
        cnt += e.count;
    }
    if (total_ms) *total_ms = t;
    if (launches) *launches = cnt;
    return CDLM_OK;
}
