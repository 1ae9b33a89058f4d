// cdlm_internal.h -- shared host/device helpers of libcdlm (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/cdlm.h"

struct cdlm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;
    int num_cu = 0;
    double last_scan_ms = 0.0;   // device time of the last PWM hit-scan kernel
    double last_topology_ms = 0.0;   // device time of the last activation-topology kernel
    double last_pvalue_ms = 0.0;     // device time of the last p-value threshold kernels
    void *pwm_scratch = nullptr;     // this device's address of the fixed PWM-scan scratch array
};

// Resolves the device address of the PWM scan's fixed scratch array for ctx's device (current)
// into ctx->pwm_scratch; once per device, from cdlm_create.
cdlm_status cdlm_pwm_scratch_init(cdlm_ctx *ctx);

// Records an error message on the context and returns the status.
cdlm_status cdlm_fail(cdlm_ctx *ctx, cdlm_status st, const char *fmt, ...);

// cdlm_problem_create from the one-hot matrix, with every device buffer carved from a caller-owned
// arena of at least cdlm_footprint_bytes(L4, N, M, gamma, 0) bytes, which the problem does not
// free (the server recycles finished jobs' memory this way instead of freeing it).
cdlm_status cdlm_problem_create_in(cdlm_ctx *ctx, const cdlm_params *prm, const float *s, int32_t L4, int32_t N,
                                   const float *D0, uint64_t seed, void *arena, size_t arena_bytes, cdlm_problem **out);

// Device bytes of the smallest window a streamed problem of this shape can use (its sequences
// in as many chunks as the dictionary means have); 0 for a bad shape.
size_t cdlm_stream_min_bytes(int32_t L4, int32_t N, int32_t M, int32_t gamma);

// Device pointer of a problem's current X_best slice (L, M, N) and its dimensions; reads the
// slice index from the device (synchronises the context stream).
cdlm_status cdlm_problem_xbest_device(cdlm_problem *p, const float **slice, int *L, int *M, int *N);

#define CDLM_HIP(ctx, expr)                                                                        \
    do {                                                                                           \
        hipError_t _e = (expr);                                                                    \
        if (_e != hipSuccess) {                                                                    \
            if (_e == hipErrorOutOfMemory || _e == hipErrorMemoryAllocation)                       \
                return cdlm_fail((ctx), CDLM_ERR_OOM, "%s: %s", #expr, hipGetErrorString(_e));     \
            return cdlm_fail((ctx), CDLM_ERR_HIP, "%s: %s", #expr, hipGetErrorString(_e));         \
        }                                                                                          \
    } while (0)

// ------------------------------------------------------------------ FFT plan
// Stockham autosort, mixed radix.  Stage st uses radix[st]; the twiddle table holds
// W_N^t = exp(-2 pi i t / N) for the table length N; a transform of length n = N / twstride
// reads entry t * twstride for W_n^t.
constexpr int kMaxStages = 24;
struct FftPlan {
    int n = 0;
    int nst = 0;
    int twstride = 1;
    int radix[kMaxStages] = {0};
};

// Stage orders: radix 8, 4 and 2 stages first, then odd primes ascending; or odd primes
// first, descending (the largest prime then runs without input twiddles).
enum FftOrder { kFftSmallFirst = 0, kFftPrimesFirst = 1, kFftOrders = 2 };

inline bool make_fft_plan(int n, int twstride, FftPlan *p, int order = kFftSmallFirst) {
    p->n = n;
    p->twstride = twstride;
    p->nst = 0;
    int m = n;
    int pw[kMaxStages], npw = 0, od[kMaxStages], nod = 0;
    // powers of two: radix 8 stages, then a 4 or a 2; 2^(3q+1) uses two 4s instead of 8 * 2
    int e = 0;
    while (m % 2 == 0) { ++e; m /= 2; }
    int q8 = e / 3, n4 = 0, n2 = 0;
    if (e % 3 == 2) n4 = 1;
    else if (e % 3 == 1) { if (q8 > 0) { --q8; n4 = 2; } else n2 = 1; }
    if (q8 + n4 + n2 > kMaxStages) return false;
    for (int i = 0; i < q8; ++i) pw[npw++] = 8;
    for (int i = 0; i < n4; ++i) pw[npw++] = 4;
    for (int i = 0; i < n2; ++i) pw[npw++] = 2;
    for (int q = 3; m > 1; q += 2)
        while (m % q == 0) { if (nod >= kMaxStages) return false; od[nod++] = q; m /= q; }
    if (npw + nod > kMaxStages) return false;
    if (order == kFftPrimesFirst) {
        for (int i = nod - 1; i >= 0; --i) p->radix[p->nst++] = od[i];
        for (int i = 0; i < npw; ++i) p->radix[p->nst++] = pw[i];
    } else {
        for (int i = 0; i < npw; ++i) p->radix[p->nst++] = pw[i];
        for (int i = 0; i < nod; ++i) p->radix[p->nst++] = od[i];
    }
    return true;
}

// ------------------------------------------------------------------ device side
#if defined(__HIPCC__)

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

template <bool INV>
__device__ __forceinline__ float2 twid(const float2 *tw, int idx) {
    float2 w = tw[idx];
    if (INV) w.y = -w.y;
    return w;
}

// Makes this wave's LDS writes visible to its own lanes.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// q = x / d and r = x - q*d for 0 <= x < 2^24, with rcp = 1.0f / d: the float quotient is
// within one of the true one and is corrected.
__device__ __forceinline__ int fdivmod(int x, int d, float rcp, int &r) {
    int q = (int)((float)x * rcp);
    r = x - q * d;
    if (r < 0) { --q; r += d; }
    else if (r >= d) { ++q; r -= d; }
    return q;
}

// One radix-R butterfly j (position k = j mod Ns) of a Stockham stage with sub-length Ns:
// inputs a[j + r T] times W_{Ns R}^{r k} (twiddle table index r * kst, kst = k * table step;
// TW = false in the first stage, where k = 0), outputs b[(j - k) R + k + q Ns].
// rs is the table step of the radix-R roots W_R (odd R).
template <int R, bool INV, bool TW>
__device__ __forceinline__ void fft_bfly(const float2 *__restrict__ a, float2 *__restrict__ b, int j, int k, int kst, int T,
                                         int Ns, int rs, const float2 *__restrict__ tw) {
    const int ob = (j - k) * R + k;
    float2 v[R];
    v[0] = a[j];
#pragma unroll
    for (int r = 1; r < R; ++r) {
        const float2 x = a[j + r * T];
        v[r] = TW ? cmul(x, twid<INV>(tw, r * kst)) : x;
    }
    if constexpr (R == 2) {
        b[ob] = cadd(v[0], v[1]);
        b[ob + Ns] = csub(v[0], v[1]);
    } else if constexpr (R == 4) {
        float2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        float2 s13 = cadd(v[1], v[3]), d13 = csub(v[1], v[3]);
        // forward: W_4 = -i ; d13 * (-i) = (d13.y, -d13.x)
        float2 rot = INV ? make_float2(-d13.y, d13.x) : make_float2(d13.y, -d13.x);
        b[ob] = cadd(s02, s13);
        b[ob + Ns] = cadd(d02, rot);
        b[ob + 2 * Ns] = csub(s02, s13);
        b[ob + 3 * Ns] = csub(d02, rot);
    } else if constexpr (R == 8) {
        // decimation in frequency: y_{2m} = DFT4(v_n + v_{n+4})_m,
        // y_{2m+1} = DFT4((v_n - v_{n+4}) W_8^n)_m
        constexpr float c = 0.70710678118654752f;
        float2 e[4], o[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            e[q] = cadd(v[q], v[q + 4]);
            o[q] = csub(v[q], v[q + 4]);
        }
        if (INV) {
            o[1] = make_float2(c * (o[1].x - o[1].y), c * (o[1].x + o[1].y));
            o[2] = make_float2(-o[2].y, o[2].x);
            o[3] = make_float2(-c * (o[3].x + o[3].y), c * (o[3].x - o[3].y));
        } else {
            o[1] = make_float2(c * (o[1].x + o[1].y), c * (o[1].y - o[1].x));
            o[2] = make_float2(o[2].y, -o[2].x);
            o[3] = make_float2(c * (o[3].y - o[3].x), -c * (o[3].x + o[3].y));
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const float2 *u = h ? o : e;
            const float2 s02 = cadd(u[0], u[2]), d02 = csub(u[0], u[2]);
            const float2 s13 = cadd(u[1], u[3]), d13 = csub(u[1], u[3]);
            const float2 rot = INV ? make_float2(-d13.y, d13.x) : make_float2(d13.y, -d13.x);
            b[ob + h * Ns] = cadd(s02, s13);
            b[ob + (2 + h) * Ns] = cadd(d02, rot);
            b[ob + (4 + h) * Ns] = csub(s02, s13);
            b[ob + (6 + h) * Ns] = csub(d02, rot);
        }
    } else {
        // odd R: outputs q and R - q from a_r = v_r + v_{R-r} and b_r = v_r - v_{R-r}:
        //   y_q = v_0 + sum_r a_r cos(2 pi r q / R) -/+ i sum_r b_r sin(2 pi r q / R)
        constexpr int H = (R - 1) / 2;
        float cs[H], sn[H];
#pragma unroll
        for (int m = 1; m <= H; ++m) {
            const float2 w = tw[m * rs];   // (cos, -sin)(2 pi m / R)
            cs[m - 1] = w.x;
            sn[m - 1] = -w.y;
        }
        float2 ap[H], bm[H];
        float2 y0 = v[0];
#pragma unroll
        for (int r = 1; r <= H; ++r) {
            ap[r - 1] = cadd(v[r], v[R - r]);
            bm[r - 1] = csub(v[r], v[R - r]);
            y0 = cadd(y0, ap[r - 1]);
        }
        b[ob] = y0;
#pragma unroll
        for (int q = 1; q <= H; ++q) {
            float2 re = v[0], im = make_float2(0.f, 0.f);
#pragma unroll
            for (int r = 1; r <= H; ++r) {
                const int m = (r * q) % R;
                const int mm = m <= H ? m : R - m;
                const float sg = m <= H ? 1.f : -1.f;
                re.x = fmaf(cs[mm - 1], ap[r - 1].x, re.x);
                re.y = fmaf(cs[mm - 1], ap[r - 1].y, re.y);
                im.x = fmaf(sg * sn[mm - 1], bm[r - 1].x, im.x);
                im.y = fmaf(sg * sn[mm - 1], bm[r - 1].y, im.y);
            }
            // forward: y_q = re - i im, y_{R-q} = re + i im (inverse: the reverse)
            const float2 mi = make_float2(re.x + im.y, re.y - im.x);
            const float2 pi = make_float2(re.x - im.y, re.y + im.x);
            b[ob + q * Ns] = INV ? pi : mi;
            b[ob + (R - q) * Ns] = INV ? mi : pi;
        }
    }
}

// One Stockham stage of radix R over nb transforms of length n (T = n / R butterflies each):
// thread tid takes butterflies tid, tid + nthr, ...; its (transform, j, k = j mod Ns, k * st)
// advance by additions (nthr = dB T + dT and T is a multiple of Ns), so the loop has no
// division or integer multiply.
template <int R, bool INV>
__device__ __forceinline__ void fft_stage(const float2 *__restrict__ a, float2 *__restrict__ b, int n, int nb, int T,
                                          int Ns, int st, int rs, const float2 *__restrict__ tw, int tid, int nthr) {
    const int tot = nb * T;
    if (tid >= tot) return;
    const float rT = 1.0f / (float)T, rNs = 1.0f / (float)Ns;
    int j, k, dT, dK;
    const int bb = fdivmod(tid, T, rT, j);
    (void)fdivmod(j, Ns, rNs, k);
    const int dB = fdivmod(nthr, T, rT, dT);
    (void)fdivmod(dT, Ns, rNs, dK);
    int kst = k * st;
    const int dKst = dK * st, Nst = Ns * st, dOff = dB * n;
    int off = bb * n;
    if (Ns == 1) {
        for (int jj = tid; jj < tot; jj += nthr) {
            fft_bfly<R, INV, false>(a + off, b + off, j, 0, 0, T, 1, rs, tw);
            j += dT;
            off += dOff;
            if (j >= T) { j -= T; off += n; }
        }
    } else {
        for (int jj = tid; jj < tot; jj += nthr) {
            fft_bfly<R, INV, true>(a + off, b + off, j, k, kst, T, Ns, rs, tw);
            j += dT;
            off += dOff;
            k += dK;
            kst += dKst;
            if (j >= T) { j -= T; off += n; }
            if (k >= Ns) { k -= Ns; kst -= Nst; }
        }
    }
}

// Large odd radix R: one work item per (butterfly j, output pair u); u = 0 writes y_0,
// u > 0 writes y_u and y_{R-u} with the same pairing as the small odd butterflies.
template <bool INV>
__device__ __forceinline__ void fft_pair_item(const float2 *__restrict__ a, float2 *__restrict__ b, int j, int k, int u,
                                              int R, int T, int Ns, int st, int rs,
                                              const float2 *__restrict__ tw) {
    const int ob = (j - k) * R + k;
    const float2 x0 = a[j];
    const bool twd = Ns > 1;
    if (u == 0) {
        float2 y = x0;
        for (int r = 1; r < R; ++r) {
            float2 x = a[j + r * T];
            if (twd) x = cmul(x, twid<INV>(tw, r * k * st));
            y = cadd(y, x);
        }
        b[ob] = y;
        return;
    }
    float2 re = x0, im = make_float2(0.f, 0.f);
    int e = 0;
    for (int r = 1; 2 * r < R; ++r) {
        float2 xr = a[j + r * T], xm = a[j + (R - r) * T];
        if (twd) {
            xr = cmul(xr, twid<INV>(tw, r * k * st));
            xm = cmul(xm, twid<INV>(tw, (R - r) * k * st));
        }
        e += u;
        if (e >= R) e -= R;
        const float2 w = tw[e * rs];   // (cos, -sin)(2 pi r u / R)
        const float2 ap = cadd(xr, xm), bm = csub(xr, xm);
        re.x = fmaf(w.x, ap.x, re.x);
        re.y = fmaf(w.x, ap.y, re.y);
        im.x = fmaf(-w.y, bm.x, im.x);
        im.y = fmaf(-w.y, bm.y, im.y);
    }
    const float2 mi = make_float2(re.x + im.y, re.y - im.x);
    const float2 pi = make_float2(re.x - im.y, re.y + im.x);
    b[ob + u * Ns] = INV ? pi : mi;
    b[ob + (R - u) * Ns] = INV ? mi : pi;
}

// Team-cooperative complex FFTs of length P.n on nb consecutive transforms (transform t at
// offset t*n) from a into b (ping-pong).  The team is threads [0, nthr) that all call this;
// SYNC is the team barrier.  Returns the buffer holding the (unscaled) results.
// Forward: exp(-2 pi i kl/n); INV: exp(+2 pi i kl/n).  tw is the twiddle table (LDS).
template <bool INV, typename Sync>
__device__ __forceinline__ float2 *team_fft(float2 *a, float2 *b, const FftPlan &P, const float2 *__restrict__ tw, int tid, int nthr,
                            Sync sync, int nb = 1) {
    const int n = P.n;
    const int tws = P.twstride;
    int Ns = 1;
    for (int st = 0; st < P.nst; ++st) {
        const int R = P.radix[st];
        const int T = n / R;
        const float rT = 1.0f / (float)T, rNs = 1.0f / (float)Ns;
        const int stw = (n / (Ns * R)) * tws;   // twiddle step of r*k
        const int rs = (n / R) * tws;           // twiddle step of the radix-R roots
        switch (R) {
        case 2: fft_stage<2, INV>(a, b, n, nb, T, Ns, stw, rs, tw, tid, nthr); break;
        case 3: fft_stage<3, INV>(a, b, n, nb, T, Ns, stw, rs, tw, tid, nthr); break;
        case 4: fft_stage<4, INV>(a, b, n, nb, T, Ns, stw, rs, tw, tid, nthr); break;
        case 5: fft_stage<5, INV>(a, b, n, nb, T, Ns, stw, rs, tw, tid, nthr); break;
        case 7: fft_stage<7, INV>(a, b, n, nb, T, Ns, stw, rs, tw, tid, nthr); break;
        case 8: fft_stage<8, INV>(a, b, n, nb, T, Ns, stw, rs, tw, tid, nthr); break;
        default: {
            // large odd radix: (butterfly, output pair) work items
            const int H = (R + 1) / 2;
            const int per = H * T;
            const float rper = 1.0f / (float)per;
            for (int idx = tid; idx < nb * per; idx += nthr) {
                int rem, j, k;
                const int bb = fdivmod(idx, per, rper, rem);
                const int u = fdivmod(rem, T, rT, j);
                (void)fdivmod(j, Ns, rNs, k);
                fft_pair_item<INV>(a + bb * n, b + bb * n, j, k, u, R, T, Ns, stw, rs, tw);
            }
        } break;
        }
        sync();
        float2 *t = a; a = b; b = t;
        Ns *= R;
    }
    return a;
}

struct BlockSync {
    __device__ __forceinline__ void operator()() const { __syncthreads(); }
};
struct WaveSync {
    __device__ __forceinline__ void operator()() const { wave_lds_sync(); }
};

#endif  // __HIPCC__
