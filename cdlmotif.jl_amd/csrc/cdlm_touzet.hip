// cdlm_touzet.hip -- score thresholds of PWMs for a p-value on the device
// (get_thresh / pvalue2score of CDLmotif.jl).
//
// Reference: src/SEARCH/helpers.jl:99 (get_thresh), src/SEARCH/PWM_Touzet.jl:9-44
// (best/worst score, min_score_range, round_pwm), :68-102 (score_distribution), :147-159
// (find_largest_alpha), :199-217 (pvalue2score, granularity 0.1 and the uniform background,
// src/SETUP/TYPE_setup.jl:52-54).
//
// The reference builds the score distribution of the rounded PWM as a sorted dictionary of
// Double64 keys and probabilities.  With the uniform background it is exact integer work:
// every rounded entry is q * c (q = floor(x / c), c the Float64 granularity), every key is
// K * c for an integer K, every probability is a count times 4^-m; in 106-bit arithmetic all
// of these are exact for m <= 53, and the one quotient it rounds, x / c, is never close
// enough to an integer for floor to change.  The distribution is therefore the m-fold
// convolution of the columns' 4-point integer histograms, which this file computes with
// 128-bit counts, one work-group per PWM; the column order (min_score_range) does not
// change an exact convolution, and the pruning bounds of score_distribution (alpha = the
// worst score, beta = Inf) never remove a score.
//
// find_largest_alpha walks the keys in ascending order with q = P(score >= key) and returns
// the first key with q < pval, or the largest key when there is none; the threshold is
// Float32(K * c) (the Float64 head of the Double64 key, rounded to Float32).

#include <algorithm>
#include <cmath>
#include <vector>

#include "cdlm_internal.h"

namespace {

typedef unsigned __int128 u128;
constexpr int kMaxTzLen = 53;        // exact agreement with 106-bit arithmetic
constexpr int kMaxTzMaxlen = 64;     // the PWM packing of the scans
constexpr int kTzThreads = 256;
constexpr double kMaxTzEntry = 1.0e5;          // |x / c| stays far inside int range
constexpr long long kMaxTzBins = 1LL << 26;    // all PWMs together (2 GiB of counts)

// floor(x / c) exactly: the largest integer q with q * c <= x (fma gives the sign of q*c - x
// exactly, as it rounds the exact value once).
__device__ __forceinline__ int floor_div(double x, double c) {
    double q = floor(x / c);
    if (fma(q, c, -x) > 0.0) q -= 1.0;
    else if (fma(q + 1.0, c, -x) <= 0.0) q += 1.0;
    return (int)q;
}

// Per PWM: K of the lowest score (sum of the column minima) and the number of score bins.
__global__ void k_tz_range(const float *__restrict__ pwms, const int *__restrict__ lens, int K, int maxlen, double c,
                           int *__restrict__ kmin, int *__restrict__ nbins, int *__restrict__ err) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const int m = lens[k];
    long lo = 0, span = 0;
    bool bad = false;
    for (int j = 0; j < m; ++j) {
        int qmin = 0, qmax = 0;
        for (int a = 0; a < 4; ++a) {
            const double x = pwms[((long)k * 4 + a) * maxlen + j];
            if (!(fabs(x) <= kMaxTzEntry)) {   // also NaN
                bad = true;
                continue;
            }
            const int q = floor_div(x, c);
            qmin = (a == 0 || q < qmin) ? q : qmin;
            qmax = (a == 0 || q > qmax) ? q : qmax;
        }
        lo += qmin;
        span += qmax - qmin;
    }
    if (bad) atomicOr(err, 1);
    kmin[k] = bad ? 0 : (int)lo;
    nbins[k] = bad ? 1 : (int)min(span + 1, kMaxTzBins + 1);   // the host refuses totals above kMaxTzBins
}

// tail < p * 4^m  <=>  tail < ceil(p * 4^m) for integer tail; p in [0, 1], so the bound is
// at most 4^m <= 2^106.
__device__ u128 tail_bound(double p, int m) {
    if (!(p > 0.0)) return 0;
    int e;
    const double fr = frexp(p, &e);                 // p = fr * 2^e, fr in [0.5, 1)
    const unsigned long long mant = (unsigned long long)ldexp(fr, 53);
    const int E = e - 53 + 2 * m;                    // p * 4^m = mant * 2^E
    if (E >= 0) return (u128)mant << E;
    const int sh = -E;
    if (sh > 53) return 1;                           // 0 < p * 4^m < 1
    return (u128)((mant + (1ULL << sh) - 1) >> sh);
}

// One work-group per PWM: the histogram of K - kmin over all 4^m sequences (exact counts),
// double-buffered in its slice of the workspace, then the threshold search.
__global__ __launch_bounds__(kTzThreads) void k_tz_dp(const float *__restrict__ pwms, const int *__restrict__ lens,
                                                      int maxlen, double c, double pval, const int *__restrict__ kmin,
                                                      const int *__restrict__ nbins, const long long *__restrict__ off,
                                                      u128 *__restrict__ work, float *__restrict__ thresh) {
    __shared__ int dq[4][kMaxTzLen];
    __shared__ int crange[kMaxTzLen];
    __shared__ u128 part[kTzThreads];
    __shared__ int s_cand, s_last;
    const int k = blockIdx.x, tid = threadIdx.x;
    const int m = lens[k], nb = nbins[k];
    u128 *A = work + 2 * off[k];
    u128 *B = A + nb;
    if (tid < m) {
        int q[4], qmin = 0, qmax = 0;
        for (int a = 0; a < 4; ++a) {
            q[a] = floor_div((double)pwms[((long)k * 4 + a) * maxlen + tid], c);
            qmin = (a == 0 || q[a] < qmin) ? q[a] : qmin;
            qmax = (a == 0 || q[a] > qmax) ? q[a] : qmax;
        }
        for (int a = 0; a < 4; ++a) dq[a][tid] = q[a] - qmin;
        crange[tid] = qmax - qmin;
    }
    if (tid == 0) {
        s_cand = 0x7fffffff;
        s_last = -1;
    }
    for (int s = tid; s < nb; s += kTzThreads) A[s] = (s == 0) ? (u128)1 : (u128)0;
    __syncthreads();
    int cur = 1;
    for (int j = 0; j < m; ++j) {
        const int nl = cur + crange[j];
        const int d0 = dq[0][j], d1 = dq[1][j], d2 = dq[2][j], d3 = dq[3][j];
        for (int s = tid; s < nl; s += kTzThreads) {
            u128 acc = 0;
            int t = s - d0;
            if (t >= 0 && t < cur) acc += A[t];
            t = s - d1;
            if (t >= 0 && t < cur) acc += A[t];
            t = s - d2;
            if (t >= 0 && t < cur) acc += A[t];
            t = s - d3;
            if (t >= 0 && t < cur) acc += A[t];
            B[s] = acc;
        }
        __syncthreads();
        u128 *sw = A;
        A = B;
        B = sw;
        cur = nl;
    }
    // cur == nb.  Bins with a zero count are unreachable scores (no key of the reference).
    const int chunk = (nb + kTzThreads - 1) / kTzThreads;
    const int lo = min(nb, tid * chunk), hi = min(nb, lo + chunk);
    u128 sum = 0;
    for (int b = lo; b < hi; ++b) sum += A[b];
    part[tid] = sum;
    __syncthreads();
    u128 running = sum;                               // P(score >= bin lo) * 4^m
    for (int t = tid + 1; t < kTzThreads; ++t) running += part[t];
    const u128 bound = tail_bound(pval, m);
    int cand = -1, last = -1;
    for (int b = lo; b < hi; ++b) {
        const u128 v = A[b];
        if (v != 0) {
            if (running < bound) {
                cand = b;
                break;
            }
            last = b;
        }
        running -= v;
    }
    if (cand >= 0) atomicMin(&s_cand, cand);
    else if (last >= 0) atomicMax(&s_last, last);
    __syncthreads();
    if (tid == 0) {
        const int bin = s_cand != 0x7fffffff ? s_cand : s_last;
        thresh[k] = (float)((double)(kmin[k] + bin) * c);
    }
}

struct TzBuffers {
    float *pwms = nullptr, *thresh = nullptr;
    int *lens = nullptr, *kmin = nullptr, *nb = nullptr, *err = nullptr;
    long long *off = nullptr;
    u128 *work = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    ~TzBuffers() {
        for (void *q : {(void *)pwms, (void *)thresh, (void *)lens, (void *)kmin, (void *)nb, (void *)err, (void *)off,
                        (void *)work})
            if (q) (void)hipFree(q);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

}  // namespace

extern "C" cdlm_status cdlm_pwm_pvalue_thresh(cdlm_ctx *ctx, const float *pwms, const int32_t *lens, int32_t K,
                                              int32_t maxlen, double pval, double eps, float *thresh) {
    if (!ctx) return CDLM_ERR_ARG;
    if (!pwms || !lens || !thresh) return cdlm_fail(ctx, CDLM_ERR_ARG, "null argument");
    if (K < 0) return cdlm_fail(ctx, CDLM_ERR_ARG, "K must be >= 0 (got %d)", K);
    if (maxlen <= 0 || maxlen > kMaxTzMaxlen)
        return cdlm_fail(ctx, CDLM_ERR_UNSUPPORTED, "maxlen %d outside 1..%d", maxlen, kMaxTzMaxlen);
    if (!(pval >= 0.0 && pval <= 1.0)) return cdlm_fail(ctx, CDLM_ERR_ARG, "pvalue must be in [0,1] (got %g)", pval);
    if (!(eps > 0.0) || !std::isfinite(eps)) return cdlm_fail(ctx, CDLM_ERR_ARG, "granularity must be positive");
    if (eps < 1e-3) return cdlm_fail(ctx, CDLM_ERR_UNSUPPORTED, "granularity %g below 1e-3", eps);
    for (int k = 0; k < K; ++k) {
        if (lens[k] <= 0 || lens[k] > maxlen) return cdlm_fail(ctx, CDLM_ERR_ARG, "lens[%d]=%d outside 1..maxlen", k, lens[k]);
        if (lens[k] > kMaxTzLen)
            return cdlm_fail(ctx, CDLM_ERR_UNSUPPORTED, "lens[%d]=%d: PWMs longer than %d are not supported", k, lens[k],
                             kMaxTzLen);
    }
    if (K == 0) return CDLM_OK;
    hipStream_t st = ctx->stream;
    TzBuffers d;
    const size_t pw = sizeof(float) * 4 * (size_t)maxlen * K;
    CDLM_HIP(ctx, hipMalloc(&d.pwms, pw));
    CDLM_HIP(ctx, hipMalloc(&d.lens, sizeof(int) * K));
    CDLM_HIP(ctx, hipMalloc(&d.kmin, sizeof(int) * K));
    CDLM_HIP(ctx, hipMalloc(&d.nb, sizeof(int) * K));
    CDLM_HIP(ctx, hipMalloc(&d.err, sizeof(int)));
    CDLM_HIP(ctx, hipMalloc(&d.off, sizeof(long long) * K));
    CDLM_HIP(ctx, hipMalloc(&d.thresh, sizeof(float) * K));
    CDLM_HIP(ctx, hipMemcpyAsync(d.pwms, pwms, pw, hipMemcpyHostToDevice, st));
    CDLM_HIP(ctx, hipMemcpyAsync(d.lens, lens, sizeof(int) * K, hipMemcpyHostToDevice, st));
    CDLM_HIP(ctx, hipMemsetAsync(d.err, 0, sizeof(int), st));
    bool timed = true;
    for (hipEvent_t &e : d.ev)
        if (hipEventCreate(&e) != hipSuccess) {
            (void)hipGetLastError();
            e = nullptr;
            timed = false;
        }
    hipEvent_t e0 = timed ? d.ev[0] : nullptr, e1 = d.ev[1], e2 = d.ev[2], e3 = d.ev[3];
    if (e0) (void)hipEventRecord(e0, st);
    hipLaunchKernelGGL(k_tz_range, dim3((K + 63) / 64), dim3(64), 0, st, d.pwms, d.lens, K, maxlen, eps, d.kmin, d.nb,
                       d.err);
    CDLM_HIP(ctx, hipGetLastError());
    if (e0) (void)hipEventRecord(e1, st);
    std::vector<int> nb(K);
    int err = 0;
    CDLM_HIP(ctx, hipMemcpyAsync(nb.data(), d.nb, sizeof(int) * K, hipMemcpyDeviceToHost, st));
    CDLM_HIP(ctx, hipMemcpyAsync(&err, d.err, sizeof(int), hipMemcpyDeviceToHost, st));
    CDLM_HIP(ctx, hipStreamSynchronize(st));
    cdlm_status s = CDLM_OK;
    if (err) s = cdlm_fail(ctx, CDLM_ERR_ARG, "PWM entries must be finite with magnitude <= %g", kMaxTzEntry);
    std::vector<long long> off(K);
    long long tot = 0;
    for (int k = 0; k < K; ++k) {
        off[k] = tot;
        tot += nb[k];
    }
    if (s == CDLM_OK && tot > kMaxTzBins)
        s = cdlm_fail(ctx, CDLM_ERR_UNSUPPORTED, "%lld score bins over all PWMs (at most %lld): split the PWM set", tot,
                      kMaxTzBins);
    if (s == CDLM_OK) {
        hipError_t e = hipMalloc(&d.work, sizeof(u128) * 2 * (size_t)tot);
        if (e == hipSuccess) e = hipMemcpyAsync(d.off, off.data(), sizeof(long long) * K, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) {
            if (e0) (void)hipEventRecord(e2, st);
            hipLaunchKernelGGL(k_tz_dp, dim3(K), dim3(kTzThreads), 0, st, d.pwms, d.lens, maxlen, eps, pval, d.kmin, d.nb,
                               d.off, d.work, d.thresh);
            e = hipGetLastError();
            if (e0) (void)hipEventRecord(e3, st);
        }
        if (e == hipSuccess) e = hipMemcpyAsync(thresh, d.thresh, sizeof(float) * K, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            s = cdlm_fail(ctx, e == hipErrorOutOfMemory ? CDLM_ERR_OOM : CDLM_ERR_HIP, "p-value thresholds: %s",
                          hipGetErrorString(e));
        }
    }
    float ms_a = 0.f, ms_b = 0.f;
    if (s == CDLM_OK && e0 && hipEventElapsedTime(&ms_a, e0, e1) == hipSuccess &&
        hipEventElapsedTime(&ms_b, e2, e3) == hipSuccess)
        ctx->last_pvalue_ms = (double)ms_a + ms_b;
    (void)hipStreamSynchronize(st);
    return s;
}

extern "C" cdlm_status cdlm_pwm_pvalue_last_ms(cdlm_ctx *ctx, double *ms) {
    if (!ctx || !ms) return CDLM_ERR_ARG;
    *ms = ctx->last_pvalue_ms;
    return CDLM_OK;
}
