// cdlm_pwm.hip -- PWM scans of CDLmotif.jl for gfx950.
//
// Reference: src/SEARCH/scan.jl:10-21 (scan_n), :89-121 (overlapping_scan!),
// :176-239 (scan_w_gpu! / greedy_search!), src/SEARCH/helpers.jl:16 (pwm_complement).
//
// Scores are FP32 sums of PWM entries selected by the bases (the reference multiplies by
// one-hot 0/1 entries; a product with 0 adds a signed zero, which never changes a sum that
// starts at +0).  Bit-exactness needs the reference's summation order:
//   position-major (greedy_search!, scan.jl:231-234): acc += pwm[b(l+j)][j], j = 0..m-1
//   nucleotide-major (scan_n, scan.jl:15-16): for a in A,C,G,T: for j: acc += pwm[a][j]*[b==a]
// The reverse-complement PWM is pwm_c[a][j] = pwm[3-a][m-1-j].
//
// Sequences: base codes 0..3, one byte each, (N, L) row-major.

#include <cstring>

#include "cdlm_internal.h"

namespace {

constexpr int kPwmThreads = 256;
constexpr int kMaxPwmLen = 64;

// ---------------------------------------------------------------------------- hit list
// Block (n-tile of one sequence, k): one thread per window.  Records (k, n, pos, strand).
__global__ __launch_bounds__(kPwmThreads) void k_pwm_hits(const uint8_t *__restrict__ bases, int L, long N,
                                                          const float *__restrict__ pwms, const int *__restrict__ lens,
                                                          const float *__restrict__ thresh, int K, int maxlen,
                                                          int both, int *hk, long long *hn, int *hp, signed char *hs,
                                                          long long cap, unsigned long long *count) {
    __shared__ float pw[4 * kMaxPwmLen];
    const long n = blockIdx.x;
    const int k = blockIdx.y;
    const int m = lens[k];
    for (int i = threadIdx.x; i < 4 * maxlen; i += kPwmThreads) pw[i] = pwms[(size_t)k * 4 * maxlen + i];
    __syncthreads();
    const uint8_t *seq = bases + n * (long)L;
    const float th = thresh[k];
    for (int l0 = threadIdx.x; l0 + m <= L; l0 += kPwmThreads) {
        float f = 0.f;
        for (int j = 0; j < m; ++j) {
            const int b = seq[l0 + j];
            f = f + (b < 4 ? pw[b * maxlen + j] : 0.f);
        }
        if (f > th) {
            unsigned long long slot = atomicAdd(count, 1ULL);
            if ((long long)slot < cap) { hk[slot] = k; hn[slot] = n; hp[slot] = l0; hs[slot] = 0; }
        }
        if (both) {
            float r = 0.f;
            for (int j = 0; j < m; ++j) {
                const int b = seq[l0 + j];
                r = r + (b < 4 ? pw[(3 - b) * maxlen + (m - 1 - j)] : 0.f);
            }
            if (r > th) {
                unsigned long long slot = atomicAdd(count, 1ULL);
                if ((long long)slot < cap) { hk[slot] = k; hn[slot] = n; hp[slot] = l0; hs[slot] = 1; }
            }
        }
    }
}

// ---------------------------------------------------------------------------- GPU argmax
// One wave per (k, n): v(l) = (score > thresh ? score : 0) for l <= L-m, 0 past it;
// first argmax over l; index 0 -> 0 ("no hit", scan.jl:197), else l + 1.
__global__ __launch_bounds__(kPwmThreads) void k_pwm_argmax(const uint8_t *__restrict__ bases, int L, long N,
                                                            const float *__restrict__ pwms, const int *__restrict__ lens,
                                                            const float *__restrict__ thresh, int K, int maxlen,
                                                            int *out) {
    const int lane = threadIdx.x & 63;
    const long wid = (long)blockIdx.x * (kPwmThreads / 64) + (threadIdx.x >> 6);
    if (wid >= (long)K * N) return;
    const int k = (int)(wid / N);
    const long n = wid % N;
    const int m = lens[k];
    const float *pw = pwms + (size_t)k * 4 * maxlen;
    const uint8_t *seq = bases + n * (long)L;
    const float th = thresh[k];
    float best = 0.f;
    int bi = 0x7fffffff;
    for (int l0 = lane; l0 < L; l0 += 64) {
        float v = 0.f;
        if (l0 + m <= L) {
            float f = 0.f;
            for (int j = 0; j < m; ++j) {
                const int b = seq[l0 + j];
                f = f + (b < 4 ? pw[b * maxlen + j] : 0.f);
            }
            v = f > th ? f : 0.f;
        }
        if (bi == 0x7fffffff || v > best) { best = v; bi = l0; }
    }
    // wave reduction: larger value wins, ties -> smaller index
    for (int off = 32; off > 0; off >>= 1) {
        const float ob = __shfl_xor(best, off);
        const int oi = __shfl_xor(bi, off);
        if (oi != 0x7fffffff && (bi == 0x7fffffff || ob > best || (ob == best && oi < bi))) { best = ob; bi = oi; }
    }
    if (lane == 0) out[(size_t)k * N + n] = (bi == 0 || bi == 0x7fffffff) ? 0 : bi + 1;
}

// ---------------------------------------------------------------------------- best hit, both strands
// overlapping_scan! with scan_n's nucleotide-major order; one wave per (k, n).
__device__ __forceinline__ float nm_score(const uint8_t *seq, int l0, const float *pw, int maxlen, int m, bool comp) {
    float acc = 0.f;
    for (int a = 0; a < 4; ++a)
        for (int j = 0; j < m; ++j) {
            if (seq[l0 + j] != a) continue;
            acc += comp ? pw[(3 - a) * maxlen + (m - 1 - j)] : pw[a * maxlen + j];
        }
    return acc;
}

__global__ __launch_bounds__(kPwmThreads) void k_pwm_best(const uint8_t *__restrict__ bases, int L, long N,
                                                          const float *__restrict__ pwms, const int *__restrict__ lens,
                                                          const float *__restrict__ thresh, int K, int maxlen,
                                                          int *opos, signed char *ocomp, float *oscore) {
    const int lane = threadIdx.x & 63;
    const long wid = (long)blockIdx.x * (kPwmThreads / 64) + (threadIdx.x >> 6);
    if (wid >= (long)K * N) return;
    const int k = (int)(wid / N);
    const long n = wid % N;
    const int m = lens[k];
    const float *pw = pwms + (size_t)k * 4 * maxlen;
    const uint8_t *seq = bases + n * (long)L;
    float bf = -INFINITY, bc = -INFINITY;
    int pf = 0x7fffffff, pc = 0x7fffffff;
    for (int l0 = lane; l0 + m <= L; l0 += 64) {
        const float f = nm_score(seq, l0, pw, maxlen, m, false);
        const float c = nm_score(seq, l0, pw, maxlen, m, true);
        if (pf == 0x7fffffff || f > bf) { bf = f; pf = l0; }
        if (pc == 0x7fffffff || c > bc) { bc = c; pc = l0; }
    }
    for (int off = 32; off > 0; off >>= 1) {
        float ob = __shfl_xor(bf, off);
        int oi = __shfl_xor(pf, off);
        if (oi != 0x7fffffff && (pf == 0x7fffffff || ob > bf || (ob == bf && oi < pf))) { bf = ob; pf = oi; }
        ob = __shfl_xor(bc, off);
        oi = __shfl_xor(pc, off);
        if (oi != 0x7fffffff && (pc == 0x7fffffff || ob > bc || (ob == bc && oi < pc))) { bc = ob; pc = oi; }
    }
    if (lane == 0) {
        const size_t o = (size_t)k * N + n;
        const float th = thresh[k];
        int pos = 0, comp = 0;
        float sc = 0.f;
        // scan_n keeps -Inf / -1 when no window fits; the tie rule then picks forward
        if (bf >= bc) {
            if (pf != 0x7fffffff && bf > th) { pos = pf + 1; comp = 0; sc = bf; }
        } else {
            if (pc != 0x7fffffff && bc > th) { pos = pc + 1; comp = 1; sc = bc; }
        }
        opos[o] = pos;
        ocomp[o] = (signed char)comp;
        oscore[o] = sc;
    }
}

// One-hot (4L, N) columns -> base codes (N, L); no 1 in a group -> 4.
__global__ void k_onehot_to_bases(const float *__restrict__ oh, int L4, long N, uint8_t *out) {
    const int L = L4 / 4;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)L * N) return;
    const long n = i / L;
    const int l = (int)(i % L);
    const float4 v = reinterpret_cast<const float4 *>(oh + n * (long)L4)[l];
    uint8_t b = 4;
    if (v.x == 1.f) b = 0;
    else if (v.y == 1.f) b = 1;
    else if (v.z == 1.f) b = 2;
    else if (v.w == 1.f) b = 3;
    out[i] = b;
}

struct DevPwm {
    uint8_t *bases = nullptr;
    float *pwms = nullptr, *thresh = nullptr;
    int *lens = nullptr;
};

cdlm_status check_pwm_args(cdlm_ctx *ctx, const void *bases, int L, long N, const float *pwms, const int *lens,
                           const float *thresh, int K, int maxlen) {
    if (!bases || !pwms || !lens || !thresh) return cdlm_fail(ctx, CDLM_ERR_ARG, "null argument");
    if (L <= 0 || N < 0 || K <= 0) return cdlm_fail(ctx, CDLM_ERR_ARG, "bad sizes L=%d N=%ld K=%d", L, N, K);
    if (maxlen <= 0 || maxlen > kMaxPwmLen) return cdlm_fail(ctx, CDLM_ERR_UNSUPPORTED, "maxlen %d outside 1..%d", maxlen, kMaxPwmLen);
    for (int k = 0; k < K; ++k)
        if (lens[k] <= 0 || lens[k] > maxlen) return cdlm_fail(ctx, CDLM_ERR_ARG, "lens[%d]=%d outside 1..maxlen", k, lens[k]);
    return CDLM_OK;
}

// Uploads the scan inputs; bases optional (device-resident callers pass their own).
cdlm_status upload(cdlm_ctx *ctx, DevPwm &d, const uint8_t *bases, int L, long N, const float *pwms, const int *lens,
                   const float *thresh, int K, int maxlen) {
    hipStream_t st = ctx->stream;
    if (bases) {
        CDLM_HIP(ctx, hipMalloc(&d.bases, (size_t)L * N + 16));
        CDLM_HIP(ctx, hipMemcpyAsync(d.bases, bases, (size_t)L * N, hipMemcpyHostToDevice, st));
    }
    CDLM_HIP(ctx, hipMalloc(&d.pwms, sizeof(float) * 4 * maxlen * (size_t)K));
    CDLM_HIP(ctx, hipMalloc(&d.thresh, sizeof(float) * K));
    CDLM_HIP(ctx, hipMalloc(&d.lens, sizeof(int) * K));
    CDLM_HIP(ctx, hipMemcpyAsync(d.pwms, pwms, sizeof(float) * 4 * maxlen * (size_t)K, hipMemcpyHostToDevice, st));
    CDLM_HIP(ctx, hipMemcpyAsync(d.thresh, thresh, sizeof(float) * K, hipMemcpyHostToDevice, st));
    CDLM_HIP(ctx, hipMemcpyAsync(d.lens, lens, sizeof(int) * K, hipMemcpyHostToDevice, st));
    return CDLM_OK;
}

void release(cdlm_ctx *ctx, DevPwm &d, bool own_bases) {
    (void)hipStreamSynchronize(ctx->stream);
    if (own_bases && d.bases) (void)hipFree(d.bases);
    if (d.pwms) (void)hipFree(d.pwms);
    if (d.thresh) (void)hipFree(d.thresh);
    if (d.lens) (void)hipFree(d.lens);
}

cdlm_status launch_hits(cdlm_ctx *ctx, const uint8_t *d_bases, int L, long N, const DevPwm &d, int K, int maxlen,
                        int both, int *hk, long long *hn, int *hp, signed char *hs, long long cap,
                        unsigned long long *d_count) {
    hipStream_t st = ctx->stream;
    CDLM_HIP(ctx, hipMemsetAsync(d_count, 0, sizeof(unsigned long long), st));
    if (N > 0) {
        // grid.x * block.x must stay below 2^32 work-items
        if ((unsigned long long)N * kPwmThreads >= (1ULL << 32))
            return cdlm_fail(ctx, CDLM_ERR_UNSUPPORTED, "N=%ld too large for one launch", N);
        hipLaunchKernelGGL(k_pwm_hits, dim3((unsigned)N, K), dim3(kPwmThreads), 0, st, d_bases, L, N, d.pwms, d.lens,
                           d.thresh, K, maxlen, both, hk, hn, hp, hs, cap, d_count);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return cdlm_fail(ctx, CDLM_ERR_HIP, "k_pwm_hits: %s", hipGetErrorString(e));
    return CDLM_OK;
}

}  // namespace

extern "C" cdlm_status cdlm_pwm_hits(cdlm_ctx *ctx, const uint8_t *bases, int32_t L, int64_t N, const float *pwms,
                                     const int32_t *lens, const float *thresh, int32_t K, int32_t maxlen,
                                     int32_t both_strands, int32_t *hit_k, int64_t *hit_n, int32_t *hit_pos,
                                     int8_t *hit_strand, int64_t cap, int64_t *n_hits) {
    if (!ctx) return CDLM_ERR_ARG;
    cdlm_status s = check_pwm_args(ctx, bases, L, N, pwms, lens, thresh, K, maxlen);
    if (s != CDLM_OK) return s;
    if (!n_hits || cap < 0 || (cap > 0 && (!hit_k || !hit_n || !hit_pos || !hit_strand)))
        return cdlm_fail(ctx, CDLM_ERR_ARG, "bad hit buffers");
    DevPwm d;
    s = upload(ctx, d, bases, L, N, pwms, lens, thresh, K, maxlen);
    int *hk = nullptr, *hp = nullptr;
    long long *hn = nullptr;
    signed char *hs = nullptr;
    unsigned long long *cnt = nullptr;
    const size_t c = cap > 0 ? (size_t)cap : 1;
    auto cleanup = [&]() {
        release(ctx, d, true);
        if (hk) (void)hipFree(hk);
        if (hp) (void)hipFree(hp);
        if (hn) (void)hipFree(hn);
        if (hs) (void)hipFree(hs);
        if (cnt) (void)hipFree(cnt);
    };
    if (s == CDLM_OK && (hipMalloc(&hk, c * 4) != hipSuccess || hipMalloc(&hp, c * 4) != hipSuccess ||
                         hipMalloc(&hn, c * 8) != hipSuccess || hipMalloc(&hs, c) != hipSuccess ||
                         hipMalloc(&cnt, 8) != hipSuccess)) {
        (void)hipGetLastError();
        s = cdlm_fail(ctx, CDLM_ERR_OOM, "no device memory for %lld hit records", (long long)cap);
    }
    if (s == CDLM_OK) s = launch_hits(ctx, d.bases, L, N, d, K, maxlen, both_strands, hk, hn, hp, hs, cap, cnt);
    unsigned long long total = 0;
    if (s == CDLM_OK) {
        hipStream_t st = ctx->stream;
        if (hipMemcpyAsync(&total, cnt, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            s = cdlm_fail(ctx, CDLM_ERR_HIP, "count readback failed");
        const size_t w = total < (unsigned long long)cap ? total : (size_t)cap;
        if (s == CDLM_OK && w > 0) {
            if (hipMemcpyAsync(hit_k, hk, w * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipMemcpyAsync(hit_n, hn, w * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipMemcpyAsync(hit_pos, hp, w * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipMemcpyAsync(hit_strand, hs, w, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                s = cdlm_fail(ctx, CDLM_ERR_HIP, "hit readback failed");
        }
    }
    cleanup();
    if (s != CDLM_OK) return s;
    *n_hits = (int64_t)total;
    if ((int64_t)total > cap) return cdlm_fail(ctx, CDLM_ERR_OOM, "%llu hits exceed cap %lld", total, (long long)cap);
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_pwm_hits_device(cdlm_ctx *ctx, const uint8_t *d_bases, int32_t L, int64_t N,
                                            const float *pwms, const int32_t *lens, const float *thresh, int32_t K,
                                            int32_t maxlen, int32_t both_strands, int32_t *d_hit_k, int64_t *d_hit_n,
                                            int32_t *d_hit_pos, int8_t *d_hit_strand, int64_t cap, int64_t *n_hits) {
    if (!ctx) return CDLM_ERR_ARG;
    cdlm_status s = check_pwm_args(ctx, d_bases, L, N, pwms, lens, thresh, K, maxlen);
    if (s != CDLM_OK) return s;
    if (!n_hits) return cdlm_fail(ctx, CDLM_ERR_ARG, "null n_hits");
    DevPwm d;
    s = upload(ctx, d, nullptr, L, N, pwms, lens, thresh, K, maxlen);
    unsigned long long *cnt = nullptr;
    if (s == CDLM_OK && hipMalloc(&cnt, 8) != hipSuccess) {
        (void)hipGetLastError();
        s = cdlm_fail(ctx, CDLM_ERR_OOM, "counter");
    }
    if (s == CDLM_OK)
        s = launch_hits(ctx, d_bases, L, N, d, K, maxlen, both_strands, d_hit_k, (long long *)d_hit_n, d_hit_pos,
                        (signed char *)d_hit_strand, cap, cnt);
    unsigned long long total = 0;
    if (s == CDLM_OK && (hipMemcpyAsync(&total, cnt, 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
                         hipStreamSynchronize(ctx->stream) != hipSuccess))
        s = cdlm_fail(ctx, CDLM_ERR_HIP, "count readback failed");
    release(ctx, d, false);
    if (cnt) (void)hipFree(cnt);
    if (s != CDLM_OK) return s;
    *n_hits = (int64_t)total;
    if ((int64_t)total > cap) return cdlm_fail(ctx, CDLM_ERR_OOM, "%llu hits exceed cap %lld", total, (long long)cap);
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_pwm_argmax(cdlm_ctx *ctx, const uint8_t *bases, int32_t L, int64_t N, const float *pwms,
                                       const int32_t *lens, const float *thresh, int32_t K, int32_t maxlen,
                                       int32_t *out_pos) {
    if (!ctx) return CDLM_ERR_ARG;
    cdlm_status s = check_pwm_args(ctx, bases, L, N, pwms, lens, thresh, K, maxlen);
    if (s != CDLM_OK) return s;
    if (!out_pos) return cdlm_fail(ctx, CDLM_ERR_ARG, "null out_pos");
    DevPwm d;
    s = upload(ctx, d, bases, L, N, pwms, lens, thresh, K, maxlen);
    int *o = nullptr;
    if (s == CDLM_OK && hipMalloc(&o, sizeof(int) * (size_t)K * N + 4) != hipSuccess) {
        (void)hipGetLastError();
        s = cdlm_fail(ctx, CDLM_ERR_OOM, "out");
    }
    if (s == CDLM_OK && N > 0) {
        const long waves = (long)K * N;
        const long blocks = (waves + 3) / 4;
        hipLaunchKernelGGL(k_pwm_argmax, dim3((unsigned)blocks), dim3(kPwmThreads), 0, ctx->stream, d.bases, L, N,
                           d.pwms, d.lens, d.thresh, K, maxlen, o);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(out_pos, o, sizeof(int) * (size_t)K * N, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            s = cdlm_fail(ctx, CDLM_ERR_HIP, "k_pwm_argmax failed");
    }
    release(ctx, d, true);
    if (o) (void)hipFree(o);
    return s;
}

extern "C" cdlm_status cdlm_pwm_best(cdlm_ctx *ctx, const uint8_t *bases, int32_t L, int64_t N, const float *pwms,
                                     const int32_t *lens, const float *thresh, int32_t K, int32_t maxlen,
                                     int32_t *out_pos, int8_t *out_strand, float *out_score) {
    if (!ctx) return CDLM_ERR_ARG;
    cdlm_status s = check_pwm_args(ctx, bases, L, N, pwms, lens, thresh, K, maxlen);
    if (s != CDLM_OK) return s;
    if (!out_pos || !out_strand || !out_score) return cdlm_fail(ctx, CDLM_ERR_ARG, "null output");
    DevPwm d;
    s = upload(ctx, d, bases, L, N, pwms, lens, thresh, K, maxlen);
    int *op = nullptr;
    signed char *oc = nullptr;
    float *os = nullptr;
    const size_t KN = (size_t)K * N;
    if (s == CDLM_OK && (hipMalloc(&op, KN * 4 + 4) != hipSuccess || hipMalloc(&oc, KN + 4) != hipSuccess ||
                         hipMalloc(&os, KN * 4 + 4) != hipSuccess)) {
        (void)hipGetLastError();
        s = cdlm_fail(ctx, CDLM_ERR_OOM, "out");
    }
    if (s == CDLM_OK && N > 0) {
        const long blocks = ((long)KN + 3) / 4;
        hipLaunchKernelGGL(k_pwm_best, dim3((unsigned)blocks), dim3(kPwmThreads), 0, ctx->stream, d.bases, L, N, d.pwms,
                           d.lens, d.thresh, K, maxlen, op, oc, os);
        hipStream_t st = ctx->stream;
        if (hipGetLastError() != hipSuccess || hipMemcpyAsync(out_pos, op, KN * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(out_strand, oc, KN, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(out_score, os, KN * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            s = cdlm_fail(ctx, CDLM_ERR_HIP, "k_pwm_best failed");
    }
    release(ctx, d, true);
    if (op) (void)hipFree(op);
    if (oc) (void)hipFree(oc);
    if (os) (void)hipFree(os);
    return s;
}

extern "C" cdlm_status cdlm_onehot_to_bases(cdlm_ctx *ctx, const float *onehot, int32_t L4, int64_t N, uint8_t *bases) {
    if (!ctx) return CDLM_ERR_ARG;
    if (!onehot || !bases || L4 <= 0 || L4 % 4 || N < 0) return cdlm_fail(ctx, CDLM_ERR_ARG, "bad arguments");
    if (N == 0) return CDLM_OK;
    const size_t L = L4 / 4;
    float *d_oh = nullptr;
    uint8_t *d_b = nullptr;
    hipStream_t st = ctx->stream;
    cdlm_status s = CDLM_OK;
    if (hipMalloc(&d_oh, sizeof(float) * L4 * (size_t)N) != hipSuccess || hipMalloc(&d_b, L * N) != hipSuccess) {
        (void)hipGetLastError();
        s = cdlm_fail(ctx, CDLM_ERR_OOM, "no memory");
    }
    if (s == CDLM_OK) {
        const long tot = (long)L * N;
        if (hipMemcpyAsync(d_oh, onehot, sizeof(float) * L4 * (size_t)N, hipMemcpyHostToDevice, st) != hipSuccess)
            s = cdlm_fail(ctx, CDLM_ERR_HIP, "upload failed");
        else {
            hipLaunchKernelGGL(k_onehot_to_bases, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, d_oh, L4, N, d_b);
            if (hipGetLastError() != hipSuccess || hipMemcpyAsync(bases, d_b, L * N, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                s = cdlm_fail(ctx, CDLM_ERR_HIP, "k_onehot_to_bases failed");
        }
    }
    if (d_oh) (void)hipFree(d_oh);
    if (d_b) (void)hipFree(d_b);
    return s;
}
