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

#include <algorithm>
#include <climits>
#include <cstring>
#include <mutex>
#include <vector>

#include "cdlm_internal.h"

#include <type_traits>

namespace {

constexpr int kPwmThreads = 256;
// Hit-list kernel shape: G tiles of 16 strand-PWMs per wave (B operands in registers) and CH
// chunks of 16 windows per pass sharing them, so a wave runs G x CH independent MFMA chains.
// Measured at 100000 x 500 bases, 512 PWMs of length 8-24, A operands built from the base
// bytes: (G, CH) = (4, 1) 94.2 ms, (2, 1) 89.3, (2, 2) 86.7, (4, 2) 87.5, (2, 4) 92.1, (1, 8) 89.1,
// (1, 4) 84.6; read as one-hot floats from LDS: (1, 2) 70.6, (1, 4) 71.4, (1, 8) 74.3, (2, 2) 79.0,
// (2, 4) 84.3.
constexpr int kPwmG = 1, kPwmCH = 2;
// LDS per wave position: the sequence expanded to one-hot floats (16 bytes; each A operand
// is one LDS load) while 4 waves' sequences fit 64 KB, else the base bytes (A operands are
// compared from them).
constexpr size_t kPwmLdsMax = 64 * 1024;
constexpr int kMaxPwmLen = 64;

// HIP events around one scan kernel on the context stream; stop() waits for the kernel and
// stores its device time in ctx->last_scan_ms (cdlm_pwm_last_ms).
struct ScanTimer {
    cdlm_ctx *ctx;
    hipEvent_t a = nullptr, b = nullptr;
    explicit ScanTimer(cdlm_ctx *c) : ctx(c) {
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
            (void)hipGetLastError();
            if (a) (void)hipEventDestroy(a);
            a = b = nullptr;
        }
        if (a) (void)hipEventRecord(a, ctx->stream);
    }
    void stop() {
        if (!a) return;
        float ms = 0.f;
        if (hipEventRecord(b, ctx->stream) == hipSuccess && hipEventSynchronize(b) == hipSuccess &&
            hipEventElapsedTime(&ms, a, b) == hipSuccess)
            ctx->last_scan_ms = ms;
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        a = b = nullptr;
    }
    ~ScanTimer() { stop(); }
};

// ---------------------------------------------------------------------------- hit list on MFMA
// One-hot im2col x PWM contraction on v_mfma_f32_16x16x4_f32.  OH: the wave's sequence is held
// in LDS as one-hot floats (position p, base a at float 4p + a), so each A operand is a
// single load; otherwise as base bytes, compared per operand.  A tile of 16 windows x 16
// strand-PWMs accumulates one motif column j per MFMA: A[i][a] = [base(w0+i+j) == a],
// B[a][c] = pwm_c[a][j].  Each K-block holds exactly one non-zero product per output (the
// base's row), so every MFMA step is the single FP32 addition acc + pwm[b][j] of the
// position-major sequential sum (greedy_search!, scan.jl:231-234); columns past a PWM's
// length are zero and add +0.
// Operand layout (16x16x4, one block): lane l supplies A[l%16][l/16] and B[l/16][l%16];
// result register r of lane l is D[4*(l/16)+r][l%16].
typedef float floatx4 __attribute__((ext_vector_type(4)));

struct PwmTiles {
    int T;                 // tiles of 16 strand-PWMs, sorted by length
    int groups;            // tile groups (G tiles per wave)
    const float *B;        // per tile: [maxlen_t][64] lane-ordered B operands
    const int *boff;       // per tile: float offset into B
    const int *tlen;       // per tile: max length (steps)
    const int *slot_k;     // per slot (T*16): PWM index, -1 for an empty slot
    const int *slot_m;     // length
    const int *slot_st;    // strand
    const float *slot_th;  // threshold
};

// Sequence bytes per wave in LDS: chunks of 16 * CH windows, each reading MAXJ bases on.
__host__ __device__ constexpr int pwm_lpad(int L, int MAXJ, int CH) {
    return ((L + 16 * CH - 1) / (16 * CH)) * (16 * CH) + MAXJ + 16;
}

// Sequence prefetch: a wave's next sequence is loaded into registers (position lane + 64 k in
// slot k, base 4 past the end) while the current one is scanned, so its loads are in flight
// together and never wait one after another.  Covers the 64 * kPwmPre positions of any
// one-hot LDS image that fits the scans' LDS budgets.
constexpr int kPwmPre = 16;
static_assert(kPwmLdsMax / (4 * sizeof(float4)) <= 64 * kPwmPre, "one-hot LDS image longer than the prefetch");
struct SeqPre {
    int b[kPwmPre];
    // src: the wave's sequence (wave-uniform), L >= 1 bases, read through a buffer resource of
    // L bytes: loads past L return 0 without touching memory
    __device__ __forceinline__ void load(const uint8_t *src, int L, int lane) {
        const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc((void *)src, (short)0, L, 0x00020000);
#pragma unroll
        for (int k = 0; k < kPwmPre; ++k) {
            const int v = __builtin_amdgcn_raw_buffer_load_b8(r, lane + 64 * k, 0, 0);
            b[k] = lane + 64 * k < L ? v : 4;
        }
    }
};

template <int G, int MAXJ, int CH, bool OH>
__global__ __launch_bounds__(256) void k_pwm_hits_mfma(const uint8_t *__restrict__ bases, int L, long N, PwmTiles P,
                                                       int *hk, long long *hn, int *hp, signed char *hs, long long cap,
                                                       unsigned long long *count) {
    extern __shared__ uint8_t seqs_lds[];
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int Lpad = pwm_lpad(L, MAXJ, CH);
    uint8_t *sq = seqs_lds + w * Lpad;
    float4 *oh = reinterpret_cast<float4 *>(seqs_lds) + (size_t)w * Lpad;
    const float *ohf = reinterpret_cast<const float *>(oh);
    const int g = blockIdx.x;
    const int t0 = g * G;
    float Bv[G][MAXJ];
    int sk[G], sm[G], sst[G];
    float sth[G];
    int gmax = 0;
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
        const int t = t0 + gi;
        const bool ok = t < P.T;
        const int len = ok ? P.tlen[t] : 0;
        gmax = len > gmax ? len : gmax;
#pragma unroll
        for (int j = 0; j < MAXJ; ++j) Bv[gi][j] = (ok && j < len) ? P.B[P.boff[t] + j * 64 + lane] : 0.f;
        const int sl = t * 16 + (lane & 15);
        sk[gi] = ok ? P.slot_k[sl] : -1;
        sm[gi] = ok ? P.slot_m[sl] : 0;
        sst[gi] = ok ? P.slot_st[sl] : 0;
        sth[gi] = ok ? P.slot_th[sl] : 0.f;
    }
    const int arow = lane & 15, abase = lane >> 4;
    const int jsteps = gmax;   // uniform: the tiles of this work-group
    const long nstride = (long)gridDim.y * 4;
    SeqPre pre;
    if (OH && (long)blockIdx.y * 4 + w < N) pre.load(bases + ((long)blockIdx.y * 4 + w) * (long)L, L, lane);
    for (long n = (long)blockIdx.y * 4 + w; n < N; n += nstride) {
        const uint8_t *src = bases + n * (long)L;
        if constexpr (OH) {
#pragma unroll
            for (int k = 0; k < kPwmPre; ++k) {
                const int i = lane + 64 * k;
                const int b = pre.b[k];
                if (64 * k < Lpad && i < Lpad)
                    oh[i] = make_float4(b == 0 ? 1.f : 0.f, b == 1 ? 1.f : 0.f, b == 2 ? 1.f : 0.f, b == 3 ? 1.f : 0.f);
            }
            if (n + nstride < N) pre.load(src + nstride * (long)L, L, lane);
        } else {
            for (int i = lane; i < Lpad; i += 64) sq[i] = i < L ? src[i] : (uint8_t)4;
        }
        wave_lds_sync();
        for (int w0 = 0; w0 < L; w0 += 16 * CH) {
            // CH chunks of 16 windows share the B operands: G x CH independent accumulators
            floatx4 acc[CH][G];
#pragma unroll
            for (int c = 0; c < CH; ++c)
#pragma unroll
                for (int gi = 0; gi < G; ++gi) acc[c][gi] = floatx4{0.f, 0.f, 0.f, 0.f};
            // the group's positions (B is zero past each PWM's length); one uniform switch per
            // chunk keeps every case branch-free, so the accumulators stay in place across the
            // MFMA chain
#define CDLM_MFMA_STEPS(JS)                                                                          \
    case JS:                                                                                         \
        _Pragma("unroll") for (int j = 0; j < JS; ++j) {                                             \
            _Pragma("unroll") for (int c = 0; c < CH; ++c) {                                         \
                const int p = w0 + 16 * c + arow + j;                                                \
                const float a = OH ? ohf[4 * p + abase] : ((sq[p] == abase) ? 1.0f : 0.0f);          \
                _Pragma("unroll") for (int gi = 0; gi < G; ++gi)                                     \
                    acc[c][gi] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, Bv[gi][j], acc[c][gi], 0, 0, 0); \
            }                                                                                        \
        }                                                                                            \
        break;
#define CDLM_MFMA_STEPS8(J0)                                                                         \
    CDLM_MFMA_STEPS(J0 + 1) CDLM_MFMA_STEPS(J0 + 2) CDLM_MFMA_STEPS(J0 + 3) CDLM_MFMA_STEPS(J0 + 4)  \
    CDLM_MFMA_STEPS(J0 + 5) CDLM_MFMA_STEPS(J0 + 6) CDLM_MFMA_STEPS(J0 + 7) CDLM_MFMA_STEPS(J0 + 8)
            switch (jsteps) {
                CDLM_MFMA_STEPS8(0) CDLM_MFMA_STEPS8(8) CDLM_MFMA_STEPS8(16) CDLM_MFMA_STEPS8(24)
                default:
                    if constexpr (MAXJ > 32) {
                        switch (jsteps) {
                            CDLM_MFMA_STEPS8(32) CDLM_MFMA_STEPS8(40) CDLM_MFMA_STEPS8(48) CDLM_MFMA_STEPS8(56)
                            default: break;
                        }
                    }
                    break;
            }
#undef CDLM_MFMA_STEPS8
#undef CDLM_MFMA_STEPS
            // hits are rare: one comparison of each group's largest score first, the per-window
            // tests only when some lane of the wave has a score above its threshold
            bool any = false;
#pragma unroll
            for (int gi = 0; gi < G; ++gi) {
                float mx = acc[0][gi][0];
#pragma unroll
                for (int c = 0; c < CH; ++c)
#pragma unroll
                    for (int r = 0; r < 4; ++r) mx = fmaxf(mx, acc[c][gi][r]);
                any = any || mx > sth[gi];
            }
            if (!__ballot(any)) continue;
            int myc = 0;
            unsigned hitmask = 0;
#pragma unroll
            for (int c = 0; c < CH; ++c)
#pragma unroll
                for (int gi = 0; gi < G; ++gi)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int l0 = w0 + 16 * c + 4 * abase + r;
                        const bool h = sk[gi] >= 0 && l0 + sm[gi] <= L && acc[c][gi][r] > sth[gi];
                        if (h) { hitmask |= 1u << ((c * G + gi) * 4 + r); ++myc; }
                    }
            if (__ballot(myc > 0)) {
                int incl = myc;
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    const int t = __shfl_up(incl, o);
                    if (lane >= o) incl += t;
                }
                const int total = __shfl(incl, 63);
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(count, (unsigned long long)total);
                base = __shfl(base, 0);
                unsigned long long slot = base + (unsigned long long)(incl - myc);
#pragma unroll
                for (int c = 0; c < CH; ++c)
#pragma unroll
                    for (int gi = 0; gi < G; ++gi)
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            if (hitmask & (1u << ((c * G + gi) * 4 + r))) {
                                if ((long long)slot < cap) {
                                    hk[slot] = sk[gi]; hn[slot] = n; hp[slot] = w0 + 16 * c + 4 * abase + r;
                                    hs[slot] = (signed char)sst[gi];
                                }
                                ++slot;
                            }
            }
        }
        wave_lds_sync();
    }
}

// ---------------------------------------------------------------------------- best window on MFMA
// Per strand-PWM slot and sequence: the first window of largest value, on the hit list's one-hot
// im2col x PWM contraction (tiles, B operands and sequence in LDS as there).
//   NUC = false, scan_w_gpu! (scan.jl:176-221): the position-major sum of greedy_search!
//     (:231-234); value = score > thresh ? score : 0, and windows past the end (l + m > L, l < L)
//     are 0 as in the reference's dense array.
//   NUC = true, scan_n (scan.jl:10-21) for overlapping_scan!: the nucleotide-major sum, one pass
//     per base a with A[i][a'] = [a' == a][base(w0 + i + j) == a], so every MFMA step still adds
//     a single product (pwm_c[a][j] or +0) in the order a = A, C, G, T, j ascending; value =
//     score; only windows that fit compete.
// The first maximum wins (strict >; across lanes the smaller window).  Output per slot s and
// sequence n: bv[s N + n], bp[s N + n] (0-based window; INT_MAX when no window competes).
template <int MAXJ, int CH, bool OH, bool NUC>
__global__ __launch_bounds__(256) void k_pwm_best_mfma(const uint8_t *__restrict__ bases, int L, long N, PwmTiles P,
                                                       float *__restrict__ bv, int *__restrict__ bp) {
    extern __shared__ uint8_t seqs_lds[];
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int Lpad = pwm_lpad(L, MAXJ, CH);
    uint8_t *sq = seqs_lds + w * Lpad;
    float4 *oh = reinterpret_cast<float4 *>(seqs_lds) + (size_t)w * Lpad;
    const float *ohf = reinterpret_cast<const float *>(oh);
    const int t = blockIdx.x;
    const bool ok = t < P.T;
    const int jsteps = ok ? P.tlen[t] : 0;    // uniform: the tile's longest PWM
    float Bv[MAXJ];
#pragma unroll
    for (int j = 0; j < MAXJ; ++j) Bv[j] = (ok && j < jsteps) ? P.B[P.boff[t] + j * 64 + lane] : 0.f;
    const int sl = t * 16 + (lane & 15);
    const int sk = ok ? P.slot_k[sl] : -1, sm = ok ? P.slot_m[sl] : 0;
    const float sth = ok ? P.slot_th[sl] : 0.f;
    const int arow = lane & 15, abase = lane >> 4;
    const long nstride = (long)gridDim.y * 4;
    SeqPre pre;
    if (OH && (long)blockIdx.y * 4 + w < N) pre.load(bases + ((long)blockIdx.y * 4 + w) * (long)L, L, lane);
    for (long n = (long)blockIdx.y * 4 + w; n < N; n += nstride) {
        const uint8_t *src = bases + n * (long)L;
        if constexpr (OH) {
#pragma unroll
            for (int k = 0; k < kPwmPre; ++k) {
                const int i = lane + 64 * k;
                const int b = pre.b[k];
                if (64 * k < Lpad && i < Lpad)
                    oh[i] = make_float4(b == 0 ? 1.f : 0.f, b == 1 ? 1.f : 0.f, b == 2 ? 1.f : 0.f, b == 3 ? 1.f : 0.f);
            }
            if (n + nstride < N) pre.load(src + nstride * (long)L, L, lane);
        } else {
            for (int i = lane; i < Lpad; i += 64) sq[i] = i < L ? src[i] : (uint8_t)4;
        }
        wave_lds_sync();
        float best = 0.f;
        int bi = INT_MAX;
        for (int w0 = 0; w0 < L; w0 += 16 * CH) {
            floatx4 acc[CH];
#pragma unroll
            for (int c = 0; c < CH; ++c) acc[c] = floatx4{0.f, 0.f, 0.f, 0.f};
            for (int a = 0; a < (NUC ? 4 : 1); ++a) {
#define CDLM_BEST_STEPS(JS)                                                                              \
    case JS:                                                                                             \
        _Pragma("unroll") for (int j = 0; j < JS; ++j) {                                                 \
            _Pragma("unroll") for (int c = 0; c < CH; ++c) {                                             \
                const int p = w0 + 16 * c + arow + j;                                                    \
                float av = OH ? ohf[4 * p + abase] : ((sq[p] == abase) ? 1.0f : 0.0f);                   \
                if (NUC) av = (abase == a) ? av : 0.0f;                                                  \
                acc[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, Bv[j], acc[c], 0, 0, 0);               \
            }                                                                                            \
        }                                                                                                \
        break;
#define CDLM_BEST_STEPS8(J0)                                                                             \
    CDLM_BEST_STEPS(J0 + 1) CDLM_BEST_STEPS(J0 + 2) CDLM_BEST_STEPS(J0 + 3) CDLM_BEST_STEPS(J0 + 4)      \
    CDLM_BEST_STEPS(J0 + 5) CDLM_BEST_STEPS(J0 + 6) CDLM_BEST_STEPS(J0 + 7) CDLM_BEST_STEPS(J0 + 8)
                switch (jsteps) {
                    CDLM_BEST_STEPS8(0) CDLM_BEST_STEPS8(8) CDLM_BEST_STEPS8(16) CDLM_BEST_STEPS8(24)
                    default:
                        if constexpr (MAXJ > 32) {
                            switch (jsteps) {
                                CDLM_BEST_STEPS8(32) CDLM_BEST_STEPS8(40) CDLM_BEST_STEPS8(48) CDLM_BEST_STEPS8(56)
                                default: break;
                            }
                        }
                        break;
                }
#undef CDLM_BEST_STEPS8
#undef CDLM_BEST_STEPS
            }
            if (!NUC) {
                // thresholded scores are rare: when no lane of the wave has a score above its
                // threshold, every window's value is 0 and only a lane's first window of the chunk
                // (w0 + 4 abase, the smallest) can be taken, as the per-window rule below would
                float mx = acc[0][0];
#pragma unroll
                for (int c = 0; c < CH; ++c)
#pragma unroll
                    for (int r = 0; r < 4; ++r) mx = fmaxf(mx, acc[c][r]);
                if (!__ballot(mx > sth)) {
                    const int lf = w0 + 4 * abase;
                    const bool take = sk >= 0 && lf < L && (bi == INT_MAX || 0.f > best);
                    best = take ? 0.f : best;
                    bi = take ? lf : bi;
                    continue;
                }
            }
            // this lane's windows of its slot, in increasing order: l = w0 + 16 c + 4 abase + r
#pragma unroll
            for (int c = 0; c < CH; ++c)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int l = w0 + 16 * c + 4 * abase + r;
                    const bool fits = l + sm <= L;
                    const bool valid = sk >= 0 && l < L && (!NUC || fits);
                    float v = acc[c][r];
                    if (!NUC) v = (fits && v > sth) ? v : 0.f;
                    const bool take = valid && (bi == INT_MAX || v > best);
                    best = take ? v : best;
                    bi = take ? l : bi;
                }
        }
        // the four lanes of a slot: larger value, then smaller window
#pragma unroll
        for (int off = 16; off < 64; off <<= 1) {
            const float ob = __shfl_xor(best, off);
            const int oi = __shfl_xor(bi, off);
            if (oi != INT_MAX && (bi == INT_MAX || ob > best || (ob == best && oi < bi))) { best = ob; bi = oi; }
        }
        if (lane < 16 && sk >= 0) {
            bv[(size_t)sl * N + n] = best;
            bp[(size_t)sl * N + n] = bi;
        }
        wave_lds_sync();
    }
}

// ---------------------------------------------------------------------------- best hit on the K=1 MFMA
// The nucleotide-major best hit (NUC mode above) on v_mfma_f32_32x32x1_2b_f32, whose single
// K row makes every product useful: each step adds pwm_c[a][j] or +0, so the four one-base passes
// cost what one position-major pass of the K=4 form costs.  Layout (measured): block b uses A and
// B lanes 32b..32b+31; result register r of lane l is block r/16, row 8((r%16)/4) + 4(l/32) + r%4,
// column l%32.  Both blocks take the tile's 32 strand-PWM slots, block b the windows
// w0 + 32b + row; a wave runs two such 64-window chunks (w0, w0 + 64) as two independent MFMA
// chains sharing each B load, base by base (a outer, j inner, so the accumulators stay in
// place).  LDS: the tile's B values [a][j][32 slots] | per wave the sequence as four one-hot
// planes of LP floats.  Same outputs as k_pwm_best_mfma (slot s = 32 t + column).
constexpr int kBest32Lds = 64 * 1024;
static_assert(kBest32Lds / (4 * 4 * sizeof(float)) <= 64 * kPwmPre, "one-hot planes longer than the prefetch");
__host__ __device__ constexpr int best32_lpad(int L, int MAXJ) { return ((L + 127) / 128) * 128 + MAXJ + 16; }

template <int MAXJ>
__global__ __launch_bounds__(256, MAXJ > 32 ? 2 : 3) void k_pwm_best32_mfma(const uint8_t *__restrict__ bases, int L, long N,
                                                         const float *__restrict__ B32, const int *__restrict__ tlen,
                                                         const int *__restrict__ slot_k, const int *__restrict__ slot_m,
                                                         int T, float *__restrict__ bv, int *__restrict__ bp) {
    typedef float floatx32 __attribute__((ext_vector_type(32)));
    extern __shared__ float lds32[];
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int LP = best32_lpad(L, MAXJ);
    float *bt = lds32;                                   // 4 * MAXJ * 32
    float *oh = lds32 + 4 * MAXJ * 32 + (size_t)w * 4 * LP;
    const int t = blockIdx.x;
    const int jsteps = t < T ? tlen[t] : 0;
    for (int i = threadIdx.x; i < 4 * MAXJ * 32; i += 256) {
        const int a = i / (MAXJ * 32), j = (i / 32) % MAXJ;
        bt[i] = (t < T && j < jsteps) ? B32[((size_t)t * 4 + a) * MAXJ * 32 + (i % (MAXJ * 32))] : 0.f;
    }
    __syncthreads();
    const int col = lane & 31, half = lane >> 5;
    const int sl = t * 32 + col;
    const int sk = t < T ? slot_k[sl] : -1, sm = t < T ? slot_m[sl] : 0;
    const long nstride = (long)gridDim.y * 4;
    SeqPre pre;
    if ((long)blockIdx.y * 4 + w < N) pre.load(bases + ((long)blockIdx.y * 4 + w) * (long)L, L, lane);
    for (long n = (long)blockIdx.y * 4 + w; n < N; n += nstride) {
        const uint8_t *src = bases + n * (long)L;
#pragma unroll
        for (int k = 0; k < kPwmPre; ++k) {
            const int i = lane + 64 * k;
            const int b = pre.b[k];
            if (64 * k < LP && i < LP) {
#pragma unroll
                for (int a = 0; a < 4; ++a) oh[a * LP + i] = b == a ? 1.f : 0.f;
            }
        }
        if (n + nstride < N) pre.load(src + nstride * (long)L, L, lane);
        wave_lds_sync();
        float best = 0.f;
        int bi = INT_MAX;
        for (int w0 = 0; w0 < L; w0 += 128) {
            // two chunks of 64 windows (w0 and w0 + 64) share each B load: two independent chains
            floatx32 acc0, acc1;
#pragma unroll
            for (int r = 0; r < 32; ++r) { acc0[r] = 0.f; acc1[r] = 0.f; }
            auto chain = [&](auto js_c) {
                constexpr int JS = decltype(js_c)::value;
#pragma unroll 1
                for (int a = 0; a < 4; ++a) {
                    const float *pa = oh + a * LP + w0 + lane;     // window w0 + lane (row lane%32 of block lane/32)
                    const float *ba = bt + a * MAXJ * 32 + col;
#pragma unroll
                    for (int j = 0; j < JS; ++j) {
                        const float bj = ba[32 * j];
                        acc0 = __builtin_amdgcn_mfma_f32_32x32x1f32(pa[j], bj, acc0, 0, 0, 0);
                        acc1 = __builtin_amdgcn_mfma_f32_32x32x1f32(pa[64 + j], bj, acc1, 0, 0, 0);
                    }
                }
            };
#define CDLM_B32_STEPS(JS) case JS: chain(std::integral_constant<int, JS>{}); break;
#define CDLM_B32_STEPS8(J0)                                                                              \
    CDLM_B32_STEPS(J0 + 1) CDLM_B32_STEPS(J0 + 2) CDLM_B32_STEPS(J0 + 3) CDLM_B32_STEPS(J0 + 4)          \
    CDLM_B32_STEPS(J0 + 5) CDLM_B32_STEPS(J0 + 6) CDLM_B32_STEPS(J0 + 7) CDLM_B32_STEPS(J0 + 8)
            switch (jsteps) {
                CDLM_B32_STEPS8(0) CDLM_B32_STEPS8(8) CDLM_B32_STEPS8(16) CDLM_B32_STEPS8(24)
                default:
                    if constexpr (MAXJ > 32) {
                        switch (jsteps) {
                            CDLM_B32_STEPS8(32) CDLM_B32_STEPS8(40) CDLM_B32_STEPS8(48) CDLM_B32_STEPS8(56)
                            default: break;
                        }
                    }
                    break;
            }
#undef CDLM_B32_STEPS8
#undef CDLM_B32_STEPS
            // this lane's windows in increasing order: chunk, block r/16, row 8((r%16)/4) + 4 half + r%4
            if (w0 + 127 + jsteps <= L) {
                // every window of both chunks fits every slot of the tile: the first largest of a
                // chunk by a pairwise tree (the later of two is taken only when strictly larger,
                // as the sequential rule below), then the chunks in order
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    float tv[32];
                    int to[32];
#pragma unroll
                    for (int r = 0; r < 32; ++r) {
                        tv[r] = c ? acc1[r] : acc0[r];
                        to[r] = 32 * (r >> 4) + 8 * ((r & 15) >> 2) + (r & 3);
                    }
#pragma unroll
                    for (int st = 1; st < 32; st <<= 1)
#pragma unroll
                        for (int i = 0; i < 32; i += 2 * st) {
                            const bool t = tv[i + st] > tv[i];
                            tv[i] = t ? tv[i + st] : tv[i];
                            to[i] = t ? to[i + st] : to[i];
                        }
                    const bool take = sk >= 0 && (bi == INT_MAX || tv[0] > best);
                    best = take ? tv[0] : best;
                    bi = take ? w0 + 64 * c + 4 * half + to[0] : bi;
                }
            } else {
#pragma unroll
                for (int c = 0; c < 2; ++c)
#pragma unroll
                    for (int r = 0; r < 32; ++r) {
                        const int l = w0 + 64 * c + 32 * (r >> 4) + 8 * ((r & 15) >> 2) + 4 * half + (r & 3);
                        const float v = c ? acc1[r] : acc0[r];
                        const bool take = sk >= 0 && l + sm <= L && (bi == INT_MAX || v > best);
                        best = take ? v : best;
                        bi = take ? l : bi;
                    }
            }
        }
        // the two lanes of a slot: larger value, then smaller window
        {
            const float ob = __shfl_xor(best, 32);
            const int oi = __shfl_xor(bi, 32);
            if (oi != INT_MAX && (bi == INT_MAX || ob > best || (ob == best && oi < bi))) { best = ob; bi = oi; }
        }
        if (half == 0 && sk >= 0) {
            bv[(size_t)sl * N + n] = best;
            bp[(size_t)sl * N + n] = bi;
        }
        wave_lds_sync();
    }
}

// scan_w_gpu! positions from the forward slots (slot_of[2k]): 1-based window, 0 for none or
// for window 1 (scan.jl:197).
__global__ void k_pwm_argmax_out(const int *__restrict__ slot_of, int K, long N, const int *__restrict__ bp,
                                 int *__restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)K * N) return;
    const int k = (int)(i / N);
    const long n = i - (long)k * N;
    const int b = bp[(size_t)slot_of[2 * k] * N + n];
    out[i] = (b == INT_MAX || b == 0) ? 0 : b + 1;
}

// overlapping_scan! (scan.jl:89-121): the forward strand wins ties (>=); a strand with no
// window scores -Inf; hit iff the chosen score > thresh.
__global__ void k_pwm_best_out(const int *__restrict__ slot_of, const float *__restrict__ thresh, int K, long N,
                               const float *__restrict__ bv, const int *__restrict__ bp, int *__restrict__ opos,
                               signed char *__restrict__ ocomp, float *__restrict__ oscore) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)K * N) return;
    const int k = (int)(i / N);
    const long n = i - (long)k * N;
    const size_t f = (size_t)slot_of[2 * k] * N + n, c = (size_t)slot_of[2 * k + 1] * N + n;
    const int pf = bp[f], pc = bp[c];
    const float vf = pf == INT_MAX ? -INFINITY : bv[f], vc = pc == INT_MAX ? -INFINITY : bv[c];
    const float th = thresh[k];
    int pos = 0, comp = 0;
    float sc = 0.f;
    if (vf >= vc) {
        if (pf != INT_MAX && vf > th) { pos = pf + 1; comp = 0; sc = vf; }
    } else {
        if (pc != INT_MAX && vc > th) { pos = pc + 1; comp = 1; sc = vc; }
    }
    opos[i] = pos;
    ocomp[i] = (signed char)comp;
    oscore[i] = sc;
}

// Lane layout and rounding probe of v_mfma_f32_16x16x4_f32: lane l supplies a[l] and b[l];
// c holds the 4 accumulator inputs of each lane; d receives the 4 results of each lane.
__global__ void k_mfma_probe(const float *a, const float *b, const float *c, float *d) {
    const int l = threadIdx.x;
    floatx4 acc = {c[4 * l], c[4 * l + 1], c[4 * l + 2], c[4 * l + 3]};
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[l], b[l], acc, 0, 0, 0);
    for (int r = 0; r < 4; ++r) d[4 * l + r] = acc[r];
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

// Strand-PWM tiles for the MFMA kernel: (k, strand) pairs sorted by length (longest
// first, then k, strand), 16 per tile.
struct HostTiles {
    std::vector<float> B;
    std::vector<int> boff, tlen, sk, sm, sst;
    std::vector<float> sth;
    std::vector<int> slot_of;   // (K, 2): slot of PWM k on strand st, -1 if not scanned
    int T = 0;
};

void build_tiles(const float *pwms, const int *lens, const float *thresh, int K, int maxlen, int both, HostTiles &ht) {
    struct SP { int k, st, m; };
    std::vector<SP> v;
    for (int k = 0; k < K; ++k)
        for (int st = 0; st < (both ? 2 : 1); ++st) v.push_back({k, st, lens[k]});
    std::stable_sort(v.begin(), v.end(), [](const SP &x, const SP &y) { return x.m > y.m; });
    ht.T = (int)((v.size() + 15) / 16);
    ht.slot_of.assign(2 * (size_t)K, -1);
    for (size_t i = 0; i < v.size(); ++i) ht.slot_of[2 * (size_t)v[i].k + v[i].st] = (int)i;
    ht.sk.assign(ht.T * 16, -1); ht.sm.assign(ht.T * 16, 0); ht.sst.assign(ht.T * 16, 0); ht.sth.assign(ht.T * 16, 0.f);
    for (int t = 0; t < ht.T; ++t) {
        int tl = 0;
        for (int c = 0; c < 16; ++c) {
            const size_t i = (size_t)t * 16 + c;
            if (i < v.size()) tl = v[i].m > tl ? v[i].m : tl;
        }
        ht.tlen.push_back(tl);
        ht.boff.push_back((int)ht.B.size());
        for (int j = 0; j < tl; ++j)
            for (int lane = 0; lane < 64; ++lane) {
                const int a = lane >> 4, c = lane & 15;
                const size_t i = (size_t)t * 16 + c;
                float val = 0.f;
                if (i < v.size() && j < v[i].m) {
                    const int k = v[i].k, m = v[i].m;
                    const float *pw = pwms + (size_t)k * 4 * maxlen;
                    val = v[i].st ? pw[(3 - a) * maxlen + (m - 1 - j)] : pw[a * maxlen + j];
                }
                ht.B.push_back(val);
            }
        for (int c = 0; c < 16; ++c) {
            const size_t i = (size_t)t * 16 + c;
            if (i >= v.size()) continue;
            ht.sk[i] = v[i].k; ht.sm[i] = v[i].m; ht.sst[i] = v[i].st; ht.sth[i] = thresh[v[i].k];
        }
    }
}

// Tiles of 32 strand-PWM slots for k_pwm_best32_mfma: B values [tile][a][j < MAXJ][slot], the
// same length-sorted slot order as build_tiles.
struct HostTiles32 {
    std::vector<float> B;
    std::vector<int> tlen, sk, sm, slot_of;
    int T = 0;
};

void build_tiles32(const float *pwms, const int *lens, int K, int maxlen, int MAXJ, HostTiles32 &ht) {
    struct SP { int k, st, m; };
    std::vector<SP> v;
    for (int k = 0; k < K; ++k)
        for (int st = 0; st < 2; ++st) v.push_back({k, st, lens[k]});
    std::stable_sort(v.begin(), v.end(), [](const SP &x, const SP &y) { return x.m > y.m; });
    ht.T = (int)((v.size() + 31) / 32);
    ht.slot_of.assign(2 * (size_t)K, -1);
    ht.sk.assign((size_t)ht.T * 32, -1);
    ht.sm.assign((size_t)ht.T * 32, 0);
    ht.B.assign((size_t)ht.T * 4 * MAXJ * 32, 0.f);
    ht.tlen.assign(ht.T, 0);
    for (size_t i = 0; i < v.size(); ++i) {
        const int t = (int)(i / 32), c = (int)(i % 32), k = v[i].k, m = v[i].m;
        ht.slot_of[2 * (size_t)k + v[i].st] = (int)i;
        ht.sk[i] = k;
        ht.sm[i] = m;
        ht.tlen[t] = std::max(ht.tlen[t], m);
        const float *pw = pwms + (size_t)k * 4 * maxlen;
        for (int a = 0; a < 4; ++a)
            for (int j = 0; j < m; ++j)
                ht.B[(((size_t)t * 4 + a) * MAXJ + j) * 32 + c] = v[i].st ? pw[(3 - a) * maxlen + (m - 1 - j)] : pw[a * maxlen + j];
    }
}

struct DevTiles {
    float *B = nullptr, *sth = nullptr;
    int *boff = nullptr, *tlen = nullptr, *sk = nullptr, *sm = nullptr, *sst = nullptr;
    void release() {
        for (void *q : {(void *)B, (void *)sth, (void *)boff, (void *)tlen, (void *)sk, (void *)sm, (void *)sst})
            if (q) (void)hipFree(q);
    }
};

template <typename T>
cdlm_status up(cdlm_ctx *ctx, T **dst, const std::vector<T> &src) {
    CDLM_HIP(ctx, hipMalloc(dst, sizeof(T) * (src.size() ? src.size() : 1)));
    if (!src.empty()) CDLM_HIP(ctx, hipMemcpyAsync(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice, ctx->stream));
    return CDLM_OK;
}

// Sequence blocks per tile group for the scan kernels, which step through the sequences with
// a fixed stride: 64 rounds of the work-groups the device holds at once.  The work-groups have
// equal work, and the finer the grain the less the device idles while the last ones finish
// (4, 8, 16, 32 and 64 rounds measured; the gain levels off at 32-64).
long scan_blocks(cdlm_ctx *ctx, const void *kfn, size_t lds, long groups, long N) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, 256, lds) != hipSuccess || per_cu < 1) {
        (void)hipGetLastError();
        per_cu = 1;
    }
    long yb = (long)ctx->num_cu * per_cu * 64 / (groups > 0 ? groups : 1);
    const long most = (N + 3) / 4;
    if (yb > most) yb = most;
    if (yb > 65535) yb = 65535;
    if (yb < 1) yb = 1;
    return yb;
}

// Launches k_pwm_hits_mfma over strand-PWM tiles already on the device (P, T tiles of up to
// maxlen columns) with the hit counter already zeroed.  timed: HIP events around the launch
// give ctx->last_scan_ms (waits for the kernel); otherwise the launch is only enqueued.
cdlm_status run_hits_kernel(cdlm_ctx *ctx, const uint8_t *d_bases, int L, long N, int maxlen, PwmTiles P, int *hk,
                            long long *hn, int *hp, signed char *hs, long long cap, unsigned long long *d_count,
                            bool timed) {
    hipStream_t st = ctx->stream;
    const int G = kPwmG;
    const int MAXJ = maxlen <= 32 ? 32 : 64;
    P.groups = (P.T + G - 1) / G;
    const int Lpad = pwm_lpad(L, MAXJ, kPwmCH);
    const bool oh = 4 * (size_t)Lpad * sizeof(float4) <= kPwmLdsMax;
    const size_t lds = 4 * (size_t)Lpad * (oh ? sizeof(float4) : 1);
    if (lds > kPwmLdsMax) return cdlm_fail(ctx, CDLM_ERR_UNSUPPORTED, "sequence length %d too long for the scan", L);
    const void *kfn = maxlen <= 32 ? (oh ? reinterpret_cast<const void *>(&k_pwm_hits_mfma<kPwmG, 32, kPwmCH, true>)
                                         : reinterpret_cast<const void *>(&k_pwm_hits_mfma<kPwmG, 32, kPwmCH, false>))
                                   : (oh ? reinterpret_cast<const void *>(&k_pwm_hits_mfma<kPwmG, 64, kPwmCH, true>)
                                         : reinterpret_cast<const void *>(&k_pwm_hits_mfma<kPwmG, 64, kPwmCH, false>));
    const long yb = scan_blocks(ctx, kfn, lds, P.groups, N);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (timed && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) {
        (void)hipGetLastError();
        if (e0) (void)hipEventDestroy(e0);
        e0 = e1 = nullptr;
    }
    if (e0) (void)hipEventRecord(e0, st);
#define L_HITS(MJ, OHV)                                                                                         \
    hipLaunchKernelGGL((k_pwm_hits_mfma<kPwmG, MJ, kPwmCH, OHV>), dim3(P.groups, (unsigned)yb), dim3(256), lds, st, \
                       d_bases, L, N, P, hk, hn, hp, hs, cap, d_count)
    if (maxlen <= 32 && oh) L_HITS(32, true);
    else if (maxlen <= 32) L_HITS(32, false);
    else if (oh) L_HITS(64, true);
    else L_HITS(64, false);
#undef L_HITS
    hipError_t e = hipGetLastError();
    cdlm_status s = e == hipSuccess ? CDLM_OK : cdlm_fail(ctx, CDLM_ERR_HIP, "k_pwm_hits_mfma: %s", hipGetErrorString(e));
    if (e1) {
        (void)hipEventRecord(e1, st);
        float ms = 0.f;
        if (hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) ctx->last_scan_ms = ms;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return s;
}

cdlm_status launch_hits(cdlm_ctx *ctx, const uint8_t *d_bases, int L, long N, const float *pwms, const int *lens,
                        const float *thresh, int K, int maxlen, int both, int *hk, long long *hn, int *hp,
                        signed char *hs, long long cap, unsigned long long *d_count) {
    hipStream_t st = ctx->stream;
    CDLM_HIP(ctx, hipMemsetAsync(d_count, 0, sizeof(unsigned long long), st));
    if (N == 0) return CDLM_OK;
    HostTiles ht;
    build_tiles(pwms, lens, thresh, K, maxlen, both, ht);
    DevTiles dt;
    cdlm_status s = CDLM_OK;
    if ((s = up(ctx, &dt.B, ht.B)) == CDLM_OK && (s = up(ctx, &dt.boff, ht.boff)) == CDLM_OK &&
        (s = up(ctx, &dt.tlen, ht.tlen)) == CDLM_OK && (s = up(ctx, &dt.sk, ht.sk)) == CDLM_OK &&
        (s = up(ctx, &dt.sm, ht.sm)) == CDLM_OK && (s = up(ctx, &dt.sst, ht.sst)) == CDLM_OK &&
        (s = up(ctx, &dt.sth, ht.sth)) == CDLM_OK) {
        PwmTiles P;
        P.T = ht.T; P.B = dt.B; P.boff = dt.boff; P.tlen = dt.tlen;
        P.slot_k = dt.sk; P.slot_m = dt.sm; P.slot_st = dt.sst; P.slot_th = dt.sth;
        s = run_hits_kernel(ctx, d_bases, L, N, maxlen, P, hk, hn, hp, hs, cap, d_count, true);
    }
    (void)hipStreamSynchronize(st);
    dt.release();
    return s;
}

// ---------------------------------------------------------------------------- fixed scratch
// The allocation-free hit list (cdlm_pwm_hits_device_noalloc) keeps its strand-PWM tiles and
// hit counter in this file-scope array instead of buffers allocated per call.  Each GPU holds
// its own copy until the process ends; host code never passes the array's name to a kernel
// (on the host it is only a stand-in for the symbol), but the device address that
// hipGetSymbolAddress returns for the GPU, resolved once per GPU when a context is created.
}  // namespace
static __device__ __align__(256) unsigned char scratch[CDLM_PWM_SCRATCH_BYTES];
namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxScratchDevices] = {};

__global__ void k_zero_count(unsigned long long *c) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0ull;
}

// Byte offsets of the tile arrays inside the scratch array (each 256-byte aligned); total
// bytes in *total.
struct ScratchLayout {
    size_t count, B, boff, tlen, sk, sm, sst, sth, total;
};
ScratchLayout scratch_layout(size_t nB, size_t T) {
    auto up256 = [](size_t x) { return (x + 255) & ~(size_t)255; };
    ScratchLayout o;
    size_t at = 0;
    o.count = at; at += up256(sizeof(unsigned long long));
    o.B = at; at += up256(nB * sizeof(float));
    o.boff = at; at += up256(T * sizeof(int));
    o.tlen = at; at += up256(T * sizeof(int));
    o.sk = at; at += up256(T * 16 * sizeof(int));
    o.sm = at; at += up256(T * 16 * sizeof(int));
    o.sst = at; at += up256(T * 16 * sizeof(int));
    o.sth = at; at += up256(T * 16 * sizeof(float));
    o.total = at;
    return o;
}


// Best window per (PWM, sequence) on MFMA: the argmax of scan_w_gpu! (nuc = false, forward
// strand) into o_pos, or the best hit of overlapping_scan! (nuc = true, both strands) into
// o_pos, o_comp, o_score; device pointers, (K, N).  CDLM_ERR_UNSUPPORTED (no error recorded)
// when four waves' sequences do not fit the LDS: the caller then runs the per-(k, n) kernels.
cdlm_status launch_best(cdlm_ctx *ctx, const uint8_t *d_bases, int L, long N, const float *pwms, const int *lens,
                        const float *thresh, const float *d_thresh, int K, int maxlen, bool nuc, int *o_pos,
                        signed char *o_comp, float *o_score) {
    const int MAXJ = maxlen <= 32 ? 32 : 64;
    const int Lpad = pwm_lpad(L, MAXJ, kPwmCH);
    const bool oh = 4 * (size_t)Lpad * sizeof(float4) <= kPwmLdsMax;
    const size_t lds = 4 * (size_t)Lpad * (oh ? sizeof(float4) : 1);
    if (lds > kPwmLdsMax) return CDLM_ERR_UNSUPPORTED;
    if (N == 0) return CDLM_OK;
    hipStream_t st = ctx->stream;
    // the best hit on the K=1 MFMA when its LDS (tile B values and four one-hot planes per
    // wave) fits
    const size_t lds32 = sizeof(float) * (4 * (size_t)MAXJ * 32 + 4 * 4 * (size_t)best32_lpad(L, MAXJ));
    if (nuc && lds32 <= (size_t)kBest32Lds) {
        HostTiles32 h32;
        build_tiles32(pwms, lens, K, maxlen, MAXJ, h32);
        float *dB = nullptr, *bval = nullptr;
        int *dtl = nullptr, *dsk = nullptr, *dsm = nullptr, *dso = nullptr, *bpos = nullptr;
        cdlm_status s = CDLM_OK;
        if ((s = up(ctx, &dB, h32.B)) == CDLM_OK && (s = up(ctx, &dtl, h32.tlen)) == CDLM_OK &&
            (s = up(ctx, &dsk, h32.sk)) == CDLM_OK && (s = up(ctx, &dsm, h32.sm)) == CDLM_OK &&
            (s = up(ctx, &dso, h32.slot_of)) == CDLM_OK) {
            const size_t nslot = (size_t)h32.T * 32;
            if (hipMalloc(&bval, sizeof(float) * nslot * N) != hipSuccess || hipMalloc(&bpos, sizeof(int) * nslot * N) != hipSuccess) {
                (void)hipGetLastError();
                s = cdlm_fail(ctx, CDLM_ERR_OOM, "no device memory for %zu best-window records", nslot * (size_t)N);
            }
        }
        if (s == CDLM_OK) {
            const long yb = scan_blocks(ctx, MAXJ == 32 ? reinterpret_cast<const void *>(&k_pwm_best32_mfma<32>)
                                                    : reinterpret_cast<const void *>(&k_pwm_best32_mfma<64>),
                                        lds32, h32.T, N);
            const long KN = (long)K * N;
            ScanTimer tm(ctx);
            if (MAXJ == 32)
                hipLaunchKernelGGL((k_pwm_best32_mfma<32>), dim3(h32.T, (unsigned)yb), dim3(256), lds32, st, d_bases, L, N,
                                   dB, dtl, dsk, dsm, h32.T, bval, bpos);
            else
                hipLaunchKernelGGL((k_pwm_best32_mfma<64>), dim3(h32.T, (unsigned)yb), dim3(256), lds32, st, d_bases, L, N,
                                   dB, dtl, dsk, dsm, h32.T, bval, bpos);
            hipLaunchKernelGGL(k_pwm_best_out, dim3((unsigned)((KN + 255) / 256)), dim3(256), 0, st, dso, d_thresh, K, N,
                               bval, bpos, o_pos, o_comp, o_score);
            const hipError_t e = hipGetLastError();
            tm.stop();
            if (e != hipSuccess) s = cdlm_fail(ctx, CDLM_ERR_HIP, "k_pwm_best32_mfma: %s", hipGetErrorString(e));
        }
        (void)hipStreamSynchronize(st);
        for (void *q : {(void *)dB, (void *)dtl, (void *)dsk, (void *)dsm, (void *)dso, (void *)bval, (void *)bpos})
            if (q) (void)hipFree(q);
        return s;
    }
    HostTiles ht;
    build_tiles(pwms, lens, thresh, K, maxlen, nuc ? 1 : 0, ht);
    DevTiles dt;
    int *d_slot = nullptr, *bpos = nullptr;
    float *bval = nullptr;
    cdlm_status s = CDLM_OK;
    if ((s = up(ctx, &dt.B, ht.B)) == CDLM_OK && (s = up(ctx, &dt.boff, ht.boff)) == CDLM_OK &&
        (s = up(ctx, &dt.tlen, ht.tlen)) == CDLM_OK && (s = up(ctx, &dt.sk, ht.sk)) == CDLM_OK &&
        (s = up(ctx, &dt.sm, ht.sm)) == CDLM_OK && (s = up(ctx, &dt.sst, ht.sst)) == CDLM_OK &&
        (s = up(ctx, &dt.sth, ht.sth)) == CDLM_OK && (s = up(ctx, &d_slot, ht.slot_of)) == CDLM_OK) {
        const size_t nslot = (size_t)ht.T * 16;
        if (hipMalloc(&bval, sizeof(float) * nslot * N) != hipSuccess || hipMalloc(&bpos, sizeof(int) * nslot * N) != hipSuccess) {
            (void)hipGetLastError();
            s = cdlm_fail(ctx, CDLM_ERR_OOM, "no device memory for %zu best-window records", nslot * (size_t)N);
        }
    }
    if (s == CDLM_OK) {
        PwmTiles P;
        P.T = ht.T; P.B = dt.B; P.boff = dt.boff; P.tlen = dt.tlen;
        P.slot_k = dt.sk; P.slot_m = dt.sm; P.slot_st = dt.sst; P.slot_th = dt.sth;
        P.groups = ht.T;
        const void *kfn =
            nuc ? (MAXJ == 32 ? (oh ? reinterpret_cast<const void *>(&k_pwm_best_mfma<32, kPwmCH, true, true>)
                                    : reinterpret_cast<const void *>(&k_pwm_best_mfma<32, kPwmCH, false, true>))
                              : (oh ? reinterpret_cast<const void *>(&k_pwm_best_mfma<64, kPwmCH, true, true>)
                                    : reinterpret_cast<const void *>(&k_pwm_best_mfma<64, kPwmCH, false, true>)))
                : (MAXJ == 32 ? (oh ? reinterpret_cast<const void *>(&k_pwm_best_mfma<32, kPwmCH, true, false>)
                                    : reinterpret_cast<const void *>(&k_pwm_best_mfma<32, kPwmCH, false, false>))
                              : (oh ? reinterpret_cast<const void *>(&k_pwm_best_mfma<64, kPwmCH, true, false>)
                                    : reinterpret_cast<const void *>(&k_pwm_best_mfma<64, kPwmCH, false, false>)));
        const long yb = scan_blocks(ctx, kfn, lds, P.groups, N);
        const long KN = (long)K * N;
        const unsigned ob = (unsigned)((KN + 255) / 256);
        ScanTimer tm(ctx);
#define L_BEST(MJ, OHV, NUCV)                                                                                      \
    hipLaunchKernelGGL((k_pwm_best_mfma<MJ, kPwmCH, OHV, NUCV>), dim3(P.groups, (unsigned)yb), dim3(256), lds, st, \
                       d_bases, L, N, P, bval, bpos)
        if (nuc) {
            if (MAXJ == 32 && oh) L_BEST(32, true, true);
            else if (MAXJ == 32) L_BEST(32, false, true);
            else if (oh) L_BEST(64, true, true);
            else L_BEST(64, false, true);
            hipLaunchKernelGGL(k_pwm_best_out, dim3(ob), dim3(256), 0, st, d_slot, d_thresh, K, N, bval, bpos, o_pos,
                               o_comp, o_score);
        } else {
            if (MAXJ == 32 && oh) L_BEST(32, true, false);
            else if (MAXJ == 32) L_BEST(32, false, false);
            else if (oh) L_BEST(64, true, false);
            else L_BEST(64, false, false);
            hipLaunchKernelGGL(k_pwm_argmax_out, dim3(ob), dim3(256), 0, st, d_slot, K, N, bpos, o_pos);
        }
#undef L_BEST
        const hipError_t e = hipGetLastError();
        tm.stop();
        if (e != hipSuccess) s = cdlm_fail(ctx, CDLM_ERR_HIP, "k_pwm_best_mfma: %s", hipGetErrorString(e));
    }
    (void)hipStreamSynchronize(st);
    dt.release();
    if (d_slot) (void)hipFree(d_slot);
    if (bval) (void)hipFree(bval);
    if (bpos) (void)hipFree(bpos);
    return s;
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
    if (s == CDLM_OK) s = launch_hits(ctx, d.bases, L, N, pwms, lens, thresh, K, maxlen, both_strands, hk, hn, hp, hs, cap, cnt);
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
    unsigned long long *cnt = nullptr;
    if (hipMalloc(&cnt, 8) != hipSuccess) {
        (void)hipGetLastError();
        s = cdlm_fail(ctx, CDLM_ERR_OOM, "counter");
    }
    if (s == CDLM_OK)
        s = launch_hits(ctx, d_bases, L, N, pwms, lens, thresh, K, maxlen, both_strands, d_hit_k, (long long *)d_hit_n,
                        d_hit_pos, (signed char *)d_hit_strand, cap, cnt);
    unsigned long long total = 0;
    if (s == CDLM_OK && (hipMemcpyAsync(&total, cnt, 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
                         hipStreamSynchronize(ctx->stream) != hipSuccess))
        s = cdlm_fail(ctx, CDLM_ERR_HIP, "count readback failed");
    if (cnt) (void)hipFree(cnt);
    if (s != CDLM_OK) return s;
    *n_hits = (int64_t)total;
    if ((int64_t)total > cap) return cdlm_fail(ctx, CDLM_ERR_OOM, "%llu hits exceed cap %lld", total, (long long)cap);
    return CDLM_OK;
}

// Resolves this GPU's address of the fixed scratch array once per device (with the context's
// device current, from cdlm_create) and keeps it on the context.
cdlm_status cdlm_pwm_scratch_init(cdlm_ctx *ctx) {
    if (ctx->device < 0 || ctx->device >= kMaxScratchDevices)
        return cdlm_fail(ctx, CDLM_ERR_UNSUPPORTED, "device %d beyond the scratch table", ctx->device);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[ctx->device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            (void)hipGetLastError();
            return cdlm_fail(ctx, CDLM_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        }
        g_scratch_ptr[ctx->device] = ptr;
    }
    ctx->pwm_scratch = g_scratch_ptr[ctx->device];
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_pwm_hits_device_noalloc(cdlm_ctx *ctx, const uint8_t *d_bases, int32_t L, int64_t N,
                                                    const float *pwms, const int32_t *lens, const float *thresh,
                                                    int32_t K, int32_t maxlen, int32_t both_strands, int32_t *d_hit_k,
                                                    int64_t *d_hit_n, int32_t *d_hit_pos, int8_t *d_hit_strand,
                                                    int64_t cap, int64_t *n_hits) {
    if (!ctx) return CDLM_ERR_ARG;
    cdlm_status s = check_pwm_args(ctx, d_bases, L, N, pwms, lens, thresh, K, maxlen);
    if (s != CDLM_OK) return s;
    if (!n_hits) return cdlm_fail(ctx, CDLM_ERR_ARG, "null n_hits");
    if (!ctx->pwm_scratch) return cdlm_fail(ctx, CDLM_ERR_STATE, "scratch array not resolved for device %d", ctx->device);
    HostTiles ht;
    build_tiles(pwms, lens, thresh, K, maxlen, both_strands, ht);
    const ScratchLayout lay = scratch_layout(ht.B.size(), (size_t)ht.T);
    if (lay.total > (size_t)CDLM_PWM_SCRATCH_BYTES)
        return cdlm_fail(ctx, CDLM_ERR_UNSUPPORTED, "%d PWMs of up to %d columns need %zu scratch bytes, more than the %d "
                         "of the allocation-free scan", K, maxlen, lay.total, (int)CDLM_PWM_SCRATCH_BYTES);
    hipStream_t st = ctx->stream;
    unsigned char *base = static_cast<unsigned char *>(ctx->pwm_scratch);
    unsigned long long *cnt = reinterpret_cast<unsigned long long *>(base + lay.count);
    hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(64), 0, st, cnt);
    CDLM_HIP(ctx, hipGetLastError());
    unsigned long long total = 0;
    if (N > 0) {
        auto put = [&](size_t off, const void *src, size_t bytes) {
            return bytes ? hipMemcpyAsync(base + off, src, bytes, hipMemcpyHostToDevice, st) : hipSuccess;
        };
        hipError_t e = put(lay.B, ht.B.data(), ht.B.size() * sizeof(float));
        if (e == hipSuccess) e = put(lay.boff, ht.boff.data(), ht.boff.size() * sizeof(int));
        if (e == hipSuccess) e = put(lay.tlen, ht.tlen.data(), ht.tlen.size() * sizeof(int));
        if (e == hipSuccess) e = put(lay.sk, ht.sk.data(), ht.sk.size() * sizeof(int));
        if (e == hipSuccess) e = put(lay.sm, ht.sm.data(), ht.sm.size() * sizeof(int));
        if (e == hipSuccess) e = put(lay.sst, ht.sst.data(), ht.sst.size() * sizeof(int));
        if (e == hipSuccess) e = put(lay.sth, ht.sth.data(), ht.sth.size() * sizeof(float));
        if (e != hipSuccess) return cdlm_fail(ctx, CDLM_ERR_HIP, "tile upload: %s", hipGetErrorString(e));
        PwmTiles P;
        P.T = ht.T;
        P.B = reinterpret_cast<const float *>(base + lay.B);
        P.boff = reinterpret_cast<const int *>(base + lay.boff);
        P.tlen = reinterpret_cast<const int *>(base + lay.tlen);
        P.slot_k = reinterpret_cast<const int *>(base + lay.sk);
        P.slot_m = reinterpret_cast<const int *>(base + lay.sm);
        P.slot_st = reinterpret_cast<const int *>(base + lay.sst);
        P.slot_th = reinterpret_cast<const float *>(base + lay.sth);
        s = run_hits_kernel(ctx, d_bases, L, N, maxlen, P, d_hit_k, (long long *)d_hit_n, d_hit_pos,
                            (signed char *)d_hit_strand, cap, cnt, false);
        if (s != CDLM_OK) {
            (void)hipStreamSynchronize(st);   // the host tiles must outlive any enqueued copy
            return s;
        }
    }
    CDLM_HIP(ctx, hipMemcpyAsync(&total, cnt, sizeof(total), hipMemcpyDeviceToHost, st));
    CDLM_HIP(ctx, hipStreamSynchronize(st));
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
        // on the MFMA tiles when the sequences fit the LDS, else one wave per (k, n)
        s = launch_best(ctx, d.bases, L, N, pwms, lens, thresh, d.thresh, K, maxlen, false, o, nullptr, nullptr);
        if (s == CDLM_ERR_UNSUPPORTED) {
            s = CDLM_OK;
            const long waves = (long)K * N;
            const long blocks = (waves + 3) / 4;
            ScanTimer tm(ctx);
            hipLaunchKernelGGL(k_pwm_argmax, dim3((unsigned)blocks), dim3(kPwmThreads), 0, ctx->stream, d.bases, L, N,
                               d.pwms, d.lens, d.thresh, K, maxlen, o);
            tm.stop();
            if (hipGetLastError() != hipSuccess) s = cdlm_fail(ctx, CDLM_ERR_HIP, "k_pwm_argmax failed");
        }
        if (s == CDLM_OK &&
            (hipMemcpyAsync(out_pos, o, sizeof(int) * (size_t)K * N, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
             hipStreamSynchronize(ctx->stream) != hipSuccess))
            s = cdlm_fail(ctx, CDLM_ERR_HIP, "copy of the argmax positions failed");
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
        // on the MFMA tiles when the sequences fit the LDS, else one wave per (k, n)
        s = launch_best(ctx, d.bases, L, N, pwms, lens, thresh, d.thresh, K, maxlen, true, op, oc, os);
        if (s == CDLM_ERR_UNSUPPORTED) {
            s = CDLM_OK;
            const long blocks = ((long)KN + 3) / 4;
            ScanTimer tm(ctx);
            hipLaunchKernelGGL(k_pwm_best, dim3((unsigned)blocks), dim3(kPwmThreads), 0, ctx->stream, d.bases, L, N,
                               d.pwms, d.lens, d.thresh, K, maxlen, op, oc, os);
            tm.stop();
            if (hipGetLastError() != hipSuccess) s = cdlm_fail(ctx, CDLM_ERR_HIP, "k_pwm_best failed");
        }
        hipStream_t st = ctx->stream;
        if (s == CDLM_OK && (hipMemcpyAsync(out_pos, op, KN * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
                             hipMemcpyAsync(out_strand, oc, KN, hipMemcpyDeviceToHost, st) != hipSuccess ||
                             hipMemcpyAsync(out_score, os, KN * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
                             hipStreamSynchronize(st) != hipSuccess))
            s = cdlm_fail(ctx, CDLM_ERR_HIP, "copy of the best hits failed");
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

extern "C" cdlm_status cdlm_selftest_mfma(cdlm_ctx *ctx, const float *a64, const float *b64, const float *c256,
                                          float *d256) {
    if (!ctx || !a64 || !b64 || !c256 || !d256) return CDLM_ERR_ARG;
    float *buf = nullptr;
    CDLM_HIP(ctx, hipMalloc(&buf, sizeof(float) * (64 + 64 + 256 + 256)));
    hipStream_t st = ctx->stream;
    cdlm_status s = CDLM_OK;
    if (hipMemcpyAsync(buf, a64, 256, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(buf + 64, b64, 256, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(buf + 128, c256, 1024, hipMemcpyHostToDevice, st) != hipSuccess)
        s = cdlm_fail(ctx, CDLM_ERR_HIP, "probe upload");
    if (s == CDLM_OK) {
        hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, st, buf, buf + 64, buf + 128, buf + 384);
        if (hipGetLastError() != hipSuccess || hipMemcpyAsync(d256, buf + 384, 1024, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            s = cdlm_fail(ctx, CDLM_ERR_HIP, "probe failed");
    }
    (void)hipFree(buf);
    return s;
}

extern "C" cdlm_status cdlm_pwm_last_ms(cdlm_ctx *ctx, double *ms) {
    if (!ctx || !ms) return CDLM_ERR_ARG;
    *ms = ctx->last_scan_ms;
    return CDLM_OK;
}
