/*
 * CPU oracle for the CDL-ADMM and PWM-scan hot path of CDLmotif.jl.
 *
 * TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py load this library as the checker.  The product (libcdlm) never links it.
 *
 * CDL part: oracle_impl.h, instantiated for float (_f32) and double (_f64).
 * PWM part (below): plain C restatements of
 *   greedy_search!  src/SEARCH/scan.jl:223-239  (position-major FP32 sum, threshold)
 *   scan_w_gpu!     src/SEARCH/scan.jl:176-221  (argmax over positions, 1 -> "no hit")
 *   scan_n          src/SEARCH/scan.jl:10-21    (nucleotide-major FP32 sum, first max)
 *   overlapping_scan! src/SEARCH/scan.jl:89-121 (both strands, forward wins ties)
 *   pwm_complement  src/SEARCH/helpers.jl:16    (reverse in both dimensions)
 *   best_score      src/SEARCH/PWM_Touzet.jl:9  (sum of column maxima)
 * Sequences are given as base codes 0..3 (A,C,G,T), the information content of the
 * reference's one-hot columns (src/LOAD/to_dummy.jl:2-9; TYPE_setup.jl:33-36).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define REAL float
#define SFX _f32
#include "oracle_impl.h"
#undef REAL
#undef SFX

#define REAL double
#define SFX _f64
#include "oracle_impl.h"
#undef REAL
#undef SFX

/* PWMs are packed (K, 4, maxlen) row-major: pwm[(k*4 + a)*maxlen + j] (zero padded). */

/* Position-major sequential FP32 score of window starting at base l0 (0-based),
 * complement selects the reverse-complement PWM. */
static inline float pm_score(const uint8_t *seq, const float *pwm, int maxlen, int m, int l0, int comp) {
    float acc = 0.0f;
    for (int j = 0; j < m; j++) {
        int b = seq[l0 + j];
        if (b > 3) { acc = acc + 0.0f; continue; }    /* all-zero one-hot column adds +0 */
        float w = comp ? pwm[(3 - b) * maxlen + (m - 1 - j)] : pwm[b * maxlen + j];
        acc = acc + w;
    }
    return acc;
}

/* Hit list: every window with score > thresh[k], both strands when both_strands != 0.
 * Output records (k, n, pos0, strand) in (k, n, pos, strand) order.  Returns the number of
 * hits; writes at most cap records. */
long orc_pwm_hits(const uint8_t *seqs, int L, long N, const float *pwms, const int *lens, const float *thresh,
                  int K, int maxlen, int both_strands, int32_t *out_k, int64_t *out_n, int32_t *out_pos,
                  int8_t *out_strand, float *out_score, long cap) {
    long cnt = 0;
    for (int k = 0; k < K; k++) {
        const float *pw = pwms + (long)k * 4 * maxlen;
        int m = lens[k];
        for (long n = 0; n < N; n++) {
            const uint8_t *seq = seqs + n * (long)L;
            for (int l0 = 0; l0 + m <= L; l0++)
                for (int st = 0; st < (both_strands ? 2 : 1); st++) {
                    float s = pm_score(seq, pw, maxlen, m, l0, st);
                    if (s > thresh[k]) {
                        if (cnt < cap) {
                            out_k[cnt] = k; out_n[cnt] = n; out_pos[cnt] = l0; out_strand[cnt] = (int8_t)st;
                            if (out_score) out_score[cnt] = s;
                        }
                        cnt++;
                    }
                }
        }
    }
    return cnt;
}

/* scan_w_gpu!: pos_scores[k,n,l] = (s > thresh ? s : 0) for l <= L-m+1, else 0;
 * argmax over l (first max); index 1 (1-based) reported as 0 = no hit.
 * out_pos[(k*N + n)] is the 1-based position or 0. */
void orc_pwm_gpu_argmax(const uint8_t *seqs, int L, long N, const float *pwms, const int *lens, const float *thresh,
                        int K, int maxlen, int32_t *out_pos) {
    for (int k = 0; k < K; k++) {
        const float *pw = pwms + (long)k * 4 * maxlen;
        int m = lens[k];
        for (long n = 0; n < N; n++) {
            const uint8_t *seq = seqs + n * (long)L;
            float best = 0.0f; int bi = 0;
            for (int l0 = 0; l0 < L; l0++) {
                float v = 0.0f;
                if (l0 + m <= L) {
                    float s = pm_score(seq, pw, maxlen, m, l0, 0);
                    v = (s > thresh[k]) ? s : 0.0f;
                }
                if (l0 == 0) { best = v; bi = 0; }
                else if (v > best) { best = v; bi = l0; }
            }
            out_pos[(long)k * N + n] = (bi == 0) ? 0 : bi + 1;
        }
    }
}

/* scan_n with nucleotide-major order; returns the 1-based start or -1, score via *out. */
static int scan_n_nm(const uint8_t *seq, int L, const float *pw, int maxlen, int m, int comp, float *out) {
    int bi = -1; float best = -INFINITY;
    for (int i = 0; i + m <= L; i++) {
        float score = 0.0f;
        for (int a = 0; a < 4; a++)
            for (int j = 0; j < m; j++) {
                if (seq[i + j] != a) continue;          /* product with a 0 entry adds a signed zero */
                float w = comp ? pw[(3 - a) * maxlen + (m - 1 - j)] : pw[a * maxlen + j];
                score += w;
            }
        if (score > best) { best = score; bi = i + 1; }
    }
    *out = best;
    return bi;
}

/* overlapping_scan!: per (k,n) best forward and best reverse-complement windows; forward
 * wins ties (score >= score_c); a hit needs score > thresh.  out_pos 1-based or 0,
 * out_comp 0/1, out_score. */
void orc_pwm_best(const uint8_t *seqs, int L, long N, const float *pwms, const int *lens, const float *thresh, int K,
                  int maxlen, int32_t *out_pos, int8_t *out_comp, float *out_score) {
    for (int k = 0; k < K; k++) {
        const float *pw = pwms + (long)k * 4 * maxlen;
        int m = lens[k];
        for (long n = 0; n < N; n++) {
            const uint8_t *seq = seqs + n * (long)L;
            float s, sc;
            int p = scan_n_nm(seq, L, pw, maxlen, m, 0, &s);
            int pc = scan_n_nm(seq, L, pw, maxlen, m, 1, &sc);
            long o = (long)k * N + n;
            out_pos[o] = 0; out_comp[o] = 0; out_score[o] = 0.0f;
            if (s >= sc) {
                if (s > thresh[k]) { out_pos[o] = p; out_comp[o] = 0; out_score[o] = s; }
            } else {
                if (sc > thresh[k]) { out_pos[o] = pc; out_comp[o] = 1; out_score[o] = sc; }
            }
        }
    }
}

/* best_score: sum over columns (sequential, FP32) of the column maximum. */
float orc_best_score(const float *pwm, int maxlen, int m) {
    float acc = 0.0f;
    for (int j = 0; j < m; j++) {
        float mx = pwm[j];
        for (int a = 1; a < 4; a++) if (pwm[a * maxlen + j] > mx) mx = pwm[a * maxlen + j];
        acc += mx;
    }
    return acc;
}

/* ------------------------------------------------------------------------------------------
 * topology_at_nth_pt! (src/SEARCH/store.jl:64-107) over all sequences, index for index.
 * start_pos (C, K, N) bytes column-major (0/1); train_num (K) 1-based; fil_lens (K);
 * cover_starts (N, T) column-major, 1-based; counts are Int32 (activations_topology{int_t},
 * SEARCH_setup.jl:73-90; int_t = Int32, TYPE_setup.jl:4): over (K, K, maxfl),
 * non_over (K, K, apart), next_to (K, K), column-major, accumulated (not cleared).
 * Every (r, c) in 1..C*K x 1..C*K is visited as the reference's threads are, including the
 * else branch that also takes distances beyond apart (store.jl:98-100).
 */
void orc_topology(const uint8_t *start_pos, const int *train_num, const int *fil_lens, const int *cover_starts, int C,
                  int K, long N, int T, int maxfl, int apart, int *over, int *non_over, int *next_to) {
    (void)T;
    const long CK = (long)C * K;
    for (long n = 1; n <= N; ++n)
        for (long r = 1; r <= CK; ++r)
            for (long c = 1; c <= CK; ++c) {
                const long pr_s_cover = (r - 1) % C + 1;
                const long fr = ((r - 1) - (pr_s_cover - 1)) / C + 1;
                const int tr = train_num[fr - 1];
                const long pr_s = pr_s_cover + cover_starts[(n - 1) + N * (tr - 1)] - 1;
                const long pr_e = pr_s + fil_lens[fr - 1] - 1;
                const long pc_s_cover = (c - 1) % C + 1;
                const long fc = ((c - 1) - (pc_s_cover - 1)) / C + 1;
                const int tc = train_num[fc - 1];
                const long pc_s = pc_s_cover + cover_starts[(n - 1) + N * (tc - 1)] - 1;
                const uint8_t ar = start_pos[(pr_s_cover - 1) + C * ((fr - 1) + K * (n - 1))];
                const uint8_t ac = start_pos[(pc_s_cover - 1) + C * ((fc - 1) + K * (n - 1))];
                if (fr != fc && pr_s <= pc_s && ar && ac && tr != tc) {
                    long d = pc_s - pr_e - 1;
                    if (!(fil_lens[fr - 1] == fil_lens[fc - 1] && pr_s == pc_s)) {
                        if (d < 0) {
                            d = -d;
                            if (d <= maxfl) over[(fr - 1) + K * ((fc - 1) + (long)K * (d - 1))] += 1;
                        } else if (d > 0 && d <= apart) {
                            non_over[(fr - 1) + K * ((fc - 1) + (long)K * (d - 1))] += 1;
                        } else {
                            next_to[(fr - 1) + K * (fc - 1)] += 1;
                        }
                    }
                }
            }
}

/* ------------------------------------------------------------------------------------------
 * get_pair_pfms! up to the PFMs (src/SEARCH/extractions.jl:18-100, kernels :115-190):
 * get_f_pair_starts, get_span_len, msa_fill (weights = mean of the two code values, c
 * ascending, Float32), the sum over sequences (here in ascending n, in Float32), the 0.1
 * pseudocount on the span (msa_pseudocount) and the division by the first column's total.
 * Inputs: start_pos (C, K, N) bytes; codes (C, M, N, T) Float32 (the reference's ar.codes with
 * the training index last); uinds (K) 1-based; train_num (K) 1-based; fil_lens (K);
 * data (4C, N) one-hot Float32; pairs (P, 3) row-major 1-based (fil1, fil2, overlap).
 * Outputs: span_len (P); pfm (P, 4, S) row-major with S = max_span_len.
 */
void orc_pair_pfms(const uint8_t *start_pos, const float *codes, const int *uinds, const int *train_num,
                   const int *fil_lens, const float *data, int C, int K, int M, long N, int P, const int *pairs, int S,
                   int *span_len, float *pfm) {
    float *msa = (float *)calloc((size_t)4 * S, sizeof(float));
    uint8_t *fps = (uint8_t *)malloc((size_t)C);
    for (int i = 0; i < P; ++i) {
        const int f1 = pairs[3 * i], f2 = pairs[3 * i + 1], olen = pairs[3 * i + 2];
        const int l1 = fil_lens[f1 - 1], l2 = fil_lens[f2 - 1];
        const int span = olen >= l2 ? olen : l1 - olen + l2;                       /* get_span_len */
        span_len[i] = span;
        memset(msa, 0, sizeof(float) * 4 * S);
        const int t1 = train_num[f1 - 1], t2 = train_num[f2 - 1];
        const int u1 = uinds[f1 - 1], u2 = uinds[f2 - 1];
        for (long n = 0; n < N; ++n) {
            memset(fps, 0, (size_t)C);
            for (int c = 1; c <= C; ++c) {                                         /* get_f_pair_starts */
                const int at = olen + c - l1;
                if (start_pos[(c - 1) + (size_t)C * ((f2 - 1) + (size_t)K * n)] && at > 0 &&
                    start_pos[(at - 1) + (size_t)C * ((f1 - 1) + (size_t)K * n)])
                    fps[at - 1] = 1;
            }
            for (int l = 1; l <= 4 * S; ++l) {                                     /* msa_fill */
                float acc = 0.f;
                for (int c = 1; c <= C - span + 1; ++c)
                    if (fps[c - 1] && l <= 4 * span) {
                        const int s2 = c + l1 - olen;
                        const float w = (codes[(c - 1) + (size_t)C * ((u1 - 1) + (size_t)M * (n + (size_t)N * (t1 - 1)))] +
                                         codes[(s2 - 1) + (size_t)C * ((u2 - 1) + (size_t)M * (n + (size_t)N * (t2 - 1)))]) /
                                        2.0f;
                        acc += w * data[(size_t)(c - 1) * 4 + (l - 1) + (size_t)4 * C * n];
                    }
                msa[l - 1] += acc;                                                 /* sum over n */
            }
        }
        for (int l = 1; l <= 4 * span && l <= 4 * S; ++l) msa[l - 1] += 0.1f;       /* msa_pseudocount */
        const float tot = msa[0] + msa[1] + msa[2] + msa[3];                       /* first column */
        for (int a = 0; a < 4; ++a)
            for (int s = 0; s < S; ++s) pfm[((size_t)i * 4 + a) * S + s] = msa[a + 4 * s] / tot;
    }
    free(msa);
    free(fps);
}
