// cdlm_record.hip -- activation record and activation topology for gfx950.
//
// Reference (CDLmotif.jl): activation_record / activations_topology
// (src/SETUP/SEARCH_setup.jl:6-90), record_training_result! (src/SEARCH/store.jl:1-40),
// topology_at_nth_pt! and get_spatial_info! (store.jl:64-144).
//
// The record keeps each training's X_best slice (L, M, N) on the device (the reference's
// ar.codes) with its code threshold, the list of filters with any activation (uinds, their
// training number and length) and the cover starts.  A position p of filter k in sequence n
// is active when codes[t_k][p, uind_k, n] > thresh_t (Float32 code against the threshold in
// Float64, as Julia promotes the comparison).
//
// Topology: the reference launches one thread per (r, c, n) in C*K x C*K x N and counts,
// with atomics, every ordered pair of active (position, filter) entries of a sequence that
// come from different trainings (store.jl:64-107).  Only active entries can count, so here a
// work-group per sequence first compacts that sequence's active entries into a list and then
// visits the ordered pairs of the list: the same pairs and the same Int32 increments (integer
// atomics, so the totals do not depend on the visiting order) at a cost of A^2 instead of
// (C K)^2 per sequence for A active entries.

#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "cdlm_internal.h"

namespace {

constexpr int kTopoThreads = 256;

struct TopoArgs {
    int C, K, T, M, maxfl, apart;
    long N;
    const int *train;      // (K) 1-based training of each filter
    const int *flen;       // (K) filter length
    const int *uind;       // (K) 0-based filter index within its training (codes mode)
    const int *cs;         // (N, T) cover starts, 1-based
    const uint8_t *sp;     // (C, K, N) activity bytes, or null: from codes
    const float *codes;    // (C, M, N, T) when sp is null
    const double *thr;     // (T)
    int *over, *non, *nxt; // (K, K, maxfl), (K, K, apart), (K, K)
    int *scratch;          // C*K ints per work-group
};

__device__ __forceinline__ bool active_of(const TopoArgs &A, long e, long n) {
    if (A.sp) return A.sp[e + (long)A.C * A.K * n] != 0;
    const int p = (int)(e % A.C), k = (int)(e / A.C);
    const int t = A.train[k] - 1;
    const float x = A.codes[(size_t)p + (size_t)A.C * ((size_t)A.uind[k] + (size_t)A.M * ((size_t)n + (size_t)A.N * t))];
    return (double)x > A.thr[t];
}

// One work-group per sequence (grid-stride): compaction of the active entries, then all
// ordered pairs (store.jl:77-102, index for index).
__global__ __launch_bounds__(kTopoThreads) void k_topology(TopoArgs A) {
    __shared__ int cnt;
    const long CK = (long)A.C * A.K;
    int *lst = A.scratch + (size_t)blockIdx.x * CK;
    for (long n = blockIdx.x; n < A.N; n += gridDim.x) {
        if (threadIdx.x == 0) cnt = 0;
        __syncthreads();
        for (long e = threadIdx.x; e < CK; e += kTopoThreads)
            if (active_of(A, e, n)) lst[atomicAdd(&cnt, 1)] = (int)e;
        __syncthreads();
        const int na = cnt;
        const long np = (long)na * na;
        for (long idx = threadIdx.x; idx < np; idx += kTopoThreads) {
            const int er = lst[idx / na], ec = lst[idx % na];
            const int fr = er / A.C, fc = ec / A.C;
            if (fr == fc) continue;
            const int tr = A.train[fr], tc = A.train[fc];
            if (tr == tc) continue;
            // 1-based starts: (p + 1) + cover_start - 1
            const long rs = (er % A.C) + (long)A.cs[n + A.N * (tr - 1)];
            const long cst = (ec % A.C) + (long)A.cs[n + A.N * (tc - 1)];
            if (rs > cst) continue;
            const int lr = A.flen[fr];
            if (lr == A.flen[fc] && rs == cst) continue;          // exact matches are excluded
            const long d = cst - (rs + lr - 1) - 1;
            if (d < 0) {
                if (-d <= A.maxfl) atomicAdd(&A.over[fr + (size_t)A.K * (fc + (size_t)A.K * (-d - 1))], 1);
            } else if (d > 0 && d <= A.apart) {
                atomicAdd(&A.non[fr + (size_t)A.K * (fc + (size_t)A.K * (d - 1))], 1);
            } else {
                atomicAdd(&A.nxt[fr + (size_t)A.K * fc], 1);           // d == 0, and d beyond apart
            }
        }
        __syncthreads();
    }
}

// flags[m] = 1 when codes[p, m, n] > thr for some p, n.
__global__ void k_any_active(const float *__restrict__ codes, int C, int M, long N, double thr, int *flags) {
    const long tot = (long)C * M * N;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += (long)gridDim.x * blockDim.x) {
        if ((double)codes[i] > thr) {
            const int m = (int)((i / C) % M);
            if (flags[m] == 0) atomicOr(&flags[m], 1);
        }
    }
}

// Activity bytes (C, K, cnt) of sequences n0 .. n0+cnt-1.
__global__ void k_start_positions(TopoArgs A, long n0, long cnt, uint8_t *out) {
    const long CK = (long)A.C * A.K;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < CK * cnt; i += (long)gridDim.x * blockDim.x)
        out[i] = active_of(A, i % CK, n0 + i / CK) ? 1 : 0;
}

// get_pair_pfms! counts (extractions.jl:115-190): one wave per (pair, chunk of sequences).
// Per sequence, lanes evaluate the pair-start flag of 64 positions at once (get_f_pair_starts:
// f1 active at a and f2 active at a + len1 - overlap; positions a <= C - span, msa_fill's
// range), then every lane l < 4 span adds weight x data[4a + l] for the flagged a in ascending
// order (msa_fill), weight = (code1 + code2) / 2.  The per-sequence rows are summed over the
// chunk's sequences in ascending order into partial[pair][chunk][l].
struct PairArgs {
    int C, M, S, P, nchunk;
    long N;
    const int *pairs;      // (P, 3) column-major, 1-based (fil1, fil2, overlap)
    const int *uind, *train, *flen;
    const float *codes;    // (C, M, N, T)
    const double *thr;
    const float *data;     // (4C, N)
    float *partial;        // (P, nchunk, 4S)
};

__global__ __launch_bounds__(256) void k_pair_msa(PairArgs A) {
    const int lane = threadIdx.x & 63;
    const long wv = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wv >= (long)A.P * A.nchunk) return;
    const int i = (int)(wv / A.nchunk), ch = (int)(wv % A.nchunk);
    const int f1 = A.pairs[i] - 1, f2 = A.pairs[i + A.P] - 1, olen = A.pairs[i + 2 * A.P];
    const int l1 = A.flen[f1], l2 = A.flen[f2];
    const int span = olen >= l2 ? olen : l1 - olen + l2;      // get_span_len
    const int t1 = A.train[f1] - 1, t2 = A.train[f2] - 1;
    const double th1 = A.thr[t1], th2 = A.thr[t2];
    const long n0 = (long)ch * A.N / A.nchunk, n1 = (long)(ch + 1) * A.N / A.nchunk;
    const int last = A.C - span;                                // a in [0, C - span]
    float sum = 0.f;
    for (long n = n0; n < n1; ++n) {
        const float *c1 = A.codes + (size_t)A.C * (A.uind[f1] + (size_t)A.M * ((size_t)n + (size_t)A.N * t1));
        const float *c2 = A.codes + (size_t)A.C * (A.uind[f2] + (size_t)A.M * ((size_t)n + (size_t)A.N * t2));
        const float *dn = A.data + (size_t)4 * A.C * n;
        float acc = 0.f;
        for (int base = 0; base <= last; base += 64) {
            const int a = base + lane;
            bool f = false;
            float w = 0.f;
            if (a <= last) {
                const int b = a + l1 - olen;
                if (b < A.C) {
                    const float x1 = c1[a], x2 = c2[b];
                    f = (double)x1 > th1 && (double)x2 > th2;
                    w = (x1 + x2) / 2.0f;
                }
            }
            unsigned long long mask = __ballot(f);
            while (mask) {
                const int j = __ffsll((long long)mask) - 1;
                mask &= mask - 1;
                const float wj = __shfl(w, j);
                if (lane < 4 * span) acc += wj * dn[4 * (base + j) + lane];
            }
        }
        sum += acc;
    }
    if (lane < 4 * A.S) A.partial[((size_t)i * A.nchunk + ch) * 4 * A.S + lane] = sum;
}

struct DevMem {
    std::vector<void *> ptrs;
    cdlm_ctx *ctx;
    explicit DevMem(cdlm_ctx *c) : ctx(c) {}
    ~DevMem() {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <typename T>
    cdlm_status alloc(T **out, size_t count) {
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, count ? count * sizeof(T) : 4);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return cdlm_fail(ctx, CDLM_ERR_OOM, "hipMalloc(%zu bytes) failed", count * sizeof(T));
        }
        ptrs.push_back(q);
        *out = static_cast<T *>(q);
        return CDLM_OK;
    }
};

template <typename T>
cdlm_status upload(cdlm_ctx *ctx, DevMem &m, T **dst, const T *src, size_t count) {
    cdlm_status st = m.alloc(dst, count);
    if (st != CDLM_OK) return st;
    if (count) CDLM_HIP(ctx, hipMemcpyAsync(*dst, src, count * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    return CDLM_OK;
}

// Runs k_topology for the prepared arguments (device pointers of the inputs set) and copies
// the counts to the host outputs.
cdlm_status run_topology(cdlm_ctx *ctx, DevMem &m, TopoArgs A, int32_t *over, int32_t *non_over, int32_t *next_to) {
    const size_t K2 = (size_t)A.K * A.K;
    const size_t no = K2 * (size_t)A.maxfl, nn = K2 * (size_t)(A.apart > 0 ? A.apart : 0);
    cdlm_status st;
    if ((st = m.alloc(&A.over, no)) != CDLM_OK || (st = m.alloc(&A.non, nn)) != CDLM_OK ||
        (st = m.alloc(&A.nxt, K2)) != CDLM_OK)
        return st;
    long grid = (long)(ctx->num_cu > 0 ? ctx->num_cu : 256) * 4;
    if (grid > A.N) grid = A.N;
    if (grid < 1) grid = 1;
    if ((st = m.alloc(&A.scratch, (size_t)grid * A.C * A.K)) != CDLM_OK) return st;
    hipStream_t s = ctx->stream;
    CDLM_HIP(ctx, hipMemsetAsync(A.over, 0, no * sizeof(int), s));
    if (nn) CDLM_HIP(ctx, hipMemsetAsync(A.non, 0, nn * sizeof(int), s));
    CDLM_HIP(ctx, hipMemsetAsync(A.nxt, 0, K2 * sizeof(int), s));
    if (A.N > 0 && A.K > 0) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
            (void)hipGetLastError();
            e0 = e1 = nullptr;
        }
        if (e0) (void)hipEventRecord(e0, s);
        hipLaunchKernelGGL(k_topology, dim3((unsigned)grid), dim3(kTopoThreads), 0, s, A);
        hipError_t e = hipGetLastError();
        if (e1) (void)hipEventRecord(e1, s);
        if (e != hipSuccess) return cdlm_fail(ctx, CDLM_ERR_HIP, "k_topology: %s", hipGetErrorString(e));
        CDLM_HIP(ctx, hipStreamSynchronize(s));
        float ms = 0.f;
        if (e0 && e1 && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) ctx->last_topology_ms = ms;
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
    if (over && no) CDLM_HIP(ctx, hipMemcpyAsync(over, A.over, no * sizeof(int), hipMemcpyDeviceToHost, s));
    if (non_over && nn) CDLM_HIP(ctx, hipMemcpyAsync(non_over, A.non, nn * sizeof(int), hipMemcpyDeviceToHost, s));
    if (next_to) CDLM_HIP(ctx, hipMemcpyAsync(next_to, A.nxt, K2 * sizeof(int), hipMemcpyDeviceToHost, s));
    CDLM_HIP(ctx, hipStreamSynchronize(s));
    return CDLM_OK;
}

}  // namespace

struct cdlm_record {
    cdlm_ctx *ctx = nullptr;
    int N = 0, C = 0, M = 0, Tmax = 0, T = 0;
    float *codes = nullptr;               // (C, M, N, Tmax)
    std::vector<double> thresh;           // (T)
    std::vector<int> uinds, train, lens;  // (K): 0-based index, 1-based training, length
    std::vector<int> cover;               // (N, Tmax), 1-based
};

extern "C" cdlm_status cdlm_record_create(cdlm_ctx *ctx, int32_t N, int32_t L, int32_t M, int32_t max_trainings,
                                          cdlm_record **out) {
    if (!ctx) return CDLM_ERR_ARG;
    if (!out || N <= 0 || L <= 0 || M <= 0 || max_trainings <= 0)
        return cdlm_fail(ctx, CDLM_ERR_ARG, "bad activation record dimensions");
    *out = nullptr;
    std::unique_ptr<cdlm_record> r(new cdlm_record());
    r->ctx = ctx;
    r->N = N; r->C = L; r->M = M; r->Tmax = max_trainings;
    const size_t bytes = (size_t)L * M * N * max_trainings * sizeof(float);
    hipError_t e = hipMalloc(&r->codes, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return cdlm_fail(ctx, CDLM_ERR_OOM, "hipMalloc(%zu bytes) for the activation record failed", bytes);
    }
    r->cover.assign((size_t)N * max_trainings, 1);
    *out = r.release();
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_record_destroy(cdlm_record *r) {
    if (!r) return CDLM_OK;
    (void)hipStreamSynchronize(r->ctx->stream);
    (void)hipFree(r->codes);
    delete r;
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_record_add(cdlm_record *r, cdlm_problem *p, double thresh, const int32_t *filter_spec,
                                       const int32_t *cover_starts, int32_t *n_added) {
    if (!r || !p || !filter_spec) return CDLM_ERR_ARG;
    cdlm_ctx *ctx = r->ctx;
    if (!(thresh > 0.0)) return cdlm_fail(ctx, CDLM_ERR_ARG, "threshold must be greater than 0");   // store.jl:8
    if (r->T >= r->Tmax) return cdlm_fail(ctx, CDLM_ERR_STATE, "activation record holds %d trainings already", r->Tmax);
    const float *slice = nullptr;
    int L = 0, M = 0, N = 0;
    cdlm_status st = cdlm_problem_xbest_device(p, &slice, &L, &M, &N);
    if (st != CDLM_OK) return st;
    if (L != r->C || M != r->M || N != r->N)
        return cdlm_fail(ctx, CDLM_ERR_ARG, "problem (L=%d, M=%d, N=%d) does not match the record (C=%d, M=%d, N=%d)", L, M,
                         N, r->C, r->M, r->N);
    const size_t cnt = (size_t)L * M * N;
    float *dst = r->codes + cnt * r->T;
    hipStream_t s = ctx->stream;
    // the problem's work is on the same context stream
    CDLM_HIP(ctx, hipMemcpyAsync(dst, slice, cnt * sizeof(float), hipMemcpyDeviceToDevice, s));
    int *flags = nullptr;
    CDLM_HIP(ctx, hipMalloc(&flags, M * sizeof(int)));
    std::vector<int> hf(M, 0);
    hipError_t e = hipMemsetAsync(flags, 0, M * sizeof(int), s);
    if (e == hipSuccess) {
        long grid = ((long)cnt + 255) / 256;
        if (grid > 65536) grid = 65536;
        hipLaunchKernelGGL(k_any_active, dim3((unsigned)grid), dim3(256), 0, s, dst, L, M, (long)N, thresh, flags);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(hf.data(), flags, M * sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(flags);
    if (e != hipSuccess) return cdlm_fail(ctx, CDLM_ERR_HIP, "record_add: %s", hipGetErrorString(e));
    int added = 0;
    for (int m = 0; m < M; ++m)
        if (hf[m]) {   // uinds = sort(unique(filter index of active entries)), store.jl:19
            r->uinds.push_back(m);
            r->train.push_back(r->T + 1);
            r->lens.push_back(filter_spec[m]);
            ++added;
        }
    for (int n = 0; n < N; ++n) r->cover[(size_t)n + (size_t)N * r->T] = cover_starts ? cover_starts[n] : 1;
    r->thresh.push_back(thresh);
    r->T += 1;
    if (n_added) *n_added = added;
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_record_dims(const cdlm_record *r, int32_t *K, int32_t *T, int32_t *C, int32_t *N) {
    if (!r) return CDLM_ERR_ARG;
    if (K) *K = (int32_t)r->uinds.size();
    if (T) *T = r->T;
    if (C) *C = r->C;
    if (N) *N = r->N;
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_record_filters(const cdlm_record *r, int32_t *uinds, int32_t *train_num,
                                           int32_t *filter_lengths) {
    if (!r) return CDLM_ERR_ARG;
    for (size_t k = 0; k < r->uinds.size(); ++k) {
        if (uinds) uinds[k] = r->uinds[k] + 1;
        if (train_num) train_num[k] = r->train[k];
        if (filter_lengths) filter_lengths[k] = r->lens[k];
    }
    return CDLM_OK;
}

namespace {
cdlm_status record_args(cdlm_record *r, DevMem &m, TopoArgs *A) {
    cdlm_ctx *ctx = r->ctx;
    TopoArgs a;
    memset(&a, 0, sizeof(a));
    a.C = r->C; a.K = (int)r->uinds.size(); a.T = r->T; a.M = r->M; a.N = r->N;
    a.codes = r->codes;
    int maxfl = 0, minfl = 1 << 30;
    for (int l : r->lens) { maxfl = std::max(maxfl, l); minfl = std::min(minfl, l); }
    a.maxfl = maxfl;
    a.apart = a.K ? r->C - minfl - 1 : 0;   // maximum_apart_length, helpers.jl:11
    cdlm_status st;
    if ((st = upload(ctx, m, const_cast<int **>(&a.train), r->train.data(), r->train.size())) != CDLM_OK) return st;
    if ((st = upload(ctx, m, const_cast<int **>(&a.flen), r->lens.data(), r->lens.size())) != CDLM_OK) return st;
    if ((st = upload(ctx, m, const_cast<int **>(&a.uind), r->uinds.data(), r->uinds.size())) != CDLM_OK) return st;
    if ((st = upload(ctx, m, const_cast<int **>(&a.cs), r->cover.data(), r->cover.size())) != CDLM_OK) return st;
    if ((st = upload(ctx, m, const_cast<double **>(&a.thr), r->thresh.data(), r->thresh.size())) != CDLM_OK) return st;
    *A = a;
    return CDLM_OK;
}
}  // namespace

extern "C" cdlm_status cdlm_record_topology_dims(const cdlm_record *r, int32_t *K, int32_t *max_fil_len,
                                                 int32_t *apart_len) {
    if (!r) return CDLM_ERR_ARG;
    int maxfl = 0, minfl = 1 << 30;
    for (int l : r->lens) { maxfl = std::max(maxfl, l); minfl = std::min(minfl, l); }
    if (K) *K = (int32_t)r->uinds.size();
    if (max_fil_len) *max_fil_len = maxfl;
    if (apart_len) *apart_len = r->uinds.empty() ? 0 : r->C - minfl - 1;
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_record_start_positions(cdlm_record *r, int32_t n0, int32_t count, uint8_t *out) {
    if (!r || !out || n0 < 0 || count < 0 || n0 + (long)count > r->N) return CDLM_ERR_ARG;
    cdlm_ctx *ctx = r->ctx;
    DevMem m(ctx);
    TopoArgs A;
    cdlm_status st = record_args(r, m, &A);
    if (st != CDLM_OK) return st;
    const size_t bytes = (size_t)r->C * A.K * count;
    if (!bytes) return CDLM_OK;
    uint8_t *d = nullptr;
    if ((st = m.alloc(&d, bytes)) != CDLM_OK) return st;
    long grid = ((long)bytes + 255) / 256;
    if (grid > 65536) grid = 65536;
    hipLaunchKernelGGL(k_start_positions, dim3((unsigned)grid), dim3(256), 0, ctx->stream, A, (long)n0, (long)count, d);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return cdlm_fail(ctx, CDLM_ERR_HIP, "k_start_positions: %s", hipGetErrorString(e));
    CDLM_HIP(ctx, hipMemcpyAsync(out, d, bytes, hipMemcpyDeviceToHost, ctx->stream));
    CDLM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_record_topology(cdlm_record *r, int32_t *overlap, int32_t *non_overlap, int32_t *next_to) {
    if (!r) return CDLM_ERR_ARG;
    DevMem m(r->ctx);
    TopoArgs A;
    cdlm_status st = record_args(r, m, &A);
    if (st != CDLM_OK) return st;
    return run_topology(r->ctx, m, A, overlap, non_overlap, next_to);
}

extern "C" cdlm_status cdlm_topology(cdlm_ctx *ctx, const uint8_t *start_pos, const int32_t *train_num,
                                     const int32_t *fil_lens, const int32_t *cover_starts, int32_t C, int32_t K,
                                     int64_t N, int32_t T, int32_t max_fil_len, int32_t apart_len, int32_t *overlap,
                                     int32_t *non_overlap, int32_t *next_to) {
    if (!ctx) return CDLM_ERR_ARG;
    if (C <= 0 || K < 0 || N < 0 || T <= 0 || max_fil_len < 0 || apart_len < 0 || (K > 0 && N > 0 && (!start_pos ||
        !train_num || !fil_lens || !cover_starts)))
        return cdlm_fail(ctx, CDLM_ERR_ARG, "bad topology arguments");
    if ((long)C * K > (1L << 30)) return cdlm_fail(ctx, CDLM_ERR_UNSUPPORTED, "C*K too large");
    for (int k = 0; k < K; ++k) {
        if (train_num[k] < 1 || train_num[k] > T) return cdlm_fail(ctx, CDLM_ERR_ARG, "training number out of range");
        if (fil_lens[k] > max_fil_len) return cdlm_fail(ctx, CDLM_ERR_ARG, "filter length beyond max_fil_len");
    }
    DevMem m(ctx);
    TopoArgs A;
    memset(&A, 0, sizeof(A));
    A.C = C; A.K = K; A.T = T; A.N = N; A.maxfl = max_fil_len; A.apart = apart_len;
    cdlm_status st;
    if ((st = upload(ctx, m, const_cast<int **>(&A.train), train_num, (size_t)K)) != CDLM_OK) return st;
    if ((st = upload(ctx, m, const_cast<int **>(&A.flen), fil_lens, (size_t)K)) != CDLM_OK) return st;
    if ((st = upload(ctx, m, const_cast<int **>(&A.cs), cover_starts, (size_t)N * T)) != CDLM_OK) return st;
    if ((st = upload(ctx, m, const_cast<uint8_t **>(&A.sp), start_pos, (size_t)C * K * N)) != CDLM_OK) return st;
    return run_topology(ctx, m, A, overlap, non_overlap, next_to);
}

extern "C" cdlm_status cdlm_topology_last_ms(cdlm_ctx *ctx, double *ms) {
    if (!ctx || !ms) return CDLM_ERR_ARG;
    *ms = ctx->last_topology_ms;
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_record_pair_pfms(cdlm_record *r, const float *data, int32_t P, const int32_t *pairs,
                                             int32_t max_span_len, int32_t *span_len, float *pfms) {
    if (!r) return CDLM_ERR_ARG;
    cdlm_ctx *ctx = r->ctx;
    const int K = (int)r->uinds.size();
    if (P < 0 || (P > 0 && (!data || !pairs || !span_len || !pfms)) || max_span_len < 1 || 4 * max_span_len > 64)
        return cdlm_fail(ctx, CDLM_ERR_ARG, "bad pair-PFM arguments (4 x max_span_len must be <= 64)");
    if (P == 0) return CDLM_OK;
    std::vector<int> span(P);
    for (int i = 0; i < P; ++i) {
        const int f1 = pairs[i], f2 = pairs[i + P], olen = pairs[i + 2 * P];
        if (f1 < 1 || f1 > K || f2 < 1 || f2 > K || olen < 1 || olen > r->lens[f1 - 1])
            return cdlm_fail(ctx, CDLM_ERR_ARG, "pair %d out of range", i);
        const int l1 = r->lens[f1 - 1], l2 = r->lens[f2 - 1];
        span[i] = olen >= l2 ? olen : l1 - olen + l2;
        if (span[i] > max_span_len) return cdlm_fail(ctx, CDLM_ERR_ARG, "span %d of pair %d exceeds max_span_len", span[i], i);
        span_len[i] = span[i];
    }
    DevMem m(ctx);
    TopoArgs T;
    cdlm_status st = record_args(r, m, &T);
    if (st != CDLM_OK) return st;
    PairArgs A;
    memset(&A, 0, sizeof(A));
    A.C = r->C; A.M = r->M; A.S = max_span_len; A.P = P; A.N = r->N;
    A.uind = T.uind; A.train = T.train; A.flen = T.flen; A.thr = T.thr; A.codes = r->codes;
    long target = (long)(ctx->num_cu > 0 ? ctx->num_cu : 256) * 16;   // waves to fill the device
    long nch = (target + P - 1) / P;
    if (nch > r->N) nch = r->N;
    if (nch < 1) nch = 1;
    A.nchunk = (int)nch;
    if ((st = upload(ctx, m, const_cast<int **>(&A.pairs), pairs, (size_t)P * 3)) != CDLM_OK) return st;
    if ((st = upload(ctx, m, const_cast<float **>(&A.data), data, (size_t)4 * r->C * r->N)) != CDLM_OK) return st;
    const size_t npart = (size_t)P * nch * 4 * max_span_len;
    if ((st = m.alloc(&A.partial, npart)) != CDLM_OK) return st;
    const long waves = (long)P * nch;
    hipLaunchKernelGGL(k_pair_msa, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, ctx->stream, A);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return cdlm_fail(ctx, CDLM_ERR_HIP, "k_pair_msa: %s", hipGetErrorString(e));
    std::vector<float> part(npart);
    CDLM_HIP(ctx, hipMemcpyAsync(part.data(), A.partial, npart * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    CDLM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    // sum over the chunks in order, pseudocount on the span (msa_pseudocount), divide by the
    // first column's total (extractions.jl:70-77)
    const int S = max_span_len;
    std::vector<float> msa(4 * S);
    for (int i = 0; i < P; ++i) {
        std::fill(msa.begin(), msa.end(), 0.f);
        for (long c = 0; c < nch; ++c)
            for (int l = 0; l < 4 * S; ++l) msa[l] += part[((size_t)i * nch + c) * 4 * S + l];
        for (int l = 0; l < 4 * span[i]; ++l) msa[l] += 0.1f;
        const float tot = msa[0] + msa[1] + msa[2] + msa[3];
        for (int a = 0; a < 4; ++a)
            for (int sp = 0; sp < S; ++sp) pfms[i + (size_t)P * (a + 4 * (size_t)sp)] = msa[a + 4 * sp] / tot;
    }
    return CDLM_OK;
}
