// cdlm_cdl.hip -- host side of the CDL ADMM for gfx950: problems, batches, the C ABI.
//
// Reference path (CDLmotif.jl): the `cdl` constructor (src/SETUP/CDL_setup.jl:165-283),
// code_update!/dict_update!/update_best! (src/CDL/train.jl:1-98), objective()
// (CDL_setup.jl:286), the training loop of cdl_optimize (src/WRAPPER/wrap.jl:4-12).
// The kernels and the formulation are in cdlm_cdl_dev.h.
//
// A problem owns its device buffers and a descriptor (CdlDev).  A launch set is a device
// table of descriptors plus prefix arrays that map a block index to (problem, unit); the
// problem's own one-entry set runs the single-dataset entry points, and a batch's set runs
// many problems of different shapes with one launch per kernel (optionally replayed from a
// captured HIP graph).

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>

#include "cdlm_cdl_dev.h"

namespace {


// HIP events around launches (per kernel kind: 0 CSC, 1 dict, 2 best, 3 dict reduce).
struct Timer {
    bool on = false;
    struct Ev {
        hipEvent_t a, b;
        int count;
    };
    std::vector<Ev> ev[4];
    void clear() {
        for (auto &v : ev) {
            for (auto &e : v) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
            v.clear();
        }
    }
};

struct TimedLaunch {
    Timer *t;
    hipStream_t s;
    int kid, count;
    hipEvent_t a = nullptr, b = nullptr;
    TimedLaunch(Timer *t_, hipStream_t s_, int kid_, int count_ = 1) : t(t_), s(s_), kid(kid_), count(count_) {
        if (t && t->on) {
            if (hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess) (void)hipEventRecord(a, s);
            else { (void)hipGetLastError(); a = b = nullptr; }
        }
    }
    ~TimedLaunch() {
        if (a && b) {
            (void)hipEventRecord(b, s);
            t->ev[kid].push_back({a, b, count});
        }
    }
};

// Launch geometry of a set of problems that run with one launch per kernel.
struct LaunchSet {
    int np = 0;
    int taps = 0, nsl = 4, B = 2, LgMax = 0, waves = 4;
    int mfma = 1;   // CSC kernel: 1 = k_csc (vector ALUs), 2 = k_csc_tz (matrix cores)
    int blk_seq = 0, blk_filt = 0, blk_cs = 0, blk_red = 0;
    void *mem = nullptr;  // descriptors followed by the prefix arrays
    CdlDev *tab = nullptr;
    int *st_seq = nullptr, *st_filt = nullptr, *st_cs = nullptr, *st_red = nullptr;
    int *slot = nullptr;  // position of each member in the caller's list (gathered outputs)
    hipStream_t stream = nullptr;  // null: the context stream
    size_t lds_init = 0, lds_csc = 0, lds_best = 0, lds_dict = 0, lds_fin = 0, lds_g1 = 0;
    void release() {
        if (mem) (void)hipFree(mem);
        mem = nullptr;
        tab = nullptr;
    }
};

constexpr size_t kLdsLimit = 160 * 1024;
// Untuned shapes run k_csc_tz (where its LDS fits) instead of k_csc.
constexpr bool kCscMfmaDefault = false;
constexpr int kTraceCap = 4096;

size_t lds_bytes_csc(int Lg, int taps, int waves = kWaves) {
    return sizeof(float) * (2 * (size_t)Lg + (size_t)(Lg + taps + 4) + waves * (size_t)(Lg + taps) + waves * (size_t)Lg +
                            4 * (size_t)Lg);
}
size_t lds_bytes_csctz(int Lg, int taps, int waves) { return sizeof(float) * csctz_lds_floats(Lg, taps, waves); }
// CSC waves per work-group where the slots fit 4 per lane: 8 when two such work-groups' LDS
// fits a CU (two sequences per CU, no register prefetch), else 16 when its LDS fits (one
// sequence per CU), else 4; a tuning of 4, 8 or 16 overrides when it fits.
int pick_csc_waves(int Lg, int taps, int nsl, int tuned) {
    const bool ok8 = nsl == 4 && lds_bytes_csc(Lg, taps, 8) <= kLdsLimit;
    const bool two8 = nsl == 4 && 2 * lds_bytes_csc(Lg, taps, 8) <= kLdsLimit;
    const bool ok16 = nsl == 4 && lds_bytes_csc(Lg, taps, 16) <= kLdsLimit;
    if (tuned == 8) return ok8 ? 8 : 4;
    if (tuned == 16) return ok16 ? 16 : 4;
    if (tuned == 4) return 4;
    return two8 ? 8 : ok16 ? 16 : 4;
}
// CSC kernel and waves per work-group.  k_csc_tz (mfma = 2, or 0 where it is the default) runs
// 8 waves where two work-groups' LDS fits a CU, else 16 (one per CU), or the tuned 8 or 16
// where that fits; otherwise, and for mfma = 1, k_csc with pick_csc_waves.
void pick_csc(int Lg, int taps, int nsl, int tuned_mfma, int tuned_waves, int *mfma, int *waves) {
    if (tuned_mfma == 2 || (tuned_mfma == 0 && kCscMfmaDefault)) {
        const size_t b8 = lds_bytes_csctz(Lg, taps, 8), b16 = lds_bytes_csctz(Lg, taps, 16);
        int wv = 0;
        if (tuned_waves == 8 && b8 <= kLdsLimit) wv = 8;
        else if (tuned_waves == 16 && b16 <= kLdsLimit) wv = 16;
        else if (2 * b8 <= kLdsLimit) wv = 8;
        else if (b16 <= kLdsLimit) wv = 16;
        else if (b8 <= kLdsLimit) wv = 8;
        if (wv) {
            *mfma = 2;
            *waves = wv;
            return;
        }
    }
    *mfma = 1;
    *waves = pick_csc_waves(Lg, taps, nsl, tuned_mfma == 2 ? 0 : tuned_waves);
}
size_t lds_bytes_init(int Lg, int taps) { return sizeof(float) * (kWaves * (size_t)(Lg + taps) + kWaves * (size_t)Lg); }
size_t lds_bytes_best(int Lg, int taps) { return sizeof(float) * (size_t)(Lg + taps + 4); }
// k_dict with W waves: twiddles, h with its tail, three spectrum accumulators and the stored
// kappa, two length-Lg complex buffers, and per wave two of length Lg/4 + 2 and the stored Y3h.
size_t lds_bytes_dict(int Lg, int taps, int W) {
    const size_t Hp = (size_t)(((Lg / 2 + 1) + 3) & ~3);
    return sizeof(float) * (2 * (size_t)Lg + (size_t)((Lg + taps + 4 + 3) & ~3) + 7 * Hp + 4 * (size_t)Lg +
                            (size_t)W * ((size_t)Lg + 8 + 2 * (size_t)dict_phx(Lg)));
}
size_t lds_bytes_half(int Lg) { return sizeof(float2) * 3 * (size_t)Lg; }
size_t lds_bytes_fin(int Lg) { return sizeof(float) * (((size_t)Lg + 3) / 4 * 4 + 6 * (size_t)Lg); }
size_t lds_bytes_g1(int Lg) { return sizeof(float) * 6 * (size_t)Lg; }

// Work-groups of W dictionary waves that fit one CU: by registers (k_dict keeps 3 waves per
// SIMD for padded lengths up to 1020 and filters up to 24, else 2) and by LDS.
int dict_wg_per_cu(int Lg, int taps, int W) {
    const size_t b = lds_bytes_dict(Lg, taps, W);
    if (b > kLdsLimit - kThreads * sizeof(double)) return 0;
    const int wps = (Lg <= 1020 && taps <= 24) ? 3 : 2;
    return std::min<int>(4 * wps / W, (int)(kLdsLimit / b));
}
// Dictionary waves per work-group (columns in flight, one per wave): the most waves per CU;
// on a tie 4, then 8.
int pick_dict_batch(int Lg, int taps) {
    int best = 0, bestv = 0;
    for (int W : {4, 8, 2}) {
        const int v = W * dict_wg_per_cu(Lg, taps, W);
        if (v > bestv) { bestv = v; best = W; }
    }
    return best;
}

// Chunks of sequences in the dictionary mean: about 4096 blocks per problem and at least 32
// chunks, a multiple of the 8 XCDs (k_dict_reduce keeps every tile of a chunk on one XCD, so the
// chunk's stored Y3 half spectra and kappa, which all its tiles read, come from HBM once into that
// XCD's L2), and no more than N rounded up to 8 (at least 32): then every chunk holds at most one
// sequence, the sums are the same as with more chunks, and the partial sums (4 doubles per bin,
// filter and chunk) stay small for small N.  A streamed problem moves through the device in a
// divisor of this many chunks, so that its means are summed exactly as when it is resident.
constexpr int kXcds = 8;
int pick_nch(int Lg, int M, int N) {
    const size_t HM = (size_t)half_bins(Lg) * M;
    const size_t tiles = (HM + kThreads - 1) / kThreads;
    int nch = (int)((4096 + tiles - 1) / tiles);
    nch = (nch + kXcds - 1) / kXcds * kXcds;
    if (nch < 32) nch = 32;
    const int cap = std::max(32, (N + kXcds - 1) / kXcds * kXcds);
    return std::min(nch, cap);
}

// Dynamic LDS above the default needs hipFuncSetAttribute.  The attribute is per kernel
// (and device), so it is only ever raised, to the largest size any launch set needs.
std::mutex g_lds_mu;
std::map<std::pair<int, const void *>, size_t> g_lds;

cdlm_status ensure_lds(cdlm_ctx *ctx, const void *fn, size_t bytes) {
    std::lock_guard<std::mutex> lk(g_lds_mu);
    size_t &cur = g_lds[{ctx->device, fn}];
    if (bytes <= cur) return CDLM_OK;
    CDLM_HIP(ctx, hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    cur = bytes;
    return CDLM_OK;
}

// Launch choices of one problem shape (Lg, M, max filter length), set by the autotuner or
// by the caller; kept per device for the life of the process.  Only the FFT stage orders
// change rounding; the others change scheduling only.
struct Tuning {
    int order = kFftSmallFirst;   // stage order of the length-Lg transforms
    int order4 = kFftSmallFirst;  // ... of the length-Lg/4 transforms (dictionary step)
    int B = 0;                    // dictionary waves per work-group, one column each (0: by occupancy)
    int nsl = 0;                  // CSC float4 slots per lane (0: the fewest that cover Lg)
    int waves = 0;                // CSC waves per work-group (4 or 16; 0: the default for the shape)
    int mfma = 0;                 // CSC kernel (1: k_csc, 2: k_csc_tz; 0: the default for the shape)
};
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int>, Tuning> g_tune;

Tuning tuning_of(int device, int Lg, int M, int gamma) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(std::make_tuple(device, Lg, M, gamma));
    return it == g_tune.end() ? Tuning() : it->second;
}

cdlm_status lcheck(cdlm_ctx *ctx, const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return cdlm_fail(ctx, CDLM_ERR_HIP, "launch of %s failed: %s", what, hipGetErrorString(e));
    return CDLM_OK;
}

}  // namespace

// ============================================================================ problem
struct cdlm_problem {
    cdlm_ctx *ctx = nullptr;
    int L4 = 0, N = 0, M = 0, Lg = 0, gamma = 0, L = 0;
    std::vector<int> fl;
    // all device buffers are carved from one arena (256-byte aligned, in the order of the
    // buffer plan); an arena handed in by the caller (the server's recycled memory) is not
    // freed with the problem
    void *arena = nullptr;
    size_t arena_bytes = 0, arena_used = 0;
    bool own_arena = true;
    size_t total_bytes = 0;
    CdlDev dv;           // descriptor with the device pointers
    LaunchSet set;       // this problem alone
    Timer timer;
    Tuning tune;         // launch choices in force for this problem
    int batches = 0;     // batches that reference this problem
    // Streamed: the state of all N sequences lives in pinned host memory and the sequences move
    // through a device window of `win` sequences in nchunk chunks (0: a resident problem).
    struct HostArr {
        char *dev[2] = {nullptr, nullptr};   // the windows' buffers (dev[1]: with two windows)
        char *host = nullptr;                // N * per_seq bytes
        size_t per_seq = 0;
    };
    // With two windows, chunk k uses window k % 2: chunk k+1 moves in and chunk k-1 out (on two
    // copy streams, the two directions of the bus at once) while chunk k's kernels run.
    int nchunk = 0, win = 0, nwin = 1;
    hipStream_t cin = nullptr, cout = nullptr;
    std::vector<hipEvent_t> ev;        // per chunk: in, kernels, out
    std::vector<int> cn0;              // chunk k: sequences [cn0[k], cn0[k+1])
    std::vector<LaunchSet> csets;      // chunk k's launch set: its N, n_off and mean chunks
    std::vector<HostArr> hstate;       // moved in and out around every chunk pass
    HostArr hs, hx0, hxs[2];           // data and initial codes (in only), X_best slices (out)
    std::vector<void *> pinned;
};

struct cdlm_batch {
    cdlm_ctx *ctx = nullptr;
    std::vector<cdlm_problem *> probs;
    // one launch set per shape class (members sharing slots per lane, dictionary batch,
    // filter taps and LDS occupancy); in a graph each class is a concurrent branch
    std::vector<LaunchSet> sets;
    std::vector<hipStream_t> streams;  // capture streams of classes 1.. (class 0: the context stream)
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> joins;
    Timer timer;
    bool use_graph = true;
    // one cached graph: (reset, iter_num, csc_iter, cdl_iter, gather count)
    hipGraphExec_t exec = nullptr;
    int key[5] = {-1, -1, -1, -1, -1};
    double *stage = nullptr;  // device gather buffer
    size_t stage_cap = 0;
    std::vector<double> host;
};

namespace {

// Buffer plan of one problem: sizes in order of allocation.  The footprint query and the
// allocator both use it, so cdlm_problem_bytes == cdlm_footprint_bytes.
struct BufPlan {
    size_t s, E, Eh, E3, Ex0, LN, LM, LMh, Lg, xs, y3s, kap, M, partials, dpart, trace, sc;
};

// N sequences resident; the per-sequence objective partials are held for Ntot (a streamed
// problem's window holds a chunk of Ntot sequences; resident: Ntot = N).
BufPlan buf_plan(int L4, int N, int M, int gamma, bool keep_x0, int Ntot = 0) {
    if (Ntot <= 0) Ntot = N;
    BufPlan b;
    const size_t Lg = (size_t)L4 + 2 * (size_t)gamma;
    b.s = (size_t)L4 * N * sizeof(float);
    b.E = Lg * M * (size_t)N * sizeof(float);
    b.Eh = (size_t)half_bins((int)Lg) * M * (size_t)N * sizeof(float2);   // g1 as half spectra
    b.E3 = 3 * (Lg / 4) * M * (size_t)N * sizeof(float);   // y3 off the stride-4 support
    b.Ex0 = keep_x0 ? b.E : 0;
    b.LN = Lg * (size_t)N * sizeof(float);
    b.LM = Lg * (size_t)M * sizeof(float);
    b.LMh = (size_t)half_bins((int)Lg) * M * sizeof(float2);
    b.Lg = Lg * sizeof(float);
    b.xs = (size_t)(L4 / 4) * M * (size_t)N * sizeof(float);
    b.y3s = (size_t)M * N * (Lg / 8 + 1) * sizeof(float2);
    b.kap = (size_t)half_bins((int)Lg) * N * sizeof(float2);
    b.M = (size_t)M * sizeof(int);
    b.partials = (size_t)Ntot * kPartStride * sizeof(double);
    b.dpart = (size_t)pick_nch((int)Lg, M, Ntot) * kDpart * ((size_t)half_bins((int)Lg) * M) * sizeof(double);
    b.trace = (size_t)kTraceCap * sizeof(double);
    b.sc = sizeof(CdlScal);
    return b;
}

constexpr size_t kBufAlign = 256;
size_t al(size_t x) { return (x + kBufAlign - 1) & ~(kBufAlign - 1); }

// Bytes of one window of per-sequence buffers (data, state, X_best slices).
size_t plan_window(const BufPlan &b) {
    return al(b.s) + 2 * al(b.E) + al(b.Eh) + al(b.E3) + (b.Ex0 ? al(b.Ex0) : 0) + 7 * al(b.LN) + 2 * al(b.xs) +
           al(b.y3s) + al(b.kap);
}
// Arena bytes of a plan: every buffer rounded up to kBufAlign, in allocation order; a streamed
// problem with two windows holds a second set of the per-sequence buffers.
size_t plan_total(const BufPlan &b, int wins = 1) {
    return (size_t)wins * plan_window(b) + 7 * al(b.LM) + 6 * al(b.LMh) + 2 * al(b.Lg) /* cS, sS */ +
           al(2 * b.Lg) /* tw */ + al(b.M) + al(b.partials) + al(b.dpart) + al(b.trace) + al(b.sc);
}

cdlm_status palloc(cdlm_problem *p, void **out, size_t bytes) {
    const size_t need = al(bytes ? bytes : 4);
    if (p->arena_used + need > p->arena_bytes)
        return cdlm_fail(p->ctx, CDLM_ERR_STATE, "buffer plan exceeds its arena (%zu + %zu > %zu bytes)", p->arena_used,
                         need, p->arena_bytes);
    *out = static_cast<char *>(p->arena) + p->arena_used;
    p->arena_used += need;
    return CDLM_OK;
}

template <typename T>
cdlm_status palloc_t(cdlm_problem *p, T **out, size_t bytes) {
    void *q = nullptr;
    cdlm_status st = palloc(p, &q, bytes);
    *out = static_cast<T *>(q);
    return st;
}

void free_problem(cdlm_problem *p) {
    if (p->cin) (void)hipStreamSynchronize(p->cin);
    if (p->cout) (void)hipStreamSynchronize(p->cout);
    for (hipEvent_t e : p->ev) (void)hipEventDestroy(e);
    p->ev.clear();
    if (p->cin) (void)hipStreamDestroy(p->cin);
    if (p->cout) (void)hipStreamDestroy(p->cout);
    p->cin = p->cout = nullptr;
    p->set.release();
    for (auto &cs : p->csets) cs.release();
    p->csets.clear();
    for (void *h : p->pinned) (void)hipHostFree(h);
    p->pinned.clear();
    p->timer.clear();
    if (p->own_arena && p->arena) (void)hipFree(p->arena);
    p->arena = nullptr;
    p->arena_bytes = p->arena_used = 0;
}

// Builds the launch set of np problems (all on ctx): maxima, LDS sizes, device table.
// slots[i] (identity when null) is member i's position in the caller's list.
cdlm_status set_build(cdlm_ctx *ctx, cdlm_problem *const *ps, int np, LaunchSet *S, const int *slots = nullptr) {
    LaunchSet T;
    T.np = np;
    for (int i = 0; i < np; ++i) {
        T.taps = std::max(T.taps, ps[i]->gamma);
        T.LgMax = std::max(T.LgMax, ps[i]->Lg);
    }
    const int Lg = T.LgMax, taps = T.taps;
    if (Lg > 2048) return cdlm_fail(ctx, CDLM_ERR_UNSUPPORTED, "padded length %d > 2048", Lg);
    // slots per lane: enough for the longest member, more if a member's tuning asks;
    // dictionary batch: the tuning of the longest member when it fits, else by occupancy
    T.nsl = (Lg / 4 <= 256) ? 4 : 8;
    int lead = 0;
    for (int i = 0; i < np; ++i) {
        T.nsl = std::max(T.nsl, ps[i]->tune.nsl);
        if (ps[i]->Lg > ps[lead]->Lg) lead = i;
    }
    T.B = pick_dict_batch(Lg, taps);
    const int tuned_B = ps[lead]->tune.B;
    if ((tuned_B == 2 || tuned_B == 4 || tuned_B == 8) && lds_bytes_dict(Lg, taps, tuned_B) <= kLdsLimit - kThreads * sizeof(double))
        T.B = tuned_B;
    if (T.B == 0) return cdlm_fail(ctx, CDLM_ERR_UNSUPPORTED, "padded length %d: dictionary step exceeds LDS", Lg);
    pick_csc(Lg, taps, T.nsl, ps[lead]->tune.mfma, ps[lead]->tune.waves, &T.mfma, &T.waves);
    T.lds_csc = T.mfma == 2 ? lds_bytes_csctz(Lg, taps, T.waves) : lds_bytes_csc(Lg, taps, T.waves);
    T.lds_init = lds_bytes_init(Lg, taps);
    T.lds_best = lds_bytes_best(Lg, taps);
    T.lds_dict = lds_bytes_dict(Lg, taps, T.B);
    T.lds_fin = lds_bytes_fin(Lg);
    T.lds_g1 = lds_bytes_g1(Lg);
    size_t mx = T.lds_csc;
    for (size_t v : {T.lds_init, T.lds_best, T.lds_dict, T.lds_fin, T.lds_g1}) mx = std::max(mx, v);
    if (mx > kLdsLimit - kThreads * sizeof(double))
        return cdlm_fail(ctx, CDLM_ERR_UNSUPPORTED, "padded length %d needs %zu B of LDS", Lg, mx);

    cdlm_status st = CDLM_OK;
#define SETA(TT)                                                                                   \
    do {                                                                                           \
        if ((st = ensure_lds(ctx, (const void *)k_init<TT>, T.lds_init)) != CDLM_OK) return st;    \
        if ((st = ensure_lds(ctx, (const void *)k_csc<TT, 4, 4>, T.lds_csc)) != CDLM_OK) return st;   \
        if ((st = ensure_lds(ctx, (const void *)k_csc<TT, 8, 4>, T.lds_csc)) != CDLM_OK) return st;   \
        if ((st = ensure_lds(ctx, (const void *)k_csc<TT, 4, 16>, T.lds_csc)) != CDLM_OK) return st;  \
        if ((st = ensure_lds(ctx, (const void *)k_csc<TT, 4, 8>, T.lds_csc)) != CDLM_OK) return st;   \
        if ((st = ensure_lds(ctx, (const void *)k_csc_tz<TT, 4, 8>, T.lds_csc)) != CDLM_OK) return st;   \
        if ((st = ensure_lds(ctx, (const void *)k_csc_tz<TT, 4, 16>, T.lds_csc)) != CDLM_OK) return st;  \
        if ((st = ensure_lds(ctx, (const void *)k_csc_tz<TT, 8, 8>, T.lds_csc)) != CDLM_OK) return st;   \
        if ((st = ensure_lds(ctx, (const void *)k_csc_tz<TT, 8, 16>, T.lds_csc)) != CDLM_OK) return st;  \
        if ((st = ensure_lds(ctx, (const void *)k_best<TT>, T.lds_best)) != CDLM_OK) return st;    \
        if ((st = ensure_lds(ctx, (const void *)k_dict<TT, 4>, T.lds_dict)) != CDLM_OK) return st;         \
        if ((st = ensure_lds(ctx, (const void *)k_dict<TT, 8>, T.lds_dict)) != CDLM_OK) return st;         \
    } while (0)
    switch (taps) {
    case 4: SETA(4); break;
    case 8: SETA(8); break;
    case 12: SETA(12); break;
    case 16: SETA(16); break;
    case 20: SETA(20); break;
    case 24: SETA(24); break;
    case 28: SETA(28); break;
    case 32: SETA(32); break;
    default: return cdlm_fail(ctx, CDLM_ERR_UNSUPPORTED, "filter length %d unsupported", taps);
    }
#undef SETA
    if ((st = ensure_lds(ctx, (const void *)k_dict_finalize, T.lds_fin)) != CDLM_OK) return st;
    if ((st = ensure_lds(ctx, (const void *)k_g1_spectrum, T.lds_g1)) != CDLM_OK) return st;
    if ((st = ensure_lds(ctx, (const void *)k_dict_to_spatial, lds_bytes_half(Lg))) != CDLM_OK) return st;

    std::vector<CdlDev> tab(np);
    std::vector<int> pre(5 * (size_t)np);
    long a_seq = 0, a_filt = 0, a_cs = 0, a_red = 0;
    for (int i = 0; i < np; ++i) {
        const cdlm_problem *p = ps[i];
        tab[i] = p->dv;
        pre[i] = (int)a_seq;
        pre[np + i] = (int)a_filt;
        pre[2 * np + i] = (int)a_cs;
        pre[3 * np + i] = (int)a_red;
        pre[4 * np + i] = slots ? slots[i] : i;
        const long HM = (long)half_bins(p->Lg) * p->M;
        a_seq += p->dv.N;
        a_filt += p->M;
        a_cs += (p->Lg + kThreads - 1) / kThreads;
        a_red += ((HM + kThreads - 1) / kThreads) * p->dv.red_cn;
    }
    if (a_seq > INT32_MAX || a_red > INT32_MAX) return cdlm_fail(ctx, CDLM_ERR_UNSUPPORTED, "batch too large for one grid");
    T.blk_seq = (int)a_seq;
    T.blk_filt = (int)a_filt;
    T.blk_cs = (int)a_cs;
    T.blk_red = (int)a_red;
    const size_t tb = ((sizeof(CdlDev) * np + 255) / 256) * 256;
    const size_t bytes = tb + pre.size() * sizeof(int);
    hipError_t e = hipMalloc(&T.mem, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return cdlm_fail(ctx, CDLM_ERR_OOM, "hipMalloc(%zu bytes) for a launch table failed", bytes);
    }
    T.tab = static_cast<CdlDev *>(T.mem);
    int *pr = reinterpret_cast<int *>(static_cast<char *>(T.mem) + tb);
    T.st_seq = pr;
    T.st_filt = pr + np;
    T.st_cs = pr + 2 * np;
    T.st_red = pr + 3 * np;
    T.slot = pr + 4 * np;
    e = hipMemcpyAsync(T.tab, tab.data(), sizeof(CdlDev) * np, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(pr, pre.data(), pre.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        T.release();
        return cdlm_fail(ctx, CDLM_ERR_HIP, "launch table upload: %s", hipGetErrorString(e));
    }
    *S = T;
    return CDLM_OK;
}

// ---------------------------------------------------------------------------- launches
hipStream_t set_stream(cdlm_ctx *ctx, const LaunchSet &S) { return S.stream ? S.stream : ctx->stream; }

cdlm_status run_g1_cs(cdlm_ctx *ctx, const LaunchSet &S, int flip) {
    hipStream_t st = set_stream(ctx, S);
    if (!flip)
        hipLaunchKernelGGL(k_g1_spectrum, dim3(S.blk_filt), dim3(kThreads), S.lds_g1, st, S.tab, S.st_filt, S.np, S.LgMax);
    hipLaunchKernelGGL(k_cs, dim3(S.blk_cs), dim3(kThreads), 0, st, S.tab, S.st_cs, S.np, flip);
    return lcheck(ctx, "k_g1_spectrum/k_cs");
}

// Objective partials, their reduction and the update_best decision.  decide = 0 only
// records the objective; slot < 0 writes no trace entry.
cdlm_status run_best_seq(cdlm_ctx *ctx, const LaunchSet &S, Timer *tm, float *Xfull) {
    hipStream_t st = set_stream(ctx, S);
    TimedLaunch tl(tm, st, 2);
#define L_BEST(T) hipLaunchKernelGGL(k_best<T>, dim3(S.blk_seq), dim3(kThreads), S.lds_best, st, S.tab, S.st_seq, S.np, Xfull)
    CDLM_TAPS_SWITCH(S.taps, L_BEST);
#undef L_BEST
    return lcheck(ctx, "k_best");
}
cdlm_status run_best_global(cdlm_ctx *ctx, const LaunchSet &S, int force, int decide, int slot) {
    hipStream_t st = set_stream(ctx, S);
    hipLaunchKernelGGL(k_reduce_partials, dim3(S.np), dim3(kThreads), 0, st, S.tab);
    hipLaunchKernelGGL(k_decide, dim3(S.np), dim3(kThreads), 0, st, S.tab, force, decide, slot);
    return lcheck(ctx, "k_decide");
}
cdlm_status run_best(cdlm_ctx *ctx, const LaunchSet &S, Timer *tm, int force, int decide, int slot, float *Xfull) {
    cdlm_status s = run_best_seq(ctx, S, tm, Xfull);
    return s != CDLM_OK ? s : run_best_global(ctx, S, force, decide, slot);
}

// iters CSC iterations in one launch (the timer counts iterations).
cdlm_status run_csc(cdlm_ctx *ctx, const LaunchSet &S, Timer *tm, int iters) {
    if (iters <= 0) return CDLM_OK;
    hipStream_t st = set_stream(ctx, S);
    {
        TimedLaunch tl(tm, st, 0, iters);
#define L_CSC(T)                                                                                                     \
    if (S.mfma == 2 && S.LgMax <= 1024 && S.waves == 16)                                                             \
        hipLaunchKernelGGL((k_csc_tz<T, 4, 16>), dim3(S.blk_seq), dim3(64 * 16), S.lds_csc, st, S.tab, S.st_seq, S.np, \
                           iters);                                                                                   \
    else if (S.mfma == 2 && S.LgMax <= 1024)                                                                         \
        hipLaunchKernelGGL((k_csc_tz<T, 4, 8>), dim3(S.blk_seq), dim3(64 * 8), S.lds_csc, st, S.tab, S.st_seq, S.np,  \
                           iters);                                                                                   \
    else if (S.mfma == 2 && S.waves == 16)                                                                           \
        hipLaunchKernelGGL((k_csc_tz<T, 8, 16>), dim3(S.blk_seq), dim3(64 * 16), S.lds_csc, st, S.tab, S.st_seq, S.np, \
                           iters);                                                                                   \
    else if (S.mfma == 2)                                                                                            \
        hipLaunchKernelGGL((k_csc_tz<T, 8, 8>), dim3(S.blk_seq), dim3(64 * 8), S.lds_csc, st, S.tab, S.st_seq, S.np,  \
                           iters);                                                                                   \
    else if (S.waves == 16)                                                                                               \
        hipLaunchKernelGGL((k_csc<T, 4, 16>), dim3(S.blk_seq), dim3(64 * 16), S.lds_csc, st, S.tab, S.st_seq, S.np,  \
                           iters);                                                                                   \
    else if (S.waves == 8)                                                                                           \
        hipLaunchKernelGGL((k_csc<T, 4, 8>), dim3(S.blk_seq), dim3(64 * 8), S.lds_csc, st, S.tab, S.st_seq, S.np,    \
                           iters);                                                                                   \
    else if (S.nsl == 4)                                                                                             \
        hipLaunchKernelGGL((k_csc<T, 4, 4>), dim3(S.blk_seq), dim3(kThreads), S.lds_csc, st, S.tab, S.st_seq, S.np,  \
                           iters);                                                                                   \
    else                                                                                                             \
        hipLaunchKernelGGL((k_csc<T, 8, 4>), dim3(S.blk_seq), dim3(kThreads), S.lds_csc, st, S.tab, S.st_seq, S.np,  \
                           iters)
        CDLM_TAPS_SWITCH(S.taps, L_CSC);
#undef L_CSC
    }
    return lcheck(ctx, "k_csc");
}

// The per-sequence part of a dictionary iteration (k_dict and the partial sums of the means).
cdlm_status run_dict_seq(cdlm_ctx *ctx, const LaunchSet &S, Timer *tm) {
    hipStream_t st = set_stream(ctx, S);
    {
        TimedLaunch tl(tm, st, 1);
#define L_DICT(T)                                                                                                    \
    if (S.LgMax <= 1020)                                                                                             \
        hipLaunchKernelGGL((k_dict<T, 4>), dim3(S.blk_seq), dim3(64 * S.B), S.lds_dict, st, S.tab, S.st_seq, S.np,   \
                           S.LgMax);                                                                                 \
    else                                                                                                             \
        hipLaunchKernelGGL((k_dict<T, 8>), dim3(S.blk_seq), dim3(64 * S.B), S.lds_dict, st, S.tab, S.st_seq, S.np,   \
                           S.LgMax)
        CDLM_TAPS_SWITCH(S.taps, L_DICT);
#undef L_DICT
    }
    {
        TimedLaunch tl(tm, st, 3);
        hipLaunchKernelGGL(k_dict_reduce, dim3(S.blk_red), dim3(kThreads), 0, st, S.tab, S.st_red, S.np);
    }
    return lcheck(ctx, "k_dict");
}
// The per-filter part: the means back to the filters, the new filter spectra.
cdlm_status run_dict_filt(cdlm_ctx *ctx, const LaunchSet &S) {
    hipStream_t st = set_stream(ctx, S);
    hipLaunchKernelGGL(k_dict_finalize, dim3(S.blk_filt), dim3(kThreads), S.lds_fin, st, S.tab, S.st_filt, S.np, S.LgMax);
    cdlm_status s = lcheck(ctx, "k_dict_finalize");
    if (s != CDLM_OK) return s;
    return run_g1_cs(ctx, S, 1);
}
cdlm_status run_dict(cdlm_ctx *ctx, const LaunchSet &S, Timer *tm) {
    cdlm_status s = run_dict_seq(ctx, S, tm);
    return s != CDLM_OK ? s : run_dict_filt(ctx, S);
}

// The constructor's device work (CDL_setup.jl:183-236): filters, |G1h|^2, codes, DX, and
// the forced first update_best.
cdlm_status run_reset_filt(cdlm_ctx *ctx, const LaunchSet &S) {
    hipStream_t st = set_stream(ctx, S);
    hipLaunchKernelGGL(k_reset, dim3(S.blk_filt), dim3(kThreads), 0, st, S.tab, S.st_filt, S.np);
    return run_g1_cs(ctx, S, 0);
}
cdlm_status run_init_seq(cdlm_ctx *ctx, const LaunchSet &S) {
    hipStream_t st = set_stream(ctx, S);
#define L_INIT(T) hipLaunchKernelGGL(k_init<T>, dim3(S.blk_seq), dim3(kThreads), S.lds_init, st, S.tab, S.st_seq, S.np)
    CDLM_TAPS_SWITCH(S.taps, L_INIT);
#undef L_INIT
    return lcheck(ctx, "k_init");
}
cdlm_status run_reset(cdlm_ctx *ctx, const LaunchSet &S) {
    cdlm_status s = run_reset_filt(ctx, S);
    if (s == CDLM_OK) s = run_init_seq(ctx, S);
    if (s != CDLM_OK) return s;
    return run_best(ctx, S, nullptr, 1, 1, -1, nullptr);
}

// wrap.jl:6-11: iter_num x (csc_iter CSC iterations, update_best, cdl_iter dictionary
// iterations); update_best of outer iteration `it` writes trace entry it.
cdlm_status run_schedule(cdlm_ctx *ctx, const LaunchSet &S, Timer *tm, int iter_num, int csc_iter, int cdl_iter) {
    cdlm_status st;
    for (int it = 0; it < iter_num; ++it) {
        if ((st = run_csc(ctx, S, tm, csc_iter)) != CDLM_OK) return st;
        if ((st = run_best(ctx, S, tm, 0, 1, it, nullptr)) != CDLM_OK) return st;
        for (int i = 0; i < cdl_iter; ++i)
            if ((st = run_dict(ctx, S, tm)) != CDLM_OK) return st;
    }
    return CDLM_OK;
}

cdlm_status sync_ctx(cdlm_ctx *ctx) {
    CDLM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CDLM_OK;
}

cdlm_status read_scal(cdlm_problem *p, CdlScal *sc) {
    CDLM_HIP(p->ctx, hipMemcpyAsync(sc, p->dv.sc, sizeof(CdlScal), hipMemcpyDeviceToHost, p->ctx->stream));
    return sync_ctx(p->ctx);
}

// std(X_best[X_best .> 0]) * mult (wrap.jl:12) from count, sum and sum of squares; the
// reference's std is Float32.
double thresh_of(double cnt, double s1, double s2, double mult) {
    if (cnt < 2) return NAN;
    double var = (s2 - s1 * s1 / cnt) / (cnt - 1.0);
    if (var < 0) var = 0;
    return mult * (double)(float)sqrt(var);
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

}  // namespace

// ============================================================================ C ABI: problem
extern "C" size_t cdlm_footprint_bytes(int32_t L4, int32_t N, int32_t M, int32_t gamma, int32_t keep_x0) {
    if (L4 <= 0 || N <= 0 || M <= 0 || gamma < 0) return 0;
    return plan_total(buf_plan(L4, N, M, gamma, keep_x0 != 0));
}

size_t cdlm_stream_min_bytes(int32_t L4, int32_t N, int32_t M, int32_t gamma) {
    if (L4 <= 0 || N <= 0 || M <= 0 || gamma < 0) return 0;
    const int nch = pick_nch(L4 + 2 * gamma, M, N);
    int mx = 1;
    for (int c = 0; c < nch; ++c) mx = std::max(mx, (int)((long)(c + 1) * N / nch) - (int)((long)c * N / nch));
    return plan_total(buf_plan(L4, mx, M, gamma, false, N));
}

namespace {
// One-hot data (4L, N) column-major from base codes (N, L) row-major: s[4l + a, n] = 1 iff
// code(n, l) == a (data_2_dummy, to_dummy.jl:2-38); a code > 3 gives an all-zero group.
__global__ void k_bases_to_onehot(const uint8_t *__restrict__ codes, int L, long N, float *__restrict__ s) {
    const long tot = (long)L * N;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += (long)gridDim.x * blockDim.x) {
        const int b = codes[i];
        float4 v = make_float4(b == 0 ? 1.f : 0.f, b == 1 ? 1.f : 0.f, b == 2 ? 1.f : 0.f, b == 3 ? 1.f : 0.f);
        reinterpret_cast<float4 *>(s)[i] = v;   // (n, l) row-major = column n, rows 4l..4l+3
    }
}

cdlm_status problem_create_impl(cdlm_ctx *ctx, const cdlm_params *prm, const float *s, const uint8_t *codes, int32_t L4,
                                int32_t N, const float *D0, const float *X0, uint64_t seed, cdlm_problem **out,
                                void *arena = nullptr, size_t arena_bytes = 0, uint64_t stream_budget = 0);
}  // namespace

extern "C" cdlm_status cdlm_problem_create_streamed(cdlm_ctx *ctx, const cdlm_params *prm, const float *s, int32_t L4,
                                                    int32_t N, const float *D0, const float *X0, uint64_t seed,
                                                    uint64_t device_budget, cdlm_problem **out) {
    if (!ctx) return CDLM_ERR_ARG;
    if (!s) return cdlm_fail(ctx, CDLM_ERR_ARG, "null argument");
    if (device_budget == 0) {
        size_t f = 0, t = 0;
        CDLM_HIP(ctx, hipMemGetInfo(&f, &t));
        device_budget = (uint64_t)(0.8 * (double)f);
    }
    return problem_create_impl(ctx, prm, s, nullptr, L4, N, D0, X0, seed, out, nullptr, 0, device_budget);
}

extern "C" cdlm_status cdlm_problem_chunks(const cdlm_problem *p, int32_t *chunks) {
    if (!p || !chunks) return CDLM_ERR_ARG;
    *chunks = p->nchunk ? p->nchunk : 1;
    return CDLM_OK;
}

cdlm_status cdlm_problem_create_in(cdlm_ctx *ctx, const cdlm_params *prm, const float *s, int32_t L4, int32_t N,
                                   const float *D0, uint64_t seed, void *arena, size_t arena_bytes, cdlm_problem **out) {
    if (!ctx) return CDLM_ERR_ARG;
    if (!s || !arena) return cdlm_fail(ctx, CDLM_ERR_ARG, "null argument");
    return problem_create_impl(ctx, prm, s, nullptr, L4, N, D0, nullptr, seed, out, arena, arena_bytes);
}

extern "C" cdlm_status cdlm_problem_create(cdlm_ctx *ctx, const cdlm_params *prm, const float *s, int32_t L4, int32_t N,
                                           const float *D0, const float *X0, uint64_t seed, cdlm_problem **out) {
    if (!ctx) return CDLM_ERR_ARG;
    if (!s) return cdlm_fail(ctx, CDLM_ERR_ARG, "null argument");
    return problem_create_impl(ctx, prm, s, nullptr, L4, N, D0, X0, seed, out);
}

extern "C" cdlm_status cdlm_problem_create_from_bases(cdlm_ctx *ctx, const cdlm_params *prm, const uint8_t *bases,
                                                      int32_t L, int32_t N, const float *D0, const float *X0, uint64_t seed,
                                                      cdlm_problem **out) {
    if (!ctx) return CDLM_ERR_ARG;
    if (!bases) return cdlm_fail(ctx, CDLM_ERR_ARG, "null argument");
    if (L <= 0 || L > (1 << 28)) return cdlm_fail(ctx, CDLM_ERR_ARG, "bad sequence length %d", L);
    return problem_create_impl(ctx, prm, nullptr, bases, 4 * L, N, D0, X0, seed, out);
}

namespace {
// ---------------------------------------------------------------------------- streamed problems
// A problem larger than its device budget keeps every per-sequence array in pinned host memory;
// each pass over the sequences moves chunk k into the device window (the state read), runs the
// per-sequence kernels on it with chunk k's launch set, and moves the state written back.  The
// per-filter state, the per-sequence objective partials (all N) and the partial sums of the
// dictionary means (all chunks) stay on the device, so the per-filter kernels and the fixed-order
// reductions see exactly what they see for a resident problem: a streamed training gives the
// resident one's results bit for bit.  Chunks are unions of the dictionary-mean chunks.

// Moves chunk k's state (and the data, and with x0 the initial codes) into its window, on stream st.
cdlm_status stream_in(cdlm_problem *p, int k, bool x0, hipStream_t st) {
    const int w = k % p->nwin;
    const size_t a = (size_t)p->cn0[k], n = (size_t)(p->cn0[k + 1] - p->cn0[k]);
    auto in = [&](const cdlm_problem::HostArr &h) {
        return hipMemcpyAsync(h.dev[w], h.host + a * h.per_seq, n * h.per_seq, hipMemcpyHostToDevice, st);
    };
    hipError_t e = in(p->hs);
    for (size_t i = 0; i < p->hstate.size() && e == hipSuccess; ++i) e = in(p->hstate[i]);
    if (e == hipSuccess && x0 && p->hx0.host) e = in(p->hx0);
    if (e != hipSuccess) return cdlm_fail(p->ctx, CDLM_ERR_HIP, "chunk upload: %s", hipGetErrorString(e));
    return CDLM_OK;
}

// Moves chunk k's state back from its window, on stream st; cand >= 0: also the candidate X_best
// slice buffer cand.
cdlm_status stream_out(cdlm_problem *p, int k, int cand, hipStream_t st) {
    const int w = k % p->nwin;
    const size_t a = (size_t)p->cn0[k], n = (size_t)(p->cn0[k + 1] - p->cn0[k]);
    auto out = [&](const cdlm_problem::HostArr &h) {
        return hipMemcpyAsync(h.host + a * h.per_seq, h.dev[w], n * h.per_seq, hipMemcpyDeviceToHost, st);
    };
    hipError_t e = hipSuccess;
    for (size_t i = 0; i < p->hstate.size() && e == hipSuccess; ++i) e = out(p->hstate[i]);
    if (e == hipSuccess && cand >= 0) e = out(p->hxs[cand & 1]);
    if (e != hipSuccess) return cdlm_fail(p->ctx, CDLM_ERR_HIP, "chunk download: %s", hipGetErrorString(e));
    return CDLM_OK;
}

// One pass over the chunks: in, fn(chunk set), out.  With one window each chunk waits for the
// last; with two, chunk k's upload (copy stream cin) waits for chunk k-2's download from the same
// window (copy stream cout), its kernels (the context stream) for its upload, and its download for
// its kernels, so uploads, kernels and downloads of three consecutive chunks overlap.
template <class F>
cdlm_status stream_pass(cdlm_problem *p, bool x0, int cand, F fn) {
    hipStream_t cs = p->ctx->stream;
    const int K = p->nchunk;
    if (p->nwin < 2) {
        for (int k = 0; k < K; ++k) {
            cdlm_status st = stream_in(p, k, x0, cs);
            if (st == CDLM_OK) st = fn(p->csets[(size_t)k], k);
            if (st == CDLM_OK) st = stream_out(p, k, cand, cs);
            if (st != CDLM_OK) return st;
        }
        return sync_ctx(p->ctx);
    }
    hipEvent_t *ev_in = p->ev.data(), *ev_run = ev_in + K, *ev_out = ev_run + K;
    auto put = [&](int k) -> cdlm_status {
        if (k >= 2) CDLM_HIP(p->ctx, hipStreamWaitEvent(p->cin, ev_out[k - 2], 0));
        cdlm_status st = stream_in(p, k, x0, p->cin);
        if (st != CDLM_OK) return st;
        CDLM_HIP(p->ctx, hipEventRecord(ev_in[k], p->cin));
        return CDLM_OK;
    };
    // the previous pass's kernels and reductions (on cs) come before anything this pass moves
    CDLM_HIP(p->ctx, hipEventRecord(ev_run[0], cs));
    CDLM_HIP(p->ctx, hipStreamWaitEvent(p->cin, ev_run[0], 0));
    cdlm_status st = put(0);
    if (st == CDLM_OK && K > 1) st = put(1);
    for (int k = 0; k < K && st == CDLM_OK; ++k) {
        CDLM_HIP(p->ctx, hipStreamWaitEvent(cs, ev_in[k], 0));
        if ((st = fn(p->csets[(size_t)k], k)) != CDLM_OK) break;
        CDLM_HIP(p->ctx, hipEventRecord(ev_run[k], cs));
        CDLM_HIP(p->ctx, hipStreamWaitEvent(p->cout, ev_run[k], 0));
        if ((st = stream_out(p, k, cand, p->cout)) != CDLM_OK) break;
        CDLM_HIP(p->ctx, hipEventRecord(ev_out[k], p->cout));
        if (k + 2 < K) st = put(k + 2);
    }
    // everything this pass moved is back before the caller reads host memory or runs the next pass
    hipError_t e = hipStreamSynchronize(p->cin);
    if (e == hipSuccess) e = hipStreamSynchronize(p->cout);
    if (st != CDLM_OK) return st;
    if (e != hipSuccess) return cdlm_fail(p->ctx, CDLM_ERR_HIP, "chunk copies: %s", hipGetErrorString(e));
    return sync_ctx(p->ctx);
}

cdlm_status mark_materialised(cdlm_ctx *ctx, const LaunchSet &S) {
    hipLaunchKernelGGL(k_mark_materialised, dim3(1), dim3(64), 0, set_stream(ctx, S), S.tab, S.np);
    return lcheck(ctx, "k_mark_materialised");
}

// Index of the X_best slice buffer k_best writes its candidate to.
cdlm_status cand_of(cdlm_problem *p, int *cand) {
    CdlScal sc;
    cdlm_status st = read_scal(p, &sc);
    if (st == CDLM_OK) *cand = (sc.best_idx & 1) ^ 1;
    return st;
}

// update_best / objective of a streamed problem (decide = 0: the objective only).
cdlm_status streamed_best(cdlm_problem *p, int force, int decide, int slot, Timer *tm) {
    int cand = 0;
    cdlm_status st = cand_of(p, &cand);
    if (st == CDLM_OK)
        st = stream_pass(p, false, cand, [&](const LaunchSet &S, int) { return run_best_seq(p->ctx, S, tm, nullptr); });
    if (st == CDLM_OK) st = run_best_global(p->ctx, p->set, force, decide, slot);
    return st;
}

cdlm_status streamed_reset(cdlm_problem *p) {
    cdlm_status st = run_reset_filt(p->ctx, p->set);
    if (st == CDLM_OK)
        st = stream_pass(p, true, -1, [&](const LaunchSet &S, int) { return run_init_seq(p->ctx, S); });
    if (st == CDLM_OK) st = streamed_best(p, 1, 1, -1, nullptr);
    if (st == CDLM_OK) st = sync_ctx(p->ctx);
    return st;
}

cdlm_status streamed_dict(cdlm_problem *p, Timer *tm) {
    cdlm_status st = stream_pass(p, false, -1, [&](const LaunchSet &S, int) { return run_dict_seq(p->ctx, S, tm); });
    return st != CDLM_OK ? st : run_dict_filt(p->ctx, p->set);
}

// The schedule of cdl_optimize with one pass over the chunks per outer iteration: CSC
// iterations, k_best (X materialised, the candidate slice) and the first dictionary iteration's
// per-sequence part run on a chunk while it is in the window; the objective decision and the
// per-filter part of the dictionary step follow over all chunks, as for a resident problem.
cdlm_status streamed_schedule(cdlm_problem *p, Timer *tm, int iter_num, int csc_iter, int cdl_iter) {
    for (int it = 0; it < iter_num; ++it) {
        int cand = 0;
        cdlm_status st = cand_of(p, &cand);
        if (st == CDLM_OK)
            st = stream_pass(p, false, cand, [&](const LaunchSet &S, int) {
                cdlm_status r = run_csc(p->ctx, S, tm, csc_iter);
                if (r == CDLM_OK) r = run_best_seq(p->ctx, S, tm, nullptr);
                if (r == CDLM_OK && cdl_iter > 0) {
                    r = mark_materialised(p->ctx, S);   // as k_reduce_partials records between the two
                    if (r == CDLM_OK) r = run_dict_seq(p->ctx, S, tm);
                }
                return r;
            });
        if (st == CDLM_OK) st = run_best_global(p->ctx, p->set, 0, 1, it);
        if (st == CDLM_OK && cdl_iter > 0) st = run_dict_filt(p->ctx, p->set);
        for (int i = 1; i < cdl_iter && st == CDLM_OK; ++i) st = streamed_dict(p, tm);
        if (st != CDLM_OK) return st;
    }
    return sync_ctx(p->ctx);
}
}  // namespace

namespace {
cdlm_status problem_create_impl(cdlm_ctx *ctx, const cdlm_params *prm, const float *s, const uint8_t *codes, int32_t L4,
                                int32_t N, const float *D0, const float *X0, uint64_t seed, cdlm_problem **out,
                                void *arena, size_t arena_bytes, uint64_t stream_budget) {
    if (!prm || !out || !prm->filter_len) return cdlm_fail(ctx, CDLM_ERR_ARG, "null argument");
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
    if (!(prm->sherman_csc > 0.f) || !(prm->sherman_cdl > 0.f)) return cdlm_fail(ctx, CDLM_ERR_ARG, "sherman constants must be positive");
    p->gamma = gamma;
    p->Lg = L4 + 2 * gamma;
    const int Lg = p->Lg, M = p->M;
    if (Lg > 2048) return cdlm_fail(ctx, CDLM_ERR_UNSUPPORTED, "padded length %d > 2048", Lg);
    CdlDev &d = p->dv;
    memset(&d, 0, sizeof(d));
    d.L4 = L4; d.N = N; d.M = M; d.Lg = Lg; d.gamma = gamma; d.L = p->L;
    d.nch = pick_nch(Lg, M, N);
    d.n_off = 0; d.Nt = N; d.red_c0 = 0; d.red_cn = d.nch;
    d.trace_cap = kTraceCap;
    d.rho = prm->rho; d.sigma = prm->sigma; d.lambda = prm->lambda;
    d.sh_csc = prm->sherman_csc; d.sh_cdl = prm->sherman_cdl;
    d.seed = seed;
    p->tune = tuning_of(ctx->device, Lg, M, gamma);
    if (!make_fft_plan(Lg, 1, &d.plan, p->tune.order) || !make_fft_plan(Lg / 4, 4, &d.plan4, p->tune.order4))
        return cdlm_fail(ctx, CDLM_ERR_UNSUPPORTED, "no FFT plan for length %d", Lg);

    // host-side masked D0 (P_D .* D, CDL_setup.jl:183)
    const size_t LM = (size_t)Lg * M;
    std::vector<float> D0m(LM, 0.f);
    if (D0) std::memcpy(D0m.data(), D0, LM * sizeof(float));
    else host_rand_probvec(D0m, Lg, M, seed);
    for (int m = 0; m < M; ++m)
        for (int l = p->fl[m]; l < Lg; ++l) D0m[(size_t)m * Lg + l] = 0.f;

    // streamed: the fewest chunks (a divisor of the mean chunks) whose two windows fit the budget
    // (copies overlapping the kernels), else the fewest whose one window does; a problem that fits
    // whole stays resident
    int Nwin = N;
    if (stream_budget && plan_total(buf_plan(L4, N, M, gamma, X0 != nullptr)) > stream_budget) {
        const int nch = d.nch;
        for (int wins = 2; wins >= 1 && !p->nchunk; --wins)
            for (int k = 2; k <= nch && !p->nchunk; ++k) {
                if (nch % k) continue;
                int mx = 1;
                for (int c = 0; c < k; ++c)
                    mx = std::max(mx, (int)((long)((c + 1) * (nch / k)) * N / nch) - (int)((long)(c * (nch / k)) * N / nch));
                if (plan_total(buf_plan(L4, mx, M, gamma, X0 != nullptr, N), wins) <= stream_budget) {
                    p->nchunk = k;
                    p->nwin = wins;
                    Nwin = mx;
                }
            }
        if (!p->nchunk)
            return cdlm_fail(ctx, CDLM_ERR_OOM, "a device window of %d chunks still exceeds the budget of %llu bytes", nch,
                             (unsigned long long)stream_budget);
        p->win = Nwin;
        for (int c = 0; c <= p->nchunk; ++c) p->cn0.push_back((int)((long)(c * (nch / p->nchunk)) * N / nch));
        d.N = Nwin;
    }
    const BufPlan b = buf_plan(L4, Nwin, M, gamma, X0 != nullptr, N);
    cdlm_status st = CDLM_OK;
    const size_t total = plan_total(b, p->nwin);
    if (arena) {
        if (arena_bytes < total)
            return cdlm_fail(ctx, CDLM_ERR_ARG, "arena of %zu bytes is smaller than the problem's %zu", arena_bytes, total);
        p->arena = arena;
        p->arena_bytes = arena_bytes;
        p->own_arena = false;
    } else {
        hipError_t e = hipMalloc(&p->arena, total);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            p->arena = nullptr;
            return cdlm_fail(ctx, CDLM_ERR_OOM, "hipMalloc(%zu bytes) failed: %s", total, hipGetErrorString(e));
        }
        p->arena_bytes = total;
        p->own_arena = true;
    }
    p->total_bytes = total;
    float *s_dev = nullptr, *X0_dev = nullptr, *D0_dev = nullptr;
    int *fl_dev = nullptr;
    float2 *tw_dev = nullptr;
#define A(ptr, bytes) do { st = palloc_t(p.get(), &(ptr), (bytes)); if (st != CDLM_OK) { free_problem(p.get()); return st; } } while (0)
    A(s_dev, b.s);
    A(d.Xt, b.E); A(d.y2, b.E); A(d.y3, b.E3); A(d.g1h, b.Eh);
    if (X0) A(X0_dev, b.Ex0);
    A(d.h, b.LN); A(d.y1, b.LN); A(d.Y1, b.LN); A(d.DX, b.LN); A(d.G2, b.LN); A(d.g2, b.LN); A(d.Td, b.LN);
    A(d.G1b[0], b.LM); A(d.G1b[1], b.LM); A(d.G3, b.LM); A(d.meand, b.LM); A(d.Dbest, b.LM); A(d.gsq, b.LM);
    A(D0_dev, b.LM);
    A(d.G1hb[0], b.LMh); A(d.G1hb[1], b.LMh); A(d.G3h, b.LMh); A(d.d13h, b.LMh); A(d.G1hp, b.LMh); A(d.Fp, b.LMh);
    A(d.cS, b.Lg); A(d.sS, b.Lg);
    A(tw_dev, 2 * b.Lg);
    A(d.xs[0], b.xs); A(d.xs[1], b.xs);
    A(d.y3s, b.y3s);
    A(d.kap, b.kap);
    A(fl_dev, b.M);
    A(d.partials, b.partials);
    A(d.dpart, b.dpart);
    A(d.trace, b.trace);
    A(d.sc, b.sc);
    // the second window's per-sequence buffers (descriptor d2: d with these in place)
    CdlDev d2 = d;
    float *s_dev2 = nullptr, *X0_dev2 = nullptr;
    if (p->nwin == 2) {
        A(s_dev2, b.s);
        A(d2.Xt, b.E); A(d2.y2, b.E); A(d2.y3, b.E3); A(d2.g1h, b.Eh);
        if (X0) A(X0_dev2, b.Ex0);
        A(d2.h, b.LN); A(d2.y1, b.LN); A(d2.Y1, b.LN); A(d2.DX, b.LN); A(d2.G2, b.LN); A(d2.g2, b.LN); A(d2.Td, b.LN);
        A(d2.xs[0], b.xs); A(d2.xs[1], b.xs);
        A(d2.y3s, b.y3s);
        A(d2.kap, b.kap);
    }
#undef A
    d.s = s_dev; d.X0 = X0_dev; d.D0 = D0_dev; d.fl = fl_dev; d.tw = tw_dev;
    d2.s = s_dev2; d2.X0 = X0_dev2; d2.D0 = D0_dev; d2.fl = fl_dev; d2.tw = tw_dev;
    hipStream_t hs = ctx->stream;
    std::vector<float2> tw(Lg);
    for (int t = 0; t < Lg; ++t) {
        const double a = -2.0 * M_PI * (double)t / (double)Lg;
        tw[t] = make_float2((float)cos(a), (float)sin(a));
    }
    hipError_t e = hipSuccess;
    if (p->nchunk) {
        // the per-sequence arrays of all N sequences in pinned host memory (data and initial
        // codes in, state in and out, X_best slices out)
        const size_t Hs = (size_t)half_bins(Lg), nslot = (size_t)Lg / 4;
        auto host = [&](cdlm_problem::HostArr &h, void *dev, void *dev2, size_t per_seq) -> bool {
            h.dev[0] = static_cast<char *>(dev);
            h.dev[1] = static_cast<char *>(dev2);
            h.per_seq = per_seq;
            void *q = nullptr;
            if (hipHostMalloc(&q, per_seq * (size_t)N ? per_seq * (size_t)N : 4) != hipSuccess) {
                (void)hipGetLastError();
                return false;
            }
            p->pinned.push_back(q);
            h.host = static_cast<char *>(q);
            return true;
        };
        const size_t LMb = (size_t)Lg * M * sizeof(float), Lb = (size_t)Lg * sizeof(float);
        struct { void *dev, *dev2; size_t per; } arrs[] = {
            {d.Xt, d2.Xt, LMb}, {d.y2, d2.y2, LMb}, {d.y3, d2.y3, 3 * nslot * M * sizeof(float)},
            {d.g1h, d2.g1h, Hs * M * sizeof(float2)}, {d.y3s, d2.y3s, (size_t)M * (Lg / 8 + 1) * sizeof(float2)},
            {d.kap, d2.kap, Hs * sizeof(float2)}, {d.h, d2.h, Lb}, {d.y1, d2.y1, Lb}, {d.Y1, d2.Y1, Lb},
            {d.DX, d2.DX, Lb}, {d.G2, d2.G2, Lb}, {d.g2, d2.g2, Lb}, {d.Td, d2.Td, Lb}};
        bool ok = host(p->hs, s_dev, s_dev2, (size_t)L4 * sizeof(float)) && (!X0 || host(p->hx0, X0_dev, X0_dev2, LMb)) &&
                  host(p->hxs[0], d.xs[0], d2.xs[0], (size_t)p->L * M * sizeof(float)) &&
                  host(p->hxs[1], d.xs[1], d2.xs[1], (size_t)p->L * M * sizeof(float));
        for (auto &a : arrs) {
            if (!ok) break;
            p->hstate.emplace_back();
            ok = host(p->hstate.back(), a.dev, a.dev2, a.per);
        }
        if (ok && p->nwin == 2) {
            // two copy streams (default priority) and three events per chunk
            ok = hipStreamCreateWithFlags(&p->cin, hipStreamNonBlocking) == hipSuccess &&
                 hipStreamCreateWithFlags(&p->cout, hipStreamNonBlocking) == hipSuccess;
            for (int i = 0; ok && i < 3 * p->nchunk; ++i) {
                hipEvent_t ev = nullptr;
                ok = hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess;
                if (ok) p->ev.push_back(ev);
            }
            if (!ok) {
                (void)hipGetLastError();
                free_problem(p.get());
                return cdlm_fail(ctx, CDLM_ERR_HIP, "copy streams for a streamed problem");
            }
        }
        if (!ok) {
            free_problem(p.get());
            return cdlm_fail(ctx, CDLM_ERR_OOM, "pinned host memory for the state of %d sequences", N);
        }
        std::memcpy(p->hs.host, s, (size_t)L4 * N * sizeof(float));
        if (X0) std::memcpy(p->hx0.host, X0, LMb * (size_t)N);
    } else if (s) {
        e = hipMemcpyAsync(s_dev, s, b.s, hipMemcpyHostToDevice, hs);
    } else {
        // base codes: N*L bytes over the bus, the one-hot expansion on the device
        uint8_t *cd = nullptr;
        const size_t nb = (size_t)(L4 / 4) * N;
        e = hipMalloc(&cd, nb);
        if (e == hipSuccess) e = hipMemcpyAsync(cd, codes, nb, hipMemcpyHostToDevice, hs);
        if (e == hipSuccess) {
            long grid = ((long)nb + 255) / 256;
            if (grid > 65536) grid = 65536;
            hipLaunchKernelGGL(k_bases_to_onehot, dim3((unsigned)grid), dim3(256), 0, hs, cd, L4 / 4, (long)N, s_dev);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(hs);
        if (cd) (void)hipFree(cd);
    }
    if (e == hipSuccess && X0 && !p->nchunk) e = hipMemcpyAsync(X0_dev, X0, b.Ex0, hipMemcpyHostToDevice, hs);
    if (e == hipSuccess) e = hipMemcpyAsync(D0_dev, D0m.data(), b.LM, hipMemcpyHostToDevice, hs);
    if (e == hipSuccess) e = hipMemcpyAsync(fl_dev, p->fl.data(), b.M, hipMemcpyHostToDevice, hs);
    if (e == hipSuccess) e = hipMemcpyAsync(tw_dev, tw.data(), Lg * sizeof(float2), hipMemcpyHostToDevice, hs);
    if (e == hipSuccess) e = hipStreamSynchronize(hs);
    if (e != hipSuccess) {
        free_problem(p.get());
        return cdlm_fail(ctx, CDLM_ERR_HIP, "upload: %s", hipGetErrorString(e));
    }
    cdlm_problem *pp = p.get();
    st = set_build(ctx, &pp, 1, &p->set);
    if (st == CDLM_OK && p->nchunk) {
        // chunk k's launch set: its sequences, their first index and its dictionary-mean chunks
        p->csets.resize((size_t)p->nchunk);
        const int per = d.nch / p->nchunk;
        const CdlDev d1 = d;
        for (int k = 0; k < p->nchunk && st == CDLM_OK; ++k) {
            d = (k % p->nwin) ? d2 : d1;   // chunk k's window
            d.N = p->cn0[k + 1] - p->cn0[k];
            d.n_off = p->cn0[k];
            d.red_c0 = k * per;
            d.red_cn = per;
            st = set_build(ctx, &pp, 1, &p->csets[(size_t)k]);
        }
        d = d1;
    }
    if (st == CDLM_OK) st = p->nchunk ? streamed_reset(p.get()) : run_reset(ctx, p->set);
    if (st == CDLM_OK) st = sync_ctx(ctx);
    if (st != CDLM_OK) {
        free_problem(p.get());
        return st;
    }
    *out = p.release();
    return CDLM_OK;
}
}  // namespace

extern "C" cdlm_status cdlm_problem_destroy(cdlm_problem *p) {
    if (!p) return CDLM_OK;
    if (p->batches > 0) return cdlm_fail(p->ctx, CDLM_ERR_STATE, "problem is still part of %d batch(es)", p->batches);
    (void)hipStreamSynchronize(p->ctx->stream);
    free_problem(p);
    delete p;
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_problem_reset(cdlm_problem *p) {
    if (!p) return CDLM_ERR_ARG;
    if (p->nchunk) return streamed_reset(p);
    cdlm_status st = run_reset(p->ctx, p->set);
    if (st != CDLM_OK) return st;
    return sync_ctx(p->ctx);
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
    if (p->nchunk)
        return stream_pass(p, false, -1, [&](const LaunchSet &S, int) { return run_csc(p->ctx, S, &p->timer, num_iter); });
    cdlm_status st = run_csc(p->ctx, p->set, &p->timer, num_iter);
    if (st != CDLM_OK) return st;
    return sync_ctx(p->ctx);
}

extern "C" cdlm_status cdlm_dict_update(cdlm_problem *p, int32_t num_iter) {
    if (!p) return CDLM_ERR_ARG;
    if (num_iter < 0) return cdlm_fail(p->ctx, CDLM_ERR_ARG, "num_iter < 0");
    for (int i = 0; i < num_iter; ++i) {
        cdlm_status st = p->nchunk ? streamed_dict(p, &p->timer) : run_dict(p->ctx, p->set, &p->timer);
        if (st != CDLM_OK) return st;
    }
    return sync_ctx(p->ctx);
}

extern "C" cdlm_status cdlm_update_best(cdlm_problem *p, double *obj, int32_t *improved) {
    if (!p) return CDLM_ERR_ARG;
    cdlm_status st = p->nchunk ? streamed_best(p, 0, 1, -1, &p->timer) : run_best(p->ctx, p->set, &p->timer, 0, 1, -1, nullptr);
    if (st != CDLM_OK) return st;
    CdlScal sc;
    if ((st = read_scal(p, &sc)) != CDLM_OK) return st;
    if (obj) *obj = sc.red[5];
    if (improved) *improved = sc.improved;
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_objective(cdlm_problem *p, double *obj) {
    if (!p || !obj) return CDLM_ERR_ARG;
    // the candidate X_best slice buffer is scratch here; no decision is taken
    cdlm_status st = p->nchunk ? streamed_best(p, 0, 0, -1, nullptr) : run_best(p->ctx, p->set, nullptr, 0, 0, -1, nullptr);
    if (st != CDLM_OK) return st;
    CdlScal sc;
    if ((st = read_scal(p, &sc)) != CDLM_OK) return st;
    *obj = sc.red[5];
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_optimize(cdlm_problem *p, int32_t iter_num, int32_t csc_iter, int32_t cdl_iter,
                                     double *obj_trace) {
    if (!p) return CDLM_ERR_ARG;
    if (iter_num < 0 || csc_iter < 0 || cdl_iter < 0) return cdlm_fail(p->ctx, CDLM_ERR_ARG, "negative iteration count");
    if (iter_num > kTraceCap) return cdlm_fail(p->ctx, CDLM_ERR_ARG, "iter_num > %d", kTraceCap);
    cdlm_status st = p->nchunk ? streamed_schedule(p, &p->timer, iter_num, csc_iter, cdl_iter)
                               : run_schedule(p->ctx, p->set, &p->timer, iter_num, csc_iter, cdl_iter);
    if (st != CDLM_OK) return st;
    if (obj_trace && iter_num > 0)
        CDLM_HIP(p->ctx, hipMemcpyAsync(obj_trace, p->dv.trace, iter_num * sizeof(double), hipMemcpyDeviceToHost, p->ctx->stream));
    return sync_ctx(p->ctx);
}

extern "C" cdlm_status cdlm_code_thresh(cdlm_problem *p, double mult, double *thresh, int64_t *count) {
    if (!p) return CDLM_ERR_ARG;
    CdlScal sc;
    cdlm_status st = read_scal(p, &sc);
    if (st != CDLM_OK) return st;
    if (count) *count = (int64_t)sc.best_stats[0];
    if (thresh) *thresh = thresh_of(sc.best_stats[0], sc.best_stats[1], sc.best_stats[2], mult);
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_obj_best(cdlm_problem *p, float *obj_best) {
    if (!p || !obj_best) return CDLM_ERR_ARG;
    CdlScal sc;
    cdlm_status st = read_scal(p, &sc);
    if (st != CDLM_OK) return st;
    *obj_best = sc.obj_best;
    return CDLM_OK;
}

namespace {
// cdlm_get of a streamed problem: per-sequence arrays from the host state, d, g1, g3 and X through
// the device window chunk by chunk, the per-filter arrays from the device.
cdlm_status streamed_get(cdlm_problem *p, int32_t field, void *dst, size_t bytes) {
    const size_t N = (size_t)p->N, E = (size_t)p->Lg * p->M * N, LN = (size_t)p->Lg * N;
    hipStream_t st = p->ctx->stream;
    auto host_copy = [&](int idx, size_t cnt) -> cdlm_status {
        if (bytes != cnt * sizeof(float))
            return cdlm_fail(p->ctx, CDLM_ERR_ARG, "field %d needs %zu bytes, got %zu", field, cnt * sizeof(float), bytes);
        std::memcpy(dst, p->hstate[(size_t)idx].host, bytes);
        return CDLM_OK;
    };
    switch (field) {
    case CDLM_F_y2: return host_copy(1, E);
    case CDLM_F_y1: return host_copy(7, LN);
    case CDLM_F_Y1: return host_copy(8, LN);
    case CDLM_F_DX: return host_copy(9, LN);
    case CDLM_F_G2: return host_copy(10, LN);
    case CDLM_F_g2: return host_copy(11, LN);
    case CDLM_F_TD: return host_copy(12, LN);
    case CDLM_F_y3: {
        if (bytes != E * sizeof(float)) return cdlm_fail(p->ctx, CDLM_ERR_ARG, "field y3 needs %zu bytes", E * sizeof(float));
        const size_t nslot = (size_t)p->Lg / 4;
        const float *packed = reinterpret_cast<const float *>(p->hstate[2].host);
        float *o = static_cast<float *>(dst);
        for (size_t c = 0; c < (size_t)p->M * N; ++c)
            for (size_t sl = 0; sl < nslot; ++sl) {
                const float *q = &packed[(c * nslot + sl) * 3];
                float *w = o + c * p->Lg + 4 * sl;
                w[0] = 0.f; w[1] = q[0]; w[2] = q[1]; w[3] = q[2];
            }
        return CDLM_OK;
    }
    case CDLM_F_XBEST_SLICE: {
        const size_t cnt = (size_t)p->L * p->M * N;
        if (bytes != cnt * sizeof(float)) return cdlm_fail(p->ctx, CDLM_ERR_ARG, "field %d needs %zu bytes", field, cnt * 4);
        CdlScal sc;
        cdlm_status s = read_scal(p, &sc);
        if (s != CDLM_OK) return s;
        std::memcpy(dst, p->hxs[sc.best_idx & 1].host, bytes);
        return CDLM_OK;
    }
    case CDLM_F_X:
    case CDLM_F_D:
    case CDLM_F_g1:
    case CDLM_F_g3: {
        if (bytes != E * sizeof(float)) return cdlm_fail(p->ctx, CDLM_ERR_ARG, "field %d needs %zu bytes", field, E * 4);
        const size_t per = (size_t)p->Lg * p->M;
        float *tmp = nullptr;
        if (hipMalloc(&tmp, per * (size_t)p->win * sizeof(float)) != hipSuccess) {
            (void)hipGetLastError();
            return cdlm_fail(p->ctx, CDLM_ERR_OOM, "no memory for field %d", field);
        }
        char *o = static_cast<char *>(dst);
        cdlm_status s = CDLM_OK;
        const int what = field == CDLM_F_D ? 0 : field == CDLM_F_g1 ? 1 : 2;
        for (int k = 0; k < p->nchunk && s == CDLM_OK; ++k) {
            const size_t a = (size_t)p->cn0[k], n = (size_t)(p->cn0[k + 1] - p->cn0[k]);
            s = stream_in(p, k, false, st);
            if (s != CDLM_OK) break;
            if (field == CDLM_F_X) {
                // the candidate slice buffer is scratch here; X is written back materialised
                s = run_best_seq(p->ctx, p->csets[(size_t)k], nullptr, tmp);
                if (s == CDLM_OK) s = stream_out(p, k, -1, st);
            } else if (n > 0) {
                hipLaunchKernelGGL(k_dict_to_spatial, dim3((unsigned)(n * p->M)), dim3(kThreads), lds_bytes_half(p->Lg), st,
                                   p->csets[(size_t)k].tab, what, tmp);
                s = lcheck(p->ctx, "k_dict_to_spatial");
            }
            if (s == CDLM_OK && n > 0 &&
                hipMemcpyAsync(o + a * per * sizeof(float), tmp, n * per * sizeof(float), hipMemcpyDeviceToHost, st) !=
                    hipSuccess)
                s = cdlm_fail(p->ctx, CDLM_ERR_HIP, "copy of field %d failed", field);
        }
        if (s == CDLM_OK && field == CDLM_F_X) s = run_best_global(p->ctx, p->set, 0, 0, -1);
        if (s == CDLM_OK) s = sync_ctx(p->ctx);
        (void)hipStreamSynchronize(st);
        (void)hipFree(tmp);
        return s;
    }
    default:
        break;
    }
    return CDLM_ERR_STATE;   // per-filter fields: read from the device as for a resident problem
}
}  // namespace

extern "C" cdlm_status cdlm_get(cdlm_problem *p, int32_t field, void *dst, size_t bytes) {
    if (!p || !dst) return CDLM_ERR_ARG;
    if (p->nchunk) {
        cdlm_status s = streamed_get(p, field, dst, bytes);
        if (s != CDLM_ERR_STATE) return s;
    }
    const size_t E = (size_t)p->Lg * p->M * p->N, LN = (size_t)p->Lg * p->N, LM = (size_t)p->Lg * p->M;
    const CdlDev &d = p->dv;
    const void *src = nullptr;
    size_t cnt = 0;
    hipStream_t st = p->ctx->stream;
    CdlScal sc;
    if (field == CDLM_F_G1 || field == CDLM_F_XBEST_SLICE) {
        cdlm_status s = read_scal(p, &sc);
        if (s != CDLM_OK) return s;
    }
    switch (field) {
    case CDLM_F_X: {
        if (bytes != E * sizeof(float)) return cdlm_fail(p->ctx, CDLM_ERR_ARG, "field X needs %zu bytes", E * sizeof(float));
        float *tmp = nullptr;
        hipError_t e = hipMalloc(&tmp, E * sizeof(float));
        if (e != hipSuccess) { (void)hipGetLastError(); return cdlm_fail(p->ctx, CDLM_ERR_OOM, "no memory for X"); }
        // the candidate slice buffer is rewritten with the same values; partials are scratch
        cdlm_status s = run_best(p->ctx, p->set, nullptr, 0, 0, -1, tmp);
        hipError_t e2 = hipMemcpyAsync(dst, tmp, bytes, hipMemcpyDeviceToHost, st);
        hipError_t e3 = hipStreamSynchronize(st);
        (void)hipFree(tmp);
        if (s != CDLM_OK) return s;
        if (e2 != hipSuccess || e3 != hipSuccess) return cdlm_fail(p->ctx, CDLM_ERR_HIP, "copy of X failed");
        return CDLM_OK;
    }
    case CDLM_F_DX: src = d.DX; cnt = LN; break;
    case CDLM_F_D:
    case CDLM_F_g1:
    case CDLM_F_g3: {
        // held as half spectra: back to the spatial domain on the device first
        if (bytes != E * sizeof(float)) return cdlm_fail(p->ctx, CDLM_ERR_ARG, "field %d needs %zu bytes", field, E * sizeof(float));
        const int what = field == CDLM_F_D ? 0 : field == CDLM_F_g1 ? 1 : 2;
        float *tmp = nullptr;
        hipError_t e = hipMalloc(&tmp, E * sizeof(float));
        if (e != hipSuccess) { (void)hipGetLastError(); return cdlm_fail(p->ctx, CDLM_ERR_OOM, "no memory for field %d", field); }
        hipLaunchKernelGGL(k_dict_to_spatial, dim3((unsigned)((size_t)p->M * p->N)), dim3(kThreads), lds_bytes_half(p->Lg),
                           st, p->set.tab, what, tmp);
        cdlm_status s = lcheck(p->ctx, "k_dict_to_spatial");
        hipError_t e2 = s == CDLM_OK ? hipMemcpyAsync(dst, tmp, bytes, hipMemcpyDeviceToHost, st) : hipSuccess;
        hipError_t e3 = hipStreamSynchronize(st);
        (void)hipFree(tmp);
        if (s != CDLM_OK) return s;
        if (e2 != hipSuccess || e3 != hipSuccess) return cdlm_fail(p->ctx, CDLM_ERR_HIP, "copy of field %d failed", field);
        return CDLM_OK;
    }
    case CDLM_F_G1: src = d.G1b[sc.g1cur & 1]; cnt = LM; break;
    case CDLM_F_G2: src = d.G2; cnt = LN; break;
    case CDLM_F_G3: src = d.G3; cnt = LM; break;
    case CDLM_F_XBEST_SLICE: src = d.xs[sc.best_idx & 1]; cnt = (size_t)p->L * p->M * p->N; break;
    case CDLM_F_DBEST: src = d.Dbest; cnt = LM; break;
    case CDLM_F_Y1: src = d.Y1; cnt = LN; break;
    case CDLM_F_y1: src = d.y1; cnt = LN; break;
    case CDLM_F_y2: src = d.y2; cnt = E; break;
    case CDLM_F_y3: {
        // stored without the stride-4 support samples (identically zero); expand
        if (bytes != E * sizeof(float)) return cdlm_fail(p->ctx, CDLM_ERR_ARG, "field y3 needs %zu bytes", E * sizeof(float));
        const size_t nslot = (size_t)p->Lg / 4;
        std::vector<float> packed(3 * nslot * p->M * (size_t)p->N);
        CDLM_HIP(p->ctx, hipMemcpyAsync(packed.data(), d.y3, packed.size() * sizeof(float), hipMemcpyDeviceToHost, st));
        cdlm_status sy = sync_ctx(p->ctx);
        if (sy != CDLM_OK) return sy;
        float *o = static_cast<float *>(dst);
        for (size_t c = 0; c < (size_t)p->M * p->N; ++c)
            for (size_t sl = 0; sl < nslot; ++sl) {
                const float *q = &packed[(c * nslot + sl) * 3];
                float *w = o + c * p->Lg + 4 * sl;
                w[0] = 0.f; w[1] = q[0]; w[2] = q[1]; w[3] = q[2];
            }
        return CDLM_OK;
    }
    case CDLM_F_g2: src = d.g2; cnt = LN; break;
    case CDLM_F_TD: src = d.Td; cnt = LN; break;
    default: return cdlm_fail(p->ctx, CDLM_ERR_ARG, "unknown field %d", field);
    }
    if (bytes != cnt * sizeof(float))
        return cdlm_fail(p->ctx, CDLM_ERR_ARG, "field %d needs %zu bytes, got %zu", field, cnt * sizeof(float), bytes);
    CDLM_HIP(p->ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
    return sync_ctx(p->ctx);
}

cdlm_status cdlm_problem_xbest_device(cdlm_problem *p, const float **slice, int *L, int *M, int *N) {
    if (p->nchunk)
        return cdlm_fail(p->ctx, CDLM_ERR_UNSUPPORTED, "a streamed problem keeps its X_best slices in host memory");
    CdlScal sc;
    cdlm_status st = read_scal(p, &sc);
    if (st != CDLM_OK) return st;
    *slice = p->dv.xs[sc.best_idx & 1];
    *L = p->L;
    *M = p->M;
    *N = p->N;
    return CDLM_OK;
}

namespace {
cdlm_status timing_read(cdlm_ctx *ctx, Timer &t, int kernel, double *total_ms, int64_t *launches) {
    CDLM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double acc = 0.0;
    int64_t cnt = 0;
    for (auto &e : t.ev[kernel]) {
        float ms = 0.f;
        CDLM_HIP(ctx, hipEventElapsedTime(&ms, e.a, e.b));
        acc += ms;
        cnt += e.count;
    }
    if (total_ms) *total_ms = acc;
    if (launches) *launches = cnt;
    return CDLM_OK;
}
}  // namespace

extern "C" cdlm_status cdlm_timing_enable(cdlm_problem *p, int32_t enable) {
    if (!p) return CDLM_ERR_ARG;
    (void)hipStreamSynchronize(p->ctx->stream);
    p->timer.clear();
    p->timer.on = enable != 0;
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_timing_read(cdlm_problem *p, int32_t kernel, double *total_ms, int64_t *launches) {
    if (!p || kernel < 0 || kernel > 3) return CDLM_ERR_ARG;
    return timing_read(p->ctx, p->timer, kernel, total_ms, launches);
}

// ============================================================================ C ABI: batch
namespace {

// Shape class of a problem run alone: members of one class share a launch set without
// lowering each other's occupancy.
struct ShapeClass {
    int nsl, B, taps, waves, mfma, occ_csc, occ_dict;
    bool operator<(const ShapeClass &o) const {
        return std::tie(nsl, B, taps, waves, mfma, occ_csc, occ_dict) <
               std::tie(o.nsl, o.B, o.taps, o.waves, o.mfma, o.occ_csc, o.occ_dict);
    }
};

ShapeClass shape_class(const cdlm_problem *p) {
    ShapeClass c;
    c.nsl = std::max((p->Lg / 4 <= 256) ? 4 : 8, p->tune.nsl);
    c.B = pick_dict_batch(p->Lg, p->gamma);
    if ((p->tune.B == 2 || p->tune.B == 4 || p->tune.B == 8) &&
        lds_bytes_dict(p->Lg, p->gamma, p->tune.B) <= kLdsLimit - kThreads * sizeof(double))
        c.B = p->tune.B;
    c.taps = p->gamma;
    pick_csc(p->Lg, p->gamma, c.nsl, p->tune.mfma, p->tune.waves, &c.mfma, &c.waves);
    if (c.mfma == 2)
        c.occ_csc = std::min<int>(c.waves == 8 ? 2 : 1, (int)(kLdsLimit / lds_bytes_csctz(p->Lg, p->gamma, c.waves)));
    else
        c.occ_csc = std::min<int>(c.waves == 16 ? 1 : c.waves == 8 ? 2 : (c.nsl == 4 && c.taps <= 20 ? 3 : 2),
                                  (int)(kLdsLimit / lds_bytes_csc(p->Lg, p->gamma, c.waves)));
    c.occ_dict = c.B > 0 ? dict_wg_per_cu(p->Lg, p->gamma, c.B) : 0;
    return c;
}

void drop_graph(cdlm_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
    for (int &k : b->key) k = -1;
}

void release_batch(cdlm_batch *b) {
    drop_graph(b);
    b->timer.clear();
    if (b->stage) (void)hipFree(b->stage);
    b->stage = nullptr;
    for (auto &S : b->sets) S.release();
    b->sets.clear();
    for (auto s : b->streams) (void)hipStreamDestroy(s);
    b->streams.clear();
    for (auto e : b->joins) (void)hipEventDestroy(e);
    b->joins.clear();
    if (b->fork) (void)hipEventDestroy(b->fork);
    b->fork = nullptr;
}

cdlm_status ensure_stage(cdlm_batch *b, int cnt) {
    const size_t need = b->probs.size() * (size_t)(cnt + kGatherExtra);
    if (need > b->stage_cap) {
        drop_graph(b);  // a cached graph holds the old pointer
        if (b->stage) (void)hipFree(b->stage);
        b->stage = nullptr;
        b->stage_cap = 0;
        hipError_t e = hipMalloc(&b->stage, need * sizeof(double));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            b->stage = nullptr;
            return cdlm_fail(b->ctx, CDLM_ERR_OOM, "hipMalloc of the gather buffer failed");
        }
        b->stage_cap = need;
    }
    b->host.resize(need);
    return CDLM_OK;
}

// The device work of one class: [reset], schedule, gather of cnt trace entries.
cdlm_status class_enqueue(cdlm_batch *b, const LaunchSet &S, int reset, int iter_num, int csc_iter, int cdl_iter, int cnt,
                          Timer *tm) {
    cdlm_status st;
    if (reset && (st = run_reset(b->ctx, S)) != CDLM_OK) return st;
    if ((st = run_schedule(b->ctx, S, tm, iter_num, csc_iter, cdl_iter)) != CDLM_OK) return st;
    hipLaunchKernelGGL(k_gather, dim3(S.np), dim3(64), 0, set_stream(b->ctx, S), S.tab, S.slot, S.np, cnt, b->stage);
    return lcheck(b->ctx, "k_gather");
}

// Eager: classes one after another on the context stream.  Captured: class 0 on the
// context stream, the others on their own streams forked from and joined back into it.
cdlm_status batch_enqueue(cdlm_batch *b, bool concurrent, int reset, int iter_num, int csc_iter, int cdl_iter, int cnt,
                          Timer *tm) {
    cdlm_ctx *ctx = b->ctx;
    const int nc = (int)b->sets.size();
    if (concurrent && nc > 1) CDLM_HIP(ctx, hipEventRecord(b->fork, ctx->stream));
    for (int c = 0; c < nc; ++c) {
        LaunchSet &S = b->sets[c];
        S.stream = (concurrent && c > 0) ? b->streams[c - 1] : nullptr;
        if (S.stream) CDLM_HIP(ctx, hipStreamWaitEvent(S.stream, b->fork, 0));
        cdlm_status st = class_enqueue(b, S, reset, iter_num, csc_iter, cdl_iter, cnt, tm);
        if (S.stream) {
            CDLM_HIP(ctx, hipEventRecord(b->joins[c - 1], S.stream));
            CDLM_HIP(ctx, hipStreamWaitEvent(ctx->stream, b->joins[c - 1], 0));
        }
        S.stream = nullptr;
        if (st != CDLM_OK) return st;
    }
    return CDLM_OK;
}

// Runs the batch work, from the cached graph when graphs are on and timing is off.
cdlm_status batch_run(cdlm_batch *b, int reset, int iter_num, int csc_iter, int cdl_iter) {
    const int cnt = iter_num;
    cdlm_status st = ensure_stage(b, cnt);
    if (st != CDLM_OK) return st;
    hipStream_t s = b->ctx->stream;
    if (!b->use_graph || b->timer.on) {
        if ((st = batch_enqueue(b, false, reset, iter_num, csc_iter, cdl_iter, cnt, &b->timer)) != CDLM_OK) return st;
    } else {
        const int key[5] = {reset, iter_num, csc_iter, cdl_iter, cnt};
        if (!b->exec || memcmp(key, b->key, sizeof(key)) != 0) {
            drop_graph(b);
            hipGraph_t g = nullptr;
            CDLM_HIP(b->ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            cdlm_status cst = batch_enqueue(b, true, reset, iter_num, csc_iter, cdl_iter, cnt, nullptr);
            hipError_t e = hipStreamEndCapture(s, &g);
            if (cst != CDLM_OK) {
                if (g) (void)hipGraphDestroy(g);
                return cst;
            }
            if (e != hipSuccess) return cdlm_fail(b->ctx, CDLM_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
            e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                b->exec = nullptr;
                return cdlm_fail(b->ctx, CDLM_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
            }
            memcpy(b->key, key, sizeof(key));
        }
        CDLM_HIP(b->ctx, hipGraphLaunch(b->exec, s));
    }
    CDLM_HIP(b->ctx, hipMemcpyAsync(b->host.data(), b->stage, b->probs.size() * (cnt + kGatherExtra) * sizeof(double),
                                    hipMemcpyDeviceToHost, s));
    return sync_ctx(b->ctx);
}

void batch_outputs(cdlm_batch *b, int cnt, double mult, double *obj_traces, double *thresh, int64_t *counts) {
    for (size_t i = 0; i < b->probs.size(); ++i) {
        const double *o = b->host.data() + i * (cnt + kGatherExtra);
        if (obj_traces && cnt > 0) memcpy(obj_traces + i * cnt, o, cnt * sizeof(double));
        if (thresh) thresh[i] = thresh_of(o[cnt + 1], o[cnt + 2], o[cnt + 3], mult);
        if (counts) counts[i] = (int64_t)o[cnt + 1];
    }
}

}  // namespace

extern "C" cdlm_status cdlm_batch_create(cdlm_ctx *ctx, cdlm_problem *const *problems, int32_t count, cdlm_batch **out) {
    if (!ctx) return CDLM_ERR_ARG;
    if (!problems || !out || count <= 0) return cdlm_fail(ctx, CDLM_ERR_ARG, "batch needs at least one problem");
    *out = nullptr;
    for (int i = 0; i < count; ++i) {
        if (!problems[i]) return cdlm_fail(ctx, CDLM_ERR_ARG, "null problem %d", i);
        if (problems[i]->ctx != ctx) return cdlm_fail(ctx, CDLM_ERR_ARG, "problem %d belongs to another context", i);
        if (problems[i]->nchunk)
            return cdlm_fail(ctx, CDLM_ERR_UNSUPPORTED, "problem %d is streamed through a device window", i);
        for (int j = 0; j < i; ++j)
            if (problems[j] == problems[i]) return cdlm_fail(ctx, CDLM_ERR_ARG, "problem %d appears twice", i);
    }
    std::unique_ptr<cdlm_batch> b(new cdlm_batch());
    b->ctx = ctx;
    b->probs.assign(problems, problems + count);
    std::map<ShapeClass, std::vector<int>> classes;
    for (int i = 0; i < count; ++i) classes[shape_class(problems[i])].push_back(i);
    cdlm_status st = CDLM_OK;
    for (auto &kv : classes) {
        std::vector<cdlm_problem *> ps;
        for (int i : kv.second) ps.push_back(problems[i]);
        LaunchSet S;
        if ((st = set_build(ctx, ps.data(), (int)ps.size(), &S, kv.second.data())) != CDLM_OK) {
            release_batch(b.get());
            return st;
        }
        b->sets.push_back(S);
    }
    hipError_t e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    for (size_t c = 1; c < b->sets.size() && e == hipSuccess; ++c) {
        hipStream_t s = nullptr;
        hipEvent_t j = nullptr;
        e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        if (e == hipSuccess) b->streams.push_back(s);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&j, hipEventDisableTiming);
        if (e == hipSuccess) b->joins.push_back(j);
    }
    if (e != hipSuccess) {
        release_batch(b.get());
        return cdlm_fail(ctx, CDLM_ERR_HIP, "batch streams: %s", hipGetErrorString(e));
    }
    for (auto *p : b->probs) ++p->batches;
    *out = b.release();
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_batch_destroy(cdlm_batch *b) {
    if (!b) return CDLM_OK;
    (void)hipStreamSynchronize(b->ctx->stream);
    release_batch(b);
    for (auto *p : b->probs) --p->batches;
    delete b;
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_batch_set_graph(cdlm_batch *b, int32_t enable) {
    if (!b) return CDLM_ERR_ARG;
    b->use_graph = enable != 0;
    if (!b->use_graph) drop_graph(b);
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_batch_classes(const cdlm_batch *b, int32_t *count) {
    if (!b || !count) return CDLM_ERR_ARG;
    *count = (int32_t)b->sets.size();
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_batch_reset(cdlm_batch *b) {
    if (!b) return CDLM_ERR_ARG;
    for (auto &S : b->sets) {
        cdlm_status st = run_reset(b->ctx, S);
        if (st != CDLM_OK) return st;
    }
    return sync_ctx(b->ctx);
}

extern "C" cdlm_status cdlm_batch_optimize(cdlm_batch *b, int32_t iter_num, int32_t csc_iter, int32_t cdl_iter,
                                           double *obj_traces) {
    if (!b) return CDLM_ERR_ARG;
    if (iter_num < 0 || csc_iter < 0 || cdl_iter < 0) return cdlm_fail(b->ctx, CDLM_ERR_ARG, "negative iteration count");
    if (iter_num > kTraceCap) return cdlm_fail(b->ctx, CDLM_ERR_ARG, "iter_num > %d", kTraceCap);
    cdlm_status st = batch_run(b, 0, iter_num, csc_iter, cdl_iter);
    if (st != CDLM_OK) return st;
    batch_outputs(b, iter_num, 1.0, obj_traces, nullptr, nullptr);
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_batch_train(cdlm_batch *b, int32_t iter_num, int32_t csc_iter, int32_t cdl_iter, double mult,
                                        double *obj_traces, double *thresh, int64_t *counts) {
    if (!b) return CDLM_ERR_ARG;
    if (iter_num < 0 || csc_iter < 0 || cdl_iter < 0) return cdlm_fail(b->ctx, CDLM_ERR_ARG, "negative iteration count");
    if (iter_num > kTraceCap) return cdlm_fail(b->ctx, CDLM_ERR_ARG, "iter_num > %d", kTraceCap);
    cdlm_status st = batch_run(b, 1, iter_num, csc_iter, cdl_iter);
    if (st != CDLM_OK) return st;
    batch_outputs(b, iter_num, mult, obj_traces, thresh, counts);
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_batch_timing_enable(cdlm_batch *b, int32_t enable) {
    if (!b) return CDLM_ERR_ARG;
    (void)hipStreamSynchronize(b->ctx->stream);
    b->timer.clear();
    b->timer.on = enable != 0;
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_batch_timing_read(cdlm_batch *b, int32_t kernel, double *total_ms, int64_t *launches) {
    if (!b || kernel < 0 || kernel > 3) return CDLM_ERR_ARG;
    return timing_read(b->ctx, b->timer, kernel, total_ms, launches);
}

// ============================================================================ C ABI: FFT self-test
extern "C" cdlm_status cdlm_selftest_fft(cdlm_ctx *ctx, int32_t n, int32_t twstride, int32_t nb, int32_t inverse,
                                         int32_t order, const float *in, float *out) {
    if (!ctx) return CDLM_ERR_ARG;
    if (!in || !out || n < 1 || twstride < 1 || nb < 1 || order < 0 || order >= kFftOrders)
        return cdlm_fail(ctx, CDLM_ERR_ARG, "bad FFT self-test arguments");
    FftPlan plan;
    if (!make_fft_plan(n, twstride, &plan, order)) return cdlm_fail(ctx, CDLM_ERR_UNSUPPORTED, "no plan for %d", n);
    const int N = n * twstride;
    const size_t lds = sizeof(float2) * ((size_t)N + 2 * (size_t)nb * n);
    if (lds > kLdsLimit) return cdlm_fail(ctx, CDLM_ERR_UNSUPPORTED, "FFT self-test needs %zu B of LDS", lds);
    cdlm_status st = ensure_lds(ctx, (const void *)k_fft_selftest, lds);
    if (st != CDLM_OK) return st;
    std::vector<float2> tw(N);
    for (int t = 0; t < N; ++t) {
        const double ang = -2.0 * M_PI * (double)t / (double)N;
        tw[t] = make_float2((float)cos(ang), (float)sin(ang));
    }
    float2 *d = nullptr;
    const size_t cnt = (size_t)nb * n;
    CDLM_HIP(ctx, hipMalloc(&d, sizeof(float2) * (2 * cnt + N)));
    hipError_t e = hipMemcpyAsync(d, in, sizeof(float2) * cnt, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d + 2 * cnt, tw.data(), sizeof(float2) * N, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_fft_selftest, dim3(1), dim3(kThreads), lds, ctx->stream, d, d + cnt, d + 2 * cnt, plan, nb,
                           inverse ? 1 : 0);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, d + cnt, sizeof(float2) * cnt, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return cdlm_fail(ctx, CDLM_ERR_HIP, "FFT self-test: %s", hipGetErrorString(e));
    return CDLM_OK;
}

// ============================================================================ C ABI: tuning
namespace {
cdlm_status tune_args(cdlm_ctx *ctx, int32_t L4, int32_t M, int32_t gamma) {
    if (L4 <= 0 || L4 % 4 || M <= 0 || gamma <= 0 || gamma % 4 || gamma > kMaxTaps || L4 + 2 * gamma > 2048)
        return cdlm_fail(ctx, CDLM_ERR_ARG, "bad tuning shape (L4=%d, M=%d, gamma=%d)", L4, M, gamma);
    return CDLM_OK;
}
}  // namespace

extern "C" cdlm_status cdlm_tuning_set(cdlm_ctx *ctx, int32_t L4, int32_t M, int32_t gamma, const cdlm_tuning *t) {
    if (!ctx) return CDLM_ERR_ARG;
    cdlm_status st = tune_args(ctx, L4, M, gamma);
    if (st != CDLM_OK) return st;
    const auto key = std::make_tuple(ctx->device, L4 + 2 * gamma, (int)M, (int)gamma);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (!t) {
        g_tune.erase(key);
        return CDLM_OK;
    }
    if (t->fft_order < 0 || t->fft_order >= kFftOrders || t->fft4_order < 0 || t->fft4_order >= kFftOrders ||
        (t->dict_batch != 0 && t->dict_batch != 2 && t->dict_batch != 4 && t->dict_batch != 8) ||
        (t->csc_slots != 0 && t->csc_slots != 4 && t->csc_slots != 8) ||
        (t->csc_waves != 0 && t->csc_waves != 4 && t->csc_waves != 8 && t->csc_waves != 16) ||
        t->csc_mfma < 0 || t->csc_mfma > 2 || (t->csc_mfma == 2 && t->csc_waves == 4))
        return cdlm_fail(ctx, CDLM_ERR_ARG, "bad tuning values");
    Tuning tu;
    tu.order = t->fft_order;
    tu.order4 = t->fft4_order;
    tu.B = t->dict_batch;
    tu.nsl = t->csc_slots;
    tu.waves = t->csc_waves;
    tu.mfma = t->csc_mfma;
    g_tune[key] = tu;
    return CDLM_OK;
}

extern "C" cdlm_status cdlm_tuning_get(cdlm_ctx *ctx, int32_t L4, int32_t M, int32_t gamma, cdlm_tuning *out) {
    if (!ctx || !out) return CDLM_ERR_ARG;
    cdlm_status st = tune_args(ctx, L4, M, gamma);
    if (st != CDLM_OK) return st;
    const Tuning tu = tuning_of(ctx->device, L4 + 2 * gamma, M, gamma);
    out->fft_order = tu.order;
    out->fft4_order = tu.order4;
    out->dict_batch = tu.B;
    out->csc_slots = tu.nsl;
    out->csc_waves = tu.waves;
    out->csc_mfma = tu.mfma;
    return CDLM_OK;
}

// Times the candidate launch choices on a scratch problem of n_seq random sequences and
// keeps the fastest for one outer iteration of 15 CSC iterations and 1 dictionary
// iteration (the reference's default schedule, WRAPPER_setup.jl:8-10).
extern "C" cdlm_status cdlm_autotune(cdlm_ctx *ctx, int32_t L4, int32_t M, int32_t gamma, int32_t n_seq,
                                     cdlm_tuning *out) {
    if (!ctx) return CDLM_ERR_ARG;
    cdlm_status st = tune_args(ctx, L4, M, gamma);
    if (st != CDLM_OK) return st;
    // six per CU: whole rounds for one, two and three resident work-groups per CU, so that no
    // candidate is timed with a part-empty last round
    if (n_seq <= 0) n_seq = 6 * (ctx->num_cu > 0 ? ctx->num_cu : 256);
    const int Lg = L4 + 2 * gamma;
    // scratch data: one-hot columns from a fixed LCG
    std::vector<float> s((size_t)L4 * n_seq, 0.f);
    uint64_t x = 0x2545F4914F6CDD1DULL;
    for (size_t i = 0; i < s.size(); i += 4) {
        x = x * 6364136223846793005ULL + 1442695040888963407ULL;
        s[i + ((x >> 33) & 3)] = 1.f;
    }
    std::vector<int32_t> fl(M, gamma);
    cdlm_params prm;
    prm.M = M; prm.filter_len = fl.data(); prm.rho = 1.f; prm.sigma = 100.f; prm.lambda = 0.05f;
    prm.sherman_csc = 2.f; prm.sherman_cdl = 2.f;
    cdlm_problem *p = nullptr;
    if ((st = cdlm_problem_create(ctx, &prm, s.data(), L4, n_seq, nullptr, nullptr, 7, &p)) != CDLM_OK) return st;
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
        (void)hipGetLastError();
        cdlm_problem_destroy(p);
        return cdlm_fail(ctx, CDLM_ERR_HIP, "event creation failed");
    }
    auto timed = [&](bool dict, int reps, float *ms) -> cdlm_status {
        cdlm_status r = dict ? run_dict(ctx, p->set, nullptr) : run_csc(ctx, p->set, nullptr, 1);   // warm-up
        if (r != CDLM_OK) return r;
        CDLM_HIP(ctx, hipEventRecord(a, ctx->stream));
        if (dict)
            for (int i = 0; i < reps && r == CDLM_OK; ++i) r = run_dict(ctx, p->set, nullptr);
        else
            r = run_csc(ctx, p->set, nullptr, reps);
        if (r != CDLM_OK) return r;
        CDLM_HIP(ctx, hipEventRecord(b, ctx->stream));
        CDLM_HIP(ctx, hipEventSynchronize(b));
        CDLM_HIP(ctx, hipEventElapsedTime(ms, a, b));
        *ms /= reps;
        return CDLM_OK;
    };
    auto apply = [&](const Tuning &tu) -> cdlm_status {
        p->tune = tu;
        if (!make_fft_plan(Lg, 1, &p->dv.plan, tu.order) || !make_fft_plan(Lg / 4, 4, &p->dv.plan4, tu.order4))
            return cdlm_fail(ctx, CDLM_ERR_UNSUPPORTED, "no FFT plan for length %d", Lg);
        (void)hipStreamSynchronize(ctx->stream);
        p->set.release();
        return set_build(ctx, &p, 1, &p->set);
    };
    const int need = (Lg / 4 <= 256) ? 4 : 8;
    // CSC candidates: (kernel, slots per lane, waves per work-group)
    const int kCsc = 6;
    const int cand_mfma[kCsc] = {1, 1, 1, 1, 2, 2}, cand_nsl[kCsc] = {4, 4, 4, 8, 0, 0},
              cand_waves[kCsc] = {4, 8, 16, 4, 8, 16};
    const int kDict = 3;
    const int cand_B[kDict] = {2, 4, 8};
    float best_csc[kFftOrders][kCsc], best_dict[kFftOrders][kFftOrders][kDict];
    Tuning best;
    float best_t = 0.f;
    bool have = false;
    // every candidate timed in each of kRounds interleaved rounds, its fastest round kept (the
    // candidates differ by a few percent, about the spread of a single timing)
    constexpr int kRounds = 3;
    for (int o = 0; o < kFftOrders; ++o) {
        for (int ci = 0; ci < kCsc; ++ci) best_csc[o][ci] = -1.f;
        for (int o4 = 0; o4 < kFftOrders; ++o4)
            for (int bi = 0; bi < kDict; ++bi) best_dict[o][o4][bi] = -1.f;
    }
    auto keep_min = [](float &slot, float t) { if (slot < 0.f || t < slot) slot = t; };
    for (int r = 0; r < kRounds && st == CDLM_OK; ++r)
        for (int o = 0; o < kFftOrders && st == CDLM_OK; ++o) {
            for (int ci = 0; ci < kCsc && st == CDLM_OK; ++ci) {
                const int nsl = cand_nsl[ci];
                if (cand_mfma[ci] == 1 && nsl < need) continue;
                int pk = 0, pw = 0;
                pick_csc(Lg, gamma, nsl < need ? need : nsl, cand_mfma[ci], cand_waves[ci], &pk, &pw);
                if (pk != cand_mfma[ci] || pw != cand_waves[ci]) continue;
                Tuning tu;
                tu.order = o; tu.nsl = nsl; tu.waves = cand_waves[ci]; tu.mfma = cand_mfma[ci];
                float t = 0.f;
                if ((st = apply(tu)) == CDLM_OK && (st = timed(false, 20, &t)) == CDLM_OK) keep_min(best_csc[o][ci], t);
            }
            for (int o4 = 0; o4 < kFftOrders && st == CDLM_OK; ++o4)
                for (int bi = 0; bi < kDict && st == CDLM_OK; ++bi) {
                    const int B = cand_B[bi];
                    if (lds_bytes_dict(Lg, gamma, B) > kLdsLimit - kThreads * sizeof(double)) continue;
                    Tuning tu;
                    tu.order = o; tu.order4 = o4; tu.B = B;
                    float t = 0.f;
                    if ((st = apply(tu)) == CDLM_OK && (st = timed(true, 2, &t)) == CDLM_OK) keep_min(best_dict[o][o4][bi], t);
                }
        }
    for (int o = 0; o < kFftOrders && st == CDLM_OK; ++o)
        for (int ci = 0; ci < kCsc; ++ci) {
            if (best_csc[o][ci] < 0) continue;
            for (int o4 = 0; o4 < kFftOrders; ++o4)
                for (int bi = 0; bi < kDict; ++bi) {
                    if (best_dict[o][o4][bi] < 0) continue;
                    const float t = 15.f * best_csc[o][ci] + best_dict[o][o4][bi];
                    if (!have || t < best_t) {
                        have = true;
                        best_t = t;
                        best.order = o; best.order4 = o4; best.nsl = cand_nsl[ci]; best.waves = cand_waves[ci];
                        best.mfma = cand_mfma[ci];
                        best.B = cand_B[bi];
                    }
                }
        }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    cdlm_problem_destroy(p);
    if (st != CDLM_OK) return st;
    if (!have) return cdlm_fail(ctx, CDLM_ERR_UNSUPPORTED, "no launch choice fits shape Lg=%d", Lg);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune[std::make_tuple(ctx->device, Lg, (int)M, (int)gamma)] = best;
    }
    if (out) {
        out->fft_order = best.order;
        out->fft4_order = best.order4;
        out->dict_batch = best.B;
        out->csc_slots = best.nsl;
        out->csc_waves = best.waves;
        out->csc_mfma = best.mfma;
    }
    return CDLM_OK;
}
