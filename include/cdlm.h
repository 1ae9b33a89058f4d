/*
 * cdlm.h -- C ABI of the MI355X (gfx950) CDL-ADMM and PWM-scan library (libcdlm.so).
 *
 * This ABI is the drop-in boundary for the hot path of CDLmotif.jl.  The reference has
 * no FFI; its hot path is a set of Julia methods on the mutable struct `cdl`
 * (src/SETUP/CDL_setup.jl:70-164) and on `good_stuff` (src/SETUP/WRAPPER_setup.jl:47-62).
 * Each entry point below names the reference method it replaces.  A Julia host binds them
 * with `ccall` (see the binding in INTEGRATION.md); Python binds them with ctypes.
 *
 * Conventions
 *  - Arrays cross the ABI as plain pointers in Julia column-major order (first index
 *    fastest), FP32, exactly the reference's layouts:
 *      s   (4L, N)        one-hot data matrix (src/LOAD/to_dummy.jl:29-38)
 *      D0  (Lg, M)        initial filters, Lg = 4L + 2*max(filter_len)
 *      X0  (Lg, M, N)     initial codes
 *  - Sizes are int32 (the reference's int_t, src/SETUP/TYPE_setup.jl:4); N may be int64
 *    for the PWM scan.  Positions returned are 0-based (the Julia side adds 1).
 *  - The library owns all device memory; host buffers are borrowed for the call only.
 *  - Every call is synchronous for the caller (results are in host memory when it
 *    returns) unless its comment says otherwise.  A handle is not thread-safe.
 *  - Errors: a cdlm_status code plus cdlm_last_error(ctx).  Out-of-memory is its own code
 *    (CDLM_ERR_OOM), never a crash.
 */
#ifndef CDLM_H
#define CDLM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    CDLM_OK = 0,
    CDLM_ERR_ARG = 1,          /* bad argument (shape, null pointer, value out of range) */
    CDLM_ERR_HIP = 2,          /* HIP runtime error */
    CDLM_ERR_OOM = 3,          /* device memory exhausted or admission budget exceeded */
    CDLM_ERR_UNSUPPORTED = 4,  /* shape outside what the kernels support (e.g. filter_len > 32) */
    CDLM_ERR_STATE = 5         /* call not valid in the current state */
} cdlm_status;

typedef struct cdlm_ctx cdlm_ctx;
typedef struct cdlm_problem cdlm_problem;

/* Hyper-parameters of one CDL problem: the fields of CDL_setup that reach the `cdl`
 * constructor (src/SETUP/WRAPPER_setup.jl:1-13; CDL_setup.jl:165-172,273-283). */
typedef struct {
    int32_t M;                  /* number of filters */
    const int32_t *filter_len;  /* M flattened filter lengths (4*filter_size), multiples of 4, <= 32 */
    float rho;                  /* CSC penalty (used by update_Y1!, csc_kernels.jl:91) */
    float sigma;                /* CDL penalty (update_g2!, cdl_kernels.jl:73) */
    float lambda;               /* sparsity weight (shrink!, csc_kernels.jl:61) */
    float sherman_csc;          /* sherman_const_csc, 2 in the reference (train.jl:4) */
    float sherman_cdl;          /* sherman_const_cdl, 2 in the reference (train.jl:49) */
} cdlm_params;

/* ---- context ---------------------------------------------------------------------- */
/* One context per device; owns the HIP stream the problems run on. */
cdlm_status cdlm_create(cdlm_ctx **out, int32_t device);
void cdlm_destroy(cdlm_ctx *ctx);
const char *cdlm_last_error(const cdlm_ctx *ctx);
/* Library version string, and a device-memory query (free, total bytes). */
const char *cdlm_version(void);
cdlm_status cdlm_mem_info(cdlm_ctx *ctx, size_t *free_bytes, size_t *total_bytes);

/* ---- CDL ADMM ----------------------------------------------------------------------- */
/* Replaces init_cdl / the cdl{T,S,Q} constructor (src/SETUP/CDL_setup.jl:165-283).
 * s: (L4, N) data.  D0: (Lg, M) filters before masking (NULL: drawn on the host from
 * seed as masked rand_probvec groups, CDL_setup.jl:59-68,183).  X0: (Lg, M, N) codes
 * before the stride-4 projection (NULL: N(0,1) drawn on the device from seed, the
 * CUDA.randn of CDL_setup.jl:191).  The problem state stays resident on the device. */
cdlm_status cdlm_problem_create(cdlm_ctx *ctx, const cdlm_params *params, const float *s, int32_t L4, int32_t N,
                                const float *D0, const float *X0, uint64_t seed, cdlm_problem **out);
/* As cdlm_problem_create, from base codes instead of the one-hot matrix: bases (N, L)
 * row-major, codes 0..3 = A,C,G,T (any other code: an all-zero 4-group); the one-hot data
 * (data_2_dummy, src/LOAD/to_dummy.jl:2-38) is built on the device, so N*L bytes cross the bus
 * instead of 16*N*L. */
cdlm_status cdlm_problem_create_from_bases(cdlm_ctx *ctx, const cdlm_params *params, const uint8_t *bases, int32_t L,
                                           int32_t N, const float *D0, const float *X0, uint64_t seed, cdlm_problem **out);
/* As cdlm_problem_create, for problems that may not fit the device: when the problem's footprint
 * (cdlm_footprint_bytes) exceeds device_budget bytes (0: 80% of the free device memory), the
 * state of all N sequences lives in pinned host memory and the sequences move through a device
 * window in chunks (the fewest, a divisor of the dictionary-mean chunk count, whose window fits
 * the budget): every pass over the sequences (code_update, update_best, dict_update, one outer
 * iteration of cdlm_optimize) moves each chunk in, runs it and moves it back (with two windows,
 * where they fit the budget, the next chunk's upload and the last one's download overlap the
 * current chunk's kernels).  The per-filter state and the fixed-order reductions stay on the
 * device, so the results equal the resident problem's bit for bit; the price is the
 * host<->device traffic of the whole state per pass.
 * The reference sizes its batches from the free memory the same way (src/SEARCH/store.jl:109-119).
 * A problem that fits stays resident.  A streamed problem cannot join a batch or an activation
 * record (its X_best slices are in host memory: read them with cdlm_get). */
cdlm_status cdlm_problem_create_streamed(cdlm_ctx *ctx, const cdlm_params *params, const float *s, int32_t L4,
                                         int32_t N, const float *D0, const float *X0, uint64_t seed,
                                         uint64_t device_budget, cdlm_problem **out);
/* Chunks a problem moves through the device in (1: resident). */
cdlm_status cdlm_problem_chunks(const cdlm_problem *p, int32_t *chunks);
cdlm_status cdlm_problem_destroy(cdlm_problem *p);
/* Re-runs the initialisation from the same D0 / X0 / seed without re-uploading s. */
cdlm_status cdlm_problem_reset(cdlm_problem *p);
/* Dimensions: Lg (padded length), gamma (max filter length), L (bases), N, M. */
cdlm_status cdlm_problem_dims(const cdlm_problem *p, int32_t *Lg, int32_t *gamma, int32_t *L, int32_t *N,
                              int32_t *M);
/* Device bytes held by the problem. */
size_t cdlm_problem_bytes(const cdlm_problem *p);
/* Device bytes a problem of this shape would hold (admission by HBM footprint). */
size_t cdlm_footprint_bytes(int32_t L4, int32_t N, int32_t M, int32_t gamma, int32_t keep_x0);

/* code_update!(c, num_iter, ComplexF32; sherman_const_csc)  (src/CDL/train.jl:1-44) */
cdlm_status cdlm_code_update(cdlm_problem *p, int32_t num_iter);
/* dict_update!(c, num_iter, ComplexF32; sherman_const_cdl)  (src/CDL/train.jl:46-89) */
cdlm_status cdlm_dict_update(cdlm_problem *p, int32_t num_iter);
/* update_best!(c, dat)  (src/CDL/train.jl:91-98).  obj: the objective (Float64, as the
 * reference's objective() returns); improved: 1 when X_best/D_best were replaced. */
cdlm_status cdlm_update_best(cdlm_problem *p, double *obj, int32_t *improved);
/* objective(c, signals)  (src/SETUP/CDL_setup.jl:286) */
cdlm_status cdlm_objective(cdlm_problem *p, double *obj);
/* The schedule of cdl_optimize for one training (src/WRAPPER/wrap.jl:5-11):
 * iter_num x { csc_iter x code_update; update_best; cdl_iter x dict_update }.
 * obj_trace (may be NULL) receives the iter_num objectives seen by update_best.
 * Runs without host synchronisation until the end. */
cdlm_status cdlm_optimize(cdlm_problem *p, int32_t iter_num, int32_t csc_iter, int32_t cdl_iter, double *obj_trace);
/* mult * std(X_best[X_best .> 0])  (src/WRAPPER/wrap.jl:12), corrected std; count of
 * positive entries in *count. */
cdlm_status cdlm_code_thresh(cdlm_problem *p, double mult, double *thresh, int64_t *count);
/* The reference's obj_best field (Float32). */
cdlm_status cdlm_obj_best(cdlm_problem *p, float *obj_best);

/* Field read-out (copies into a caller buffer of exactly the field's size in bytes). */
typedef enum {
    CDLM_F_X = 0,        /* (Lg, M, N) codes */
    CDLM_F_DX = 1,       /* (Lg, N) */
    CDLM_F_D = 2,        /* (Lg, M, N) consensus filters d */
    CDLM_F_G1 = 3,       /* (Lg, M) */
    CDLM_F_G2 = 4,       /* (Lg, N) */
    CDLM_F_G3 = 5,       /* (Lg, M) */
    CDLM_F_XBEST_SLICE = 6, /* (L, M, N) = X_best[gamma+1:4:gamma+4L, :, :] (src/SEARCH/store.jl:15) */
    CDLM_F_DBEST = 7,    /* (Lg, M) */
    CDLM_F_Y1 = 8,       /* (Lg, N) */
    CDLM_F_y1 = 11,      /* (Lg, N) */
    CDLM_F_y2 = 12,      /* (Lg, M, N) */
    CDLM_F_y3 = 13,      /* (Lg, M, N) */
    CDLM_F_g1 = 14,      /* (Lg, M, N) */
    CDLM_F_g2 = 15,      /* (Lg, N) */
    CDLM_F_g3 = 16,      /* (Lg, M, N) */
    CDLM_F_TD = 17       /* (Lg, N) */
} cdlm_field;
cdlm_status cdlm_get(cdlm_problem *p, int32_t field, void *dst, size_t bytes);

/* Kernel timing with HIP events recorded on the library's stream around each launch (the
 * reference only prints @time wall clocks, src/WRAPPER/wrap.jl:30-137).  enable != 0 starts
 * a fresh accumulation; read returns the summed device time and count of one kernel:
 * 0 = CSC iterations (one launch runs all iterations of a code_update!; the count is the
 * number of iterations), 1 = dictionary iteration (per-sequence kernel), 2 = best/objective
 * pass, 3 = dictionary mean reduction (counts are launches). */
cdlm_status cdlm_timing_enable(cdlm_problem *p, int32_t enable);
cdlm_status cdlm_timing_read(cdlm_problem *p, int32_t kernel, double *total_ms, int64_t *launches);
/* ---- batches of problems ----------------------------------------------------------------- */
/* The reference trains one dataset per cdl struct and loops over datasets on the host
 * (src/WRAPPER/wrap.jl:140-149 is called once per dataset).  A batch runs several problems
 * (of any N, L, M and filter lengths, all on one context) with one launch per kernel over
 * all of them; each problem's results are those it gets alone.  The problems stay owned by
 * the caller and must outlive the batch (cdlm_problem_destroy refuses a problem that is
 * still in a batch).  A problem must not appear twice in one batch. */
typedef struct cdlm_batch cdlm_batch;
cdlm_status cdlm_batch_create(cdlm_ctx *ctx, cdlm_problem *const *problems, int32_t count, cdlm_batch **out);
cdlm_status cdlm_batch_destroy(cdlm_batch *b);
/* enable != 0 (the default): a batch call replays a captured HIP graph of its launches
 * (captured on first use of a schedule, reused while the schedule is unchanged). */
cdlm_status cdlm_batch_set_graph(cdlm_batch *b, int32_t enable);
/* Number of shape classes the batch was split into: problems that share a launch
 * configuration run in one launch per kernel; classes run as concurrent graph branches. */
cdlm_status cdlm_batch_classes(const cdlm_batch *b, int32_t *count);
/* cdlm_problem_reset of every problem. */
cdlm_status cdlm_batch_reset(cdlm_batch *b);
/* cdlm_optimize of every problem; obj_traces (may be NULL): count x iter_num, problem-major. */
cdlm_status cdlm_batch_optimize(cdlm_batch *b, int32_t iter_num, int32_t csc_iter, int32_t cdl_iter, double *obj_traces);
/* One full training of every problem (cdl_optimize, src/WRAPPER/wrap.jl:4-12): reset,
 * the schedule, and the code thresholds mult * std(X_best[X_best .> 0]).  Outputs (each may
 * be NULL): obj_traces count x iter_num, thresh[count], counts[count]. */
cdlm_status cdlm_batch_train(cdlm_batch *b, int32_t iter_num, int32_t csc_iter, int32_t cdl_iter, double mult,
                             double *obj_traces, double *thresh, int64_t *counts);
/* Kernel timing of batch calls (as cdlm_timing_*); graphs are not used while it is on. */
cdlm_status cdlm_batch_timing_enable(cdlm_batch *b, int32_t enable);
cdlm_status cdlm_batch_timing_read(cdlm_batch *b, int32_t kernel, double *total_ms, int64_t *launches);

/* ---- serving --------------------------------------------------------------------------- */
/* A continuous-batching server of CDL training jobs on one device (the reference runs one
 * dataset at a time from find_motifs_fasta -> cdl_optimize, src/WRAPPER/wrap.jl:4-12,140-149).
 * The server owns its own context and host threads.  Jobs are admitted in arrival order
 * (smaller later jobs may pass a waiting one) while their device footprint
 * (cdlm_footprint_bytes) fits the memory budget and the resident jobs do not yet fill the
 * device (fill_seqs sequences).  By default resident jobs run concurrently, each on a worker
 * thread with its own stream, one outer iteration (csc_iter CSC iterations, update_best,
 * cdl_iter dictionary iterations) after another at its own pace, up to four at a time (or
 * max_active); with lockstep set, every tick advances all resident jobs by one outer iteration
 * in one batch instead.  A finished job's results move to host memory and its device
 * memory (one arena per job) is kept for later admissions, freed only when a new arena must be
 * allocated within the budget.  A job's results do not depend on the jobs it shares the device
 * with.  The calls below are thread-safe. */
typedef struct cdlm_server cdlm_server;
typedef struct {
    uint64_t mem_budget;   /* device bytes for resident jobs (0: 80% of the free memory at start) */
    int32_t max_active;    /* resident jobs at most (0: no limit besides memory) */
    int32_t csc_iter;      /* CSC iterations per outer iteration, all jobs (0: 15, WRAPPER_setup.jl:8) */
    int32_t cdl_iter;      /* dictionary iterations per outer iteration (0: 1, WRAPPER_setup.jl:9) */
    int32_t use_graph;     /* lockstep: replay ticks from captured HIP graphs */
    int32_t fill_seqs;     /* admit no job while resident jobs hold this many sequences
                            * (0: twelve per CU; < 0: no limit) */
    int32_t policy;        /* admission order: CDLM_POLICY_FIFO or CDLM_POLICY_SHORTEST (least work
                            * iter_num*N*Lg*M first); a job passed by 4 later arrivals goes first
                            * either way */
    int32_t lockstep;      /* 1: all resident jobs advance together, one batched outer iteration
                            * per tick; 0: concurrent jobs, each at its own pace */
} cdlm_server_opts;
enum { CDLM_POLICY_FIFO = 0, CDLM_POLICY_SHORTEST = 1 };
typedef struct {
    cdlm_params params;
    const float *s;        /* (L4, N) data, copied at submission */
    int32_t L4;
    int32_t N;
    const float *D0;       /* (Lg, M) initial filters or NULL (drawn from seed) */
    uint64_t seed;
    int32_t iter_num;      /* outer iterations (wrap.jl:6) */
    double thresh_mult;    /* code threshold multiplier (wrap.jl:12) */
    int32_t keep_codes;    /* keep the X_best slice (L, M, N) for cdlm_server_result */
} cdlm_job_spec;
typedef enum { CDLM_JOB_QUEUED = 0, CDLM_JOB_RUNNING = 1, CDLM_JOB_DONE = 2, CDLM_JOB_FAILED = 3 } cdlm_job_state;
typedef struct {
    double latency_s;      /* submission -> results in host memory */
    double queue_s;        /* submission -> admission */
    double service_s;      /* admission -> results in host memory */
    double thresh;         /* thresh_mult * std(X_best[X_best .> 0]) */
    int64_t count;         /* positive X_best entries */
    float obj_best;
    int32_t iter_num, L, M, N;
} cdlm_job_result;
typedef struct {
    int64_t completed, active, queued, ticks;   /* ticks: lockstep ticks, or outer iterations run */
    double p50_s, p99_s;   /* latency quantiles (nearest rank) of the completed jobs */
    uint64_t bytes_in_use, budget_bytes;
} cdlm_server_stats_t;
cdlm_status cdlm_server_create(int32_t device, const cdlm_server_opts *opts, cdlm_server **out);
/* Stops the thread; queued and resident jobs end as CDLM_JOB_FAILED. */
cdlm_status cdlm_server_destroy(cdlm_server *srv);
/* Message of the last failed server call or job; the string is a copy owned by the calling
 * thread, valid until that thread calls cdlm_server_last_error again. */
const char *cdlm_server_last_error(cdlm_server *srv);
/* Queues a job; CDLM_ERR_UNSUPPORTED outside the kernels' shape range.  A job whose footprint
 * exceeds the whole budget is streamed (cdlm_problem_create_streamed with the budget as its
 * device window): it waits until nothing else is resident and then runs alone, its state in
 * pinned host memory; CDLM_ERR_OOM when even its smallest window exceeds the budget. */
cdlm_status cdlm_server_submit(cdlm_server *srv, const cdlm_job_spec *job, uint64_t *id);
cdlm_status cdlm_server_poll(cdlm_server *srv, uint64_t id, int32_t *state);
/* Waits until the job is done or failed, at most timeout_ms (< 0: no limit). */
cdlm_status cdlm_server_wait(cdlm_server *srv, uint64_t id, int32_t timeout_ms, int32_t *state);
/* Results of a done job; obj_trace (iter_num entries) and codes may be NULL. */
cdlm_status cdlm_server_result(cdlm_server *srv, uint64_t id, cdlm_job_result *res, double *obj_trace, float *codes,
                               size_t codes_bytes);
/* Forgets a finished job. */
cdlm_status cdlm_server_release(cdlm_server *srv, uint64_t id);
cdlm_status cdlm_server_stats(cdlm_server *srv, cdlm_server_stats_t *out);

/* ---- launch tuning ------------------------------------------------------------------ */
/* Launch choices for problems of one shape (L4, M, max filter length gamma); problems
 * created afterwards with that shape use them.  The reference's cuFFT plans choose their
 * own algorithms (CDL_setup.jl:158-163); here the choice is explicit and measured.
 * fft_order / fft4_order: stage order of the length-Lg and length-Lg/4 transforms
 * (0 = radix-4/2 stages first, 1 = odd primes first; the only choices that change
 * rounding); dict_batch: dictionary waves per work-group, one column in flight per wave
 * (2, 4 or 8; 0 = by occupancy; it changes the order of the sum over waves, hence rounding);
 * csc_slots: CSC float4 slots per lane (4 or 8; 0 = the fewest that cover Lg);
 * csc_waves: CSC waves per work-group (4: three sequences per CU with register prefetch;
 * 8: two sequences per CU where their LDS fits, one streaming while the other transforms;
 * 16: one sequence per CU; with 8 and 16 the resident sequences' state stays in the
 * last-level cache between the CSC iterations of one launch; 0 = 8 where two work-groups'
 * LDS fits a CU, else 16 where one fits, else 4;
 * it changes the order of the sum over waves, hence rounding);
 * csc_mfma: CSC kernel (1: the filter correlation and convolution on the vector ALUs, as
 * above; 2: on the matrix cores as Toeplitz x Hankel products of 16 x 256-position blocks, one
 * column per wave, csc_waves 8 (two work-groups per CU where their LDS fits) or 16, csc_slots
 * unused; 0 = the default for the shape; the choice changes rounding). */
typedef struct {
    int32_t fft_order;
    int32_t fft4_order;
    int32_t dict_batch;
    int32_t csc_slots;
    int32_t csc_waves;
    int32_t csc_mfma;
} cdlm_tuning;
/* Times every candidate on a scratch problem of n_seq sequences (<= 0: six per CU),
 * each in three interleaved rounds (its fastest round counts), keeps the fastest for 15 CSC + 1
 * dictionary iterations, and returns it in *out (may be NULL). */
cdlm_status cdlm_autotune(cdlm_ctx *ctx, int32_t L4, int32_t M, int32_t gamma, int32_t n_seq, cdlm_tuning *out);
/* Sets (t != NULL) or clears (t == NULL) the choices for a shape; get returns those in force. */
cdlm_status cdlm_tuning_set(cdlm_ctx *ctx, int32_t L4, int32_t M, int32_t gamma, const cdlm_tuning *t);
cdlm_status cdlm_tuning_get(cdlm_ctx *ctx, int32_t L4, int32_t M, int32_t gamma, cdlm_tuning *out);

/* Self-test of the device FFT (no reference counterpart; it pins the transforms that replace
 * the reference's cuFFT plans, CDL_setup.jl:158-163,195-204): nb complex transforms of
 * length n, interleaved (re, im) FP32, unscaled, forward exp(-2 pi i kl/n) or inverse
 * exp(+2 pi i kl/n); twstride = table length / n as the kernels use it (1, or 4 for the
 * quarter-length transforms); order: 0 = radix-4/2 stages first, 1 = odd primes first. */
cdlm_status cdlm_selftest_fft(cdlm_ctx *ctx, int32_t n, int32_t twstride, int32_t nb, int32_t inverse, int32_t order,
                              const float *in, float *out);
/* Device copy of `bytes` bytes, `reps` times, on the context stream; average ms per copy.
 * A plain streaming kernel of known traffic, used to calibrate HBM byte counters. */
cdlm_status cdlm_copy_bench(cdlm_ctx *ctx, size_t bytes, int32_t reps, double *ms_per_copy);

/* ---- activation record and topology ------------------------------------------------------ */
/* activation_record{T,S}(N, C, M) (src/SETUP/SEARCH_setup.jl:6-54) on the device: up to
 * max_trainings X_best slices (L, M, N) with their thresholds, the filters with any
 * activation, their training numbers and lengths, and the cover starts.  C = L. */
typedef struct cdlm_record cdlm_record;
cdlm_status cdlm_record_create(cdlm_ctx *ctx, int32_t N, int32_t L, int32_t M, int32_t max_trainings, cdlm_record **out);
cdlm_status cdlm_record_destroy(cdlm_record *r);
/* record_training_result!(c, thresh, filter_spec, cover_starts, ar, train_num) (src/SEARCH/store.jl:1-40):
 * copies the problem's X_best slice (device to device), appends the filters m with any code
 * > thresh (ascending, store.jl:19) with training number T+1 and length filter_spec[m]
 * (filter_size in bases, wrap.jl:15), and the cover starts (N values, NULL: all 1, wrap.jl:2);
 * thresh must be > 0 (store.jl:8).  *n_added = number of filters appended. */
cdlm_status cdlm_record_add(cdlm_record *r, cdlm_problem *p, double thresh, const int32_t *filter_spec,
                            const int32_t *cover_starts, int32_t *n_added);
/* K filters recorded, T trainings, cover width C, N sequences. */
cdlm_status cdlm_record_dims(const cdlm_record *r, int32_t *K, int32_t *T, int32_t *C, int32_t *N);
/* ar.uinds (1-based), ar.training_num, ar.filter_lengths: K entries each (any may be NULL). */
cdlm_status cdlm_record_filters(const cdlm_record *r, int32_t *uinds, int32_t *train_num, int32_t *filter_lengths);
/* ar.start_positions[n] for n in [n0, n0+count): (C, K) Bool each, as bytes, (C, K, count). */
cdlm_status cdlm_record_start_positions(cdlm_record *r, int32_t n0, int32_t count, uint8_t *out);
/* Sizes of the topology: K, maximum_filter_length, maximum_apart_length = C - min length - 1
 * (src/SEARCH/helpers.jl:9-11). */
cdlm_status cdlm_record_topology_dims(const cdlm_record *r, int32_t *K, int32_t *max_fil_len, int32_t *apart_len);
/* get_spatial_info!(ar) (src/SEARCH/store.jl:121-144) with topology_at_nth_pt! (store.jl:64-107):
 * Int32 counts overlap (K, K, max_fil_len), non_overlap (K, K, apart_len), next_to (K, K),
 * column-major (activations_topology{int_t}, SEARCH_setup.jl:73-90). */
cdlm_status cdlm_record_topology(cdlm_record *r, int32_t *overlap, int32_t *non_overlap, int32_t *next_to);
/* topology_at_nth_pt! over host arrays: start_pos (C, K, N) bytes, train_num (K) 1-based,
 * fil_lens (K), cover_starts (N, T) 1-based, column-major; outputs as cdlm_record_topology. */
cdlm_status cdlm_topology(cdlm_ctx *ctx, const uint8_t *start_pos, const int32_t *train_num, const int32_t *fil_lens,
                          const int32_t *cover_starts, int32_t C, int32_t K, int64_t N, int32_t T, int32_t max_fil_len,
                          int32_t apart_len, int32_t *overlap, int32_t *non_overlap, int32_t *next_to);
/* get_pair_pfms! up to the PFMs (src/SEARCH/extractions.jl:18-100 with get_f_pair_starts,
 * get_span_len, msa_fill, msa_pseudocount, :115-190): for each pair (fil1, fil2, overlap)
 * (pairs (P, 3) column-major, 1-based indices into the record's filters; the reference takes
 * the entries of overlap > count_threshold in column-major order), the weighted one-hot
 * counts of data (4C, N) over the pair's span at every sequence where fil1 starts at a and
 * fil2 at a + len1 - overlap, summed over sequences, plus 0.1 on the span, divided by the
 * first column's total.  Outputs span_len (P) and pfms (P, 4, max_span_len) column-major,
 * zero past each span.  4 x max_span_len <= 64. */
cdlm_status cdlm_record_pair_pfms(cdlm_record *r, const float *data, int32_t P, const int32_t *pairs,
                                  int32_t max_span_len, int32_t *span_len, float *pfms);
/* Device time (HIP events) of the last topology kernel. */
cdlm_status cdlm_topology_last_ms(cdlm_ctx *ctx, double *ms);

/* ---- PWM scan ----------------------------------------------------------------------- */
/* PWMs are packed (K, 4, maxlen) row-major, pwm[(k*4 + a)*maxlen + j], zero padded past
 * lens[k]; rows a = A,C,G,T.  Sequences are base codes 0..3, (N, L) row-major, one byte
 * each (the information content of the reference's one-hot columns).  Scores are FP32. */

/* All windows with score > thresh[k]: the window score is the position-major sequential
 * FP32 sum of greedy_search! (src/SEARCH/scan.jl:223-239); strand 1 uses the reverse
 * complement PWM (pwm_complement = reverse in both dimensions, src/SEARCH/helpers.jl:16).
 * Hit records (k, n, pos, strand); order unspecified (sort on the host for a canonical
 * order).  On CDLM_OK *n_hits is the total; if it exceeds cap, only cap records were
 * written and the call returns CDLM_ERR_OOM (rerun with cap >= *n_hits). */
cdlm_status cdlm_pwm_hits(cdlm_ctx *ctx, const uint8_t *bases, int32_t L, int64_t N, const float *pwms,
                          const int32_t *lens, const float *thresh, int32_t K, int32_t maxlen, int32_t both_strands,
                          int32_t *hit_k, int64_t *hit_n, int32_t *hit_pos, int8_t *hit_strand, int64_t cap,
                          int64_t *n_hits);
/* scan_w_gpu! (src/SEARCH/scan.jl:176-221): forward strand, per (k, n) the first argmax
 * over positions of (score > thresh ? score : 0) with positions past L-m+1 at 0; a best
 * position of 1 (1-based) is "no hit".  out_pos (K, N) row-major: 1-based position or 0. */
cdlm_status cdlm_pwm_argmax(cdlm_ctx *ctx, const uint8_t *bases, int32_t L, int64_t N, const float *pwms,
                            const int32_t *lens, const float *thresh, int32_t K, int32_t maxlen, int32_t *out_pos);
/* overlapping_scan! (src/SEARCH/scan.jl:89-121 with scan_n :10-21): nucleotide-major FP32
 * sums, best window per strand (first maximum), forward wins ties, hit iff > thresh.
 * Outputs (K, N) row-major: 1-based position or 0, strand (1 = reverse complement),
 * score (0 when no hit). */
cdlm_status cdlm_pwm_best(cdlm_ctx *ctx, const uint8_t *bases, int32_t L, int64_t N, const float *pwms,
                          const int32_t *lens, const float *thresh, int32_t K, int32_t maxlen, int32_t *out_pos,
                          int8_t *out_strand, float *out_score);
/* Device time (HIP events on the context stream) of the last scan kernel: hit list
 * (cdlm_pwm_hits), argmax (cdlm_pwm_argmax) or best hit (cdlm_pwm_best). */
cdlm_status cdlm_pwm_last_ms(cdlm_ctx *ctx, double *ms);
/* get_thresh(pwms, pval, Float32) (src/SEARCH/helpers.jl:99): per PWM the Float32 of
 * pvalue2score(pwm, pval) (src/SEARCH/PWM_Touzet.jl:199-217; granularity eps, 0.1 in the
 * reference, TYPE_setup.jl:52; uniform background): the PWM rounded down to multiples of eps,
 * its score distribution over all 4^m sequences, and the lowest attainable score whose upper
 * tail probability is below pval (the highest attainable score when there is none).
 * Computed with exact integer counts, which equals the reference's Double64 arithmetic for
 * lens <= 53.  PWMs packed as for the scans; pval in [0, 1] (the reference passes a Float32,
 * pass its exact value); entries finite with |x| <= 1e5; eps >= 1e-3. */
cdlm_status cdlm_pwm_pvalue_thresh(cdlm_ctx *ctx, const float *pwms, const int32_t *lens, int32_t K, int32_t maxlen,
                                   double pval, double eps, float *thresh);
/* Device time (HIP events) of the last cdlm_pwm_pvalue_thresh kernels. */
cdlm_status cdlm_pwm_pvalue_last_ms(cdlm_ctx *ctx, double *ms);
/* One-hot (4L, N) FP32 columns to base codes (N, L); a column with no 1 in a 4-group
 * gets code 4 (never matches a PWM row).  Replaces the device side of data_2_dummy. */
cdlm_status cdlm_onehot_to_bases(cdlm_ctx *ctx, const float *onehot, int32_t L4, int64_t N, uint8_t *bases);

/* Self-test of the matrix-core operand layout and rounding the PWM hit-list kernel relies
 * on: one v_mfma_f32_16x16x4_f32 with lane l supplying a64[l] and b64[l] and accumulator
 * inputs c256[4l..4l+3]; d256 receives each lane's 4 results (no reference counterpart). */
cdlm_status cdlm_selftest_mfma(cdlm_ctx *ctx, const float *a64, const float *b64, const float *c256, float *d256);

/* ---- Device-resident variants (inputs already in HBM, for benchmarks and hosts that
 * keep data on the device).  Pointers are device pointers; outputs stay on the device
 * and the call returns after enqueueing unless noted. ------------------------------- */
/* Hit list into device buffers (hit_k, hit_n, hit_pos, hit_strand), counter in
 * d_count (device int64, zeroed by the call).  Synchronous: *n_hits is valid on return. */
cdlm_status cdlm_pwm_hits_device(cdlm_ctx *ctx, const uint8_t *d_bases, int32_t L, int64_t N, const float *d_pwms,
                                 const int32_t *lens, const float *thresh, int32_t K, int32_t maxlen,
                                 int32_t both_strands, int32_t *d_hit_k, int64_t *d_hit_n, int32_t *d_hit_pos,
                                 int8_t *d_hit_strand, int64_t cap, int64_t *n_hits);

/* Allocation-free variant of cdlm_pwm_hits_device.  It has no counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and results as cdlm_pwm_hits_device;
 * the strand-PWM tiles (B operands) and the hit counter live in a fixed library array of
 * CDLM_PWM_SCRATCH_BYTES per GPU (held from first use until the process ends) instead of buffers
 * allocated per call, and the call allocates and frees no device memory.  Limit: the tiles of
 * both strands, about 2*K*maxlen*16 bytes plus 4 KiB, must fit the array (K = 512 PWMs of up to
 * 64 columns fit); larger inputs get CDLM_ERR_UNSUPPORTED.  The counter is cleared by a kernel
 * at the start of each call.  All calls on one GPU share its array: calls that overlap on one
 * GPU (from other threads or contexts) race and must be serialised by the caller. */
#define CDLM_PWM_SCRATCH_BYTES (4 * 1024 * 1024)
cdlm_status cdlm_pwm_hits_device_noalloc(cdlm_ctx *ctx, const uint8_t *d_bases, int32_t L, int64_t N,
                                         const float *pwms, const int32_t *lens, const float *thresh, int32_t K,
                                         int32_t maxlen, int32_t both_strands, int32_t *d_hit_k, int64_t *d_hit_n,
                                         int32_t *d_hit_pos, int8_t *d_hit_strand, int64_t cap, int64_t *n_hits);

#ifdef __cplusplus
}
#endif
#endif /* CDLM_H */
