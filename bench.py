"""Benchmark of the CDL-ADMM hot path on one MI355X.

Workload (the single-dataset configuration): N=20000 sequences, L=200 bases (4L=800),
M=64 filters of flattened length 20 (Lg=840), lambda=0.05, rho=1, sigma=100; one step is
one full training of cdl_optimize (src/WRAPPER/wrap.jl:5-12): initialisation, then
10 x (20 CSC iterations + update_best + 1 dictionary iteration) = 200 ADMM iterations, then
the code threshold.  Metric: CSC-ADMM seq*iters/s = N * 200 / seconds per step.

Inputs are resident in HBM before timing: the data matrix is uploaded once; each step
re-initialises the state on the device (D0 from the host generator, X0 ~ N(0,1) drawn on
the device), as the reference's constructor does.  Every library call synchronises its
stream before returning, so the wall clock brackets complete device work.

Printed: one JSON line with the metric, a roofline object for the dominant kernel (the
fused CSC iteration, HBM-bound; per-launch average from HIP events on the library's stream
over the timed steps), the step-level and dictionary-kernel fractions, a cpu_baseline object
(the CPU oracle on bounded samples of the single-dataset workload, the CPU-runnable workload
in full and the PWM-scan workload) and the legs of the other configurations: the batched
workload (256 datasets), the served workload (Poisson arrivals), the PWM scan (10^6
sequences), the activation topology, problem creation from base codes and the reference's
step sequence on the GPU.  Legs run in priority order within a time budget (--budget-s,
counted from the start of the process); a leg that would not fit is recorded as skipped.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

T_START = time.perf_counter()
ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cdlmotif.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)

N_SEQ, L_BASES, M_FILT, FLEN = 20000, 200, 64, 20
ITER_NUM, CSC_ITER, CDL_ITER = 10, 20, 1
LAM, RHO, SIGMA = 0.05, 1.0, 100.0
HBM_PEAK_GBS = 8000.0
PWM_L, PWM_K = 500, 512
MFMA_FP32_PEAK_TFLOPS = 157.3
C3_ITER_NUM, C3_CSC_ITER, C3_CDL_ITER = 10, 15, 1


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# ------------------------------------------------------------------------- byte counts
def compulsory_csc_bytes(N, Lg, M):
    """Compulsory HBM bytes of one CSC iteration, 4 N Lg (6M + 7): X, y2 and y3 read and
    written (6M), y1, Y1 and DX read and written and the data read (7); persistent state only,
    every spectrum and auxiliary array recomputed on chip."""
    return 4.0 * N * Lg * (6 * M + 7)


def compulsory_dict_bytes(N, Lg, M):
    """Compulsory HBM bytes of one dictionary iteration, 4 N Lg (8M + 4): X, y3, d, g1 and g3
    read, d, g1 and g3 written (8M); g2 read and written, the data and Td (4)."""
    return 4.0 * N * Lg * (8 * M + 4)


def carry_csc_bytes(N, Lg, M, iters):
    """This kernel's own floor for one CSC launch of `iters` iterations, per (l, n), 4 bytes
    each: the state enters as Xt and y2 (2M) and y3 off the stride-4 support (0.75M; on the
    support y3 is identically zero) and leaves the same way; between two iterations of the
    launch it is carried as the two shrink/projection inputs u2 = X + y2 and u3 = X + y3 (2M
    written, 2M read).  So 5.5M + 4M (iters - 1) in all; once per launch y1, Y1, DX, h and the
    data read and h, DX, y1, Y1 written (9), which stay on chip between iterations."""
    return 4.0 * N * Lg * ((5.5 + 4.0 * (iters - 1)) * M + 9)


def dict_floor_bytes(N, Lg, M):
    """This dictionary kernel's own floor: per column X read (Lg), the g1 half spectrum read and
    written (2 x 2 Hs) and the stored Y3 half spectrum read and written (2 x 2 Ph); per sequence
    G2, g2 and the data read, G2, g2 and Td written (6 Lg) and kappa read and written (4 Hs)."""
    Hs = ((Lg // 2 + 1) + 1) & ~1
    Ph = Lg // 8 + 1
    return 4.0 * N * (M * (Lg + 4 * Hs + 4 * Ph) + 6 * Lg + 4 * Hs)


# ------------------------------------------------------------------------- time budget
class Budget:
    def __init__(self, seconds: float):
        self.seconds = seconds
        self.skipped = {}

    def left(self) -> float:
        return self.seconds - (time.perf_counter() - T_START)

    def fits(self, name: str, estimate_s: float) -> bool:
        if self.left() >= estimate_s:
            return True
        self.skipped[name] = "time budget: %.0f s estimated, %.0f s left of %.0f s" % (estimate_s, self.left(),
                                                                                       self.seconds)
        log("bench: skip %s (%s)" % (name, self.skipped[name]))
        return False


def cpu_info():
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return threads, model


# ------------------------------------------------------------------------- CPU baselines
def cpu_baseline(sample_n: int, pwm_seqs: int):
    """The oracle (the reference-structured FP32 restatement, OpenMP over columns) on the host
    cores.  value: the single-dataset workload's metric on a sample_n-sequence slice for one
    outer iteration (20 CSC + update_best + 1 dict), the same work per sequence-iteration as a
    step (a 1000-sequence slice for all 200 iterations takes about 25 minutes on 8 cores, far
    beyond a benchmark run; the CSC work is per sequence, so seq*iters/s carries over).  c1:
    the CPU-runnable workload in full (N=500, L=100, M=10, filter_len 16, 100 ADMM iterations,
    one training).  c3_dataset: one dataset of the batched configuration at its real shape, 512 of
    its sequences for one outer iteration (the served configuration's jobs are such datasets).
    pwm: the PWM-scan workload's hit list on pwm_seqs sequences x 512 PWMs, the oracle's scan in
    parallel threads over PWM chunks."""
    import numpy as np
    import oracle as O
    from cdlm import synth
    threads, model = cpu_info()
    bases, _ = synth.c2_dataset(seed=2, N=sample_n, L=L_BASES)
    s = synth.onehot(bases)
    Lg = 4 * L_BASES + 2 * FLEN
    fl = [FLEN] * M_FILT
    D0 = synth.rand_probvec_filters(7, M_FILT, Lg, fl)
    X0 = np.random.default_rng(8).standard_normal((sample_n, M_FILT, Lg)).astype(np.float32)
    oc = O.OracleCDL(s, D0, X0, fl, RHO, SIGMA, LAM, "f32")
    t0 = time.perf_counter()
    oc.code_update(CSC_ITER)
    oc.update_best()
    oc.dict_update(CDL_ITER)
    dt = time.perf_counter() - t0
    del oc, X0
    out = {"value": sample_n * CSC_ITER / dt, "unit": "seq*iters/s", "cores": threads, "kind": "port",
           "host_cpus": os.cpu_count(), "cpu_model": model,
           "sample": "oracle f32 reference-structured CDL on %d sequences of the single-dataset shape (L=%d, M=%d, "
                     "filter_len=%d), one outer iteration (%d CSC + update_best + %d dict) in %.1f s"
                     % (sample_n, L_BASES, M_FILT, FLEN, CSC_ITER, CDL_ITER, dt)}
    log("bench: CPU baseline, single-dataset slice: %.1f seq*iters/s (%.1f s, %d threads)" % (out["value"], dt, threads))
    # the CPU-runnable workload in full
    b1, _ = synth.c1_dataset(seed=1, N=500, L=100)
    s1 = synth.onehot(b1)
    Lg1 = 4 * 100 + 2 * 16
    D01 = synth.rand_probvec_filters(11, 10, Lg1, [16] * 10)
    X01 = np.random.default_rng(21).standard_normal((500, 10, Lg1)).astype(np.float32)
    t0 = time.perf_counter()
    o1 = O.OracleCDL(s1, D01, X01, [16] * 10, RHO, SIGMA, LAM, "f32")
    for _ in range(10):
        o1.code_update(10)
        o1.update_best()
        o1.dict_update(1)
    th1, _ = o1.code_thresh(1.0)
    dt1 = time.perf_counter() - t0
    del o1
    out["c1"] = {"value": 500 * 100 / dt1, "unit": "seq*iters/s", "seconds": dt1, "cores": threads,
                 "sample": "the CPU-runnable workload in full: N=500 L=100 M=10 filter_len=16, 10 x (10 CSC + "
                           "update_best + 1 dict) = 100 ADMM iterations, one training, code threshold %.6f" % th1}
    log("bench: CPU baseline, CPU-runnable workload in full: %.2f s (%.0f seq*iters/s)" % (dt1, 500 * 100 / dt1))
    # one dataset of the batched configuration (its real L, M, filter_len), 512 of its sequences,
    # one outer iteration of its schedule
    sp = synth.c3_specs(1)[0]
    n3, fl3 = 512, [sp["filter_len"]] * sp["M"]
    b3, _ = synth.planted_bases(sp["seed"], n3, sp["L"], sp["motif_lens"], sp["occupancy"], alpha=0.5)
    Lg3 = 4 * sp["L"] + 2 * sp["filter_len"]
    D03 = synth.rand_probvec_filters(sp["seed"], sp["M"], Lg3, fl3)
    X03 = np.random.default_rng(9).standard_normal((n3, sp["M"], Lg3)).astype(np.float32)
    o3 = O.OracleCDL(synth.onehot(b3), D03, X03, fl3, RHO, SIGMA, LAM, "f32")
    t0 = time.perf_counter()
    o3.code_update(C3_CSC_ITER)
    o3.update_best()
    o3.dict_update(C3_CDL_ITER)
    dt3 = time.perf_counter() - t0
    del o3, X03
    out["c3_dataset"] = {"value": n3 * C3_CSC_ITER / dt3, "unit": "seq*iters/s", "seconds": dt3, "cores": threads,
                         "sample": "the first dataset of the batched configuration (L=%d M=%d filter_len=%d), %d of its "
                                   "sequences, one outer iteration (%d CSC + update_best + %d dict)"
                                   % (sp["L"], sp["M"], sp["filter_len"], n3, C3_CSC_ITER, C3_CDL_ITER)}
    log("bench: CPU baseline, one batched-configuration dataset (L=%d M=%d): %.1f seq*iters/s (%.1f s)"
        % (sp["L"], sp["M"], n3 * C3_CSC_ITER / dt3, dt3))
    if pwm_seqs > 0:
        from concurrent.futures import ThreadPoolExecutor
        pb, pw, pl, pt = synth.pwm_workload(5, pwm_seqs, PWM_L, PWM_K)
        bounds = np.linspace(0, PWM_K, threads + 1).astype(int)
        chunks = [(a, b) for a, b in zip(bounds[:-1], bounds[1:]) if b > a]
        t0 = time.perf_counter()
        with ThreadPoolExecutor(len(chunks)) as ex:
            nh = sum(ex.map(lambda ab: len(O.pwm_hits(pb, pw[ab[0]:ab[1]], pl[ab[0]:ab[1]], pt[ab[0]:ab[1]])[0]),
                            chunks))
        dtp = time.perf_counter() - t0
        windows = float(sum(pwm_seqs * (PWM_L - int(m) + 1) * 2 for m in pl))
        out["pwm"] = {"value": windows / dtp, "unit": "windows/s", "seconds": dtp, "cores": len(chunks), "hits": nh,
                      "sample": "PWM-scan workload hit list, both strands: %d sequences x L=%d against 512 PWMs of "
                                "length 8-24 (the workload's 10^6 sequences scale linearly)" % (pwm_seqs, PWM_L)}
        log("bench: CPU baseline, PWM scan %d x %d x %d: %.1f s (%.3g windows/s)" % (pwm_seqs, PWM_L, PWM_K, dtp,
                                                                                     windows / dtp))
    return out


# ------------------------------------------------------------------------- device legs
# Freeing device memory starts work in the driver that slows the device for a few seconds
# afterwards (measured: about 1 s per 50 GB freed; the next batch ran 2-4 s slower after 228 GB
# were freed, and at its normal speed after a 6 s pause).  The legs pause this long, outside
# their timed regions, after steps that free tens of GB, so that one leg's frees do not land in
# the next leg's timing.
SETTLE_S = 6.0


def settle(what):
    log("bench: %.0f s pause after %s" % (SETTLE_S, what))
    time.sleep(SETTLE_S)


def warm_kernels(ctx):
    """First use of every kernel variant, outside the timed regions."""
    import numpy as np
    import cdlm
    warm = [cdlm.CDL(np.zeros((4, 4 * 8), np.float32), [f] * 2, RHO, SIGMA, LAM, seed=1, ctx=ctx) for f in (16, 20, 24)]
    wb = cdlm.Batch(warm)
    wb.train(1, 1, 1)
    wb.close()
    for w in warm:
        w.optimize(1, 1, 1)
        w.close()


class C3Set:
    """Datasets of the batched configuration, generated on the host when first needed."""

    def __init__(self, count, N_range=(200, 20000)):
        from cdlm import synth
        self.specs = synth.c3_specs(count, N_range=N_range)
        self._data = {}

    def data(self, i):
        from cdlm import synth
        if i not in self._data:
            self._data[i] = synth.c3_onehot(self.specs[i])
        return self._data[i]

    def drop(self, i):
        self._data.pop(i, None)

    def make(self, i, ctx):
        import cdlm
        sp = self.specs[i]
        return cdlm.CDL(self.data(i), [sp["filter_len"]] * sp["M"], RHO, SIGMA, LAM, seed=sp["seed"], ctx=ctx)


def batched_leg(cs: C3Set, n_plain: int, budget: Budget, wave_frac=0.25, keep_plain=0):
    """Batched workload: the datasets of cs (N in [200, 20000], L in {100..500}, M in {32, 64},
    filter_len in {16, 20, 24}), each trained once (10 x (15 CSC + update_best + 1 dict) = 150
    ADMM iterations, code threshold).  Plain loop: one dataset at a time through the
    single-problem entry points, on the first n_plain datasets (the rest extrapolated from their
    mean time: the datasets are independent draws).  Batched: all datasets packed by device
    footprint into waves of at most wave_frac of the free device memory, one batch training (a
    captured HIP graph) per wave, as many waves as the time budget allows.  Both include problem
    creation (data upload) and teardown; the host-side data generation is outside the timing."""
    import numpy as np
    import cdlm
    ctx = cdlm.default_context()
    specs = cs.specs
    shapes = sorted({(sp["L"], sp["M"], sp["filter_len"]) for sp in specs})
    t0 = time.perf_counter()
    tunings = {shape: cdlm.autotune(*shape, ctx=ctx) for shape in shapes}
    log("bench: autotuned %d shapes in %.2f s" % (len(shapes), time.perf_counter() - t0))
    warm_kernels(ctx)
    th_plain, t_each = [], []
    for i in range(min(n_plain, len(specs))):
        cs.data(i)
        t1 = time.perf_counter()
        c = cs.make(i, ctx)
        c.optimize(C3_ITER_NUM, C3_CSC_ITER, C3_CDL_ITER)
        th_plain.append(c.code_thresh(1.0)[0])
        c.close()
        t_each.append(time.perf_counter() - t1)
        if i >= keep_plain:
            cs.drop(i)
    t_plain = float(sum(t_each))
    log("bench: batched workload, plain loop: %d datasets in %.2f s (%.3f datasets/s)" % (len(t_each), t_plain,
                                                                                           len(t_each) / t_plain))
    settle("the plain loop")
    free, _total = ctx.mem_info()
    budget_b = int(free * wave_frac)
    sizes = [cdlm.footprint_bytes(sp["L"], sp["N"], sp["M"], sp["filter_len"]) for sp in specs]
    waves = cdlm.pack_by_footprint(sizes, budget_b)
    est_wave = (t_plain / max(len(t_each), 1)) * len(specs) / max(len(waves), 1)
    th_batch = {}
    t_batch, done, waves_run = 0.0, 0, 0
    for w in waves:
        if budget.left() < 1.5 * est_wave + 20:
            log("bench: batched workload: stopping after %d of %d waves (time budget)" % (waves_run, len(waves)))
            break
        for i in w:
            cs.data(i)
        t1 = time.perf_counter()
        probs = [cs.make(i, ctx) for i in w]
        b = cdlm.Batch(probs)
        _tr, th, _cnt = b.train(C3_ITER_NUM, C3_CSC_ITER, C3_CDL_ITER, 1.0)
        b.close()
        for p in probs:
            p.close()
        t_batch += time.perf_counter() - t1
        for j, i in enumerate(w):
            th_batch[i] = float(th[j])
            if i >= keep_plain:
                cs.drop(i)
        done += len(w)
        waves_run += 1
    log("bench: batched workload, batched: %d datasets in %d waves, %.2f s" % (done, waves_run, t_batch))
    seq_done = sum(specs[i]["N"] for i in th_batch)
    rate_plain = len(t_each) / t_plain
    rate_batch = done / t_batch if t_batch > 0 else 0.0
    same = all(th_batch[i] == th_plain[i] or (np.isnan(th_batch[i]) and np.isnan(th_plain[i]))
               for i in range(len(th_plain)) if i in th_batch)
    batched_leg.cache = (cs, t_each, th_plain, tunings)
    return {
        "workload": "%d datasets of the batched configuration (seeds 3..%d, N in [200, 20000]), 150 ADMM iterations each"
                    % (len(specs), 2 + len(specs)),
        "datasets": len(specs), "batched_datasets_run": done, "waves": len(waves), "waves_run": waves_run,
        "wave_budget_bytes": budget_b, "wave_fraction_of_free": wave_frac,
        "plain_loop": {"datasets_timed": len(t_each), "seconds": t_plain, "datasets_per_s": rate_plain,
                       "note": "one dataset at a time; the whole set's plain time extrapolated from the timed "
                               "datasets' mean"},
        "batched_seconds": t_batch, "batched_datasets_per_s": rate_batch,
        "speedup_vs_plain_loop": rate_batch / rate_plain if rate_plain > 0 else None,
        "batched_seq_iters_per_s": seq_done * C3_ITER_NUM * C3_CSC_ITER / t_batch if t_batch > 0 else 0.0,
        "thresholds_identical_to_plain": bool(same),
    }


def served_leg(rate: float, n_jobs: int, policy: str = "fifo"):
    """Served workload sample: the first n_jobs datasets of the batched leg arrive as a Poisson
    process of the given rate (seeded) and are submitted to the library's server (continuous
    batching, admission by device footprint, one outer iteration per tick, recycled device
    memory).  Latency = arrival -> results in host memory.  The plain per-job loop's latencies
    follow from the measured alone times of the same datasets (FIFO, one job at a time)."""
    import threading
    import numpy as np
    import cdlm
    cs, t_each, th_plain, tunings = batched_leg.cache
    n = min(n_jobs, len(t_each))
    # the launch choices the plain results were computed with (the FFT stage order changes
    # rounding, so another choice would change result bits)
    for shape, choice in tunings.items():
        cdlm.tuning_set(*shape, choice)
    for i in range(n):
        cs.data(i)
    gaps = np.random.default_rng(4).exponential(1.0 / rate, n)
    arrive = np.cumsum(gaps) - gaps[0]
    srv = cdlm.Server(csc_iter=C3_CSC_ITER, cdl_iter=C3_CDL_ITER, policy=policy)
    ids = [None] * n
    t0 = time.perf_counter()

    def feeder():
        for i in range(n):
            dt = t0 + arrive[i] - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            sp = cs.specs[i]
            ids[i] = srv.submit(cs.data(i), [sp["filter_len"]] * sp["M"], RHO, SIGMA, LAM, C3_ITER_NUM, seed=sp["seed"])

    th = threading.Thread(target=feeder)
    th.start()
    th.join()
    lat, svc, same = [], [], True
    for i in range(n):
        srv.wait(ids[i])
        r = srv.result(ids[i])
        lat.append(r["latency_s"])
        svc.append(r["service_s"])
        same = same and (r["thresh"] == th_plain[i] or (np.isnan(r["thresh"]) and np.isnan(th_plain[i])))
    t_end = time.perf_counter()
    st = srv.stats()
    srv.close()
    done, plat = 0.0, []
    for i in range(n):
        done = max(done, arrive[i]) + t_each[i]
        plat.append(done - arrive[i])
    lat, plat = np.array(lat), np.array(plat)
    q = lambda v, p: float(np.sort(v)[max(0, int(np.ceil(p * len(v))) - 1)])
    log("bench: served (%s) %d jobs at %.1f jobs/s: p50 %.2f s p99 %.2f s, %.3f datasets/s (plain loop p50 %.2f s "
        "p99 %.2f s, %.3f datasets/s)" % (policy, n, rate, q(lat, 0.5), q(lat, 0.99), n / (t_end - t0), q(plat, 0.5),
                                         q(plat, 0.99), n / float(done)))
    return {
        "workload": "%d batched-configuration datasets arriving as Poisson(%.1f jobs/s), 150 ADMM iterations each"
                    % (n, rate), "policy": policy, "jobs": n, "arrival_rate": rate,
        "p50_latency_s": q(lat, 0.5), "p99_latency_s": q(lat, 0.99), "datasets_per_s": n / (t_end - t0),
        "ticks": st["ticks"], "plain_loop_p50_latency_s": q(plat, 0.5), "plain_loop_p99_latency_s": q(plat, 0.99),
        "plain_loop_datasets_per_s": n / float(done), "results_identical_to_alone": bool(same),
        "note": "the configuration's 1000 jobs per rate were not run (about 10 minutes per rate); the 5% large jobs "
                "(N=100000, L=500, M=128: 475 GB of state, more than the device) are not in the sample: the server "
                "streams such a job through a device window with its state in pinned host memory, and 475 GB of "
                "pinned memory is more than the measuring machine allows a process",
    }


def reference_loop_leg(s, fl, D0, t_fused_step, n_batch, csc_sample):
    """The plain per-dataset GPU loop, as the reference runs it, at the single-dataset
    configuration, against the fused and batched forms on the same shape.

    Plain loop: refloop.py (the reference's code_update!/dict_update!/update_best! statement
    for statement on torch + hipFFT, a device synchronisation at each CUDA.@sync) on the full
    N=20000 problem.  Bounded sample: after one untimed outer iteration (FFT plan creation),
    csc_sample CSC iterations, update_best and one dictionary iteration are timed; a training is
    10 x (20 CSC + update_best + 1 dict), extrapolated from those per-call times.
    Initialisation and the code threshold are left out, which favours the plain loop.  Batched:
    n_batch datasets of the same shape (other seeds) trained in one batch (a captured HIP graph),
    problem creation and teardown included."""
    import torch
    import cdlm
    from cdlm import synth
    from refloop import ReferenceStructuredCDL
    N = s.shape[0]
    Lg = 4 * L_BASES + 2 * FLEN
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    X0 = torch.randn((N, M_FILT, Lg), generator=g, device="cuda", dtype=torch.float32)
    t0 = time.perf_counter()
    r = ReferenceStructuredCDL(s, fl, RHO, SIGMA, LAM, D0, X0)
    t_init = time.perf_counter() - t0
    del X0
    r.code_update(1)
    r.update_best()
    r.dict_update(1)
    r.sync()
    t0 = time.perf_counter()
    r.code_update(csc_sample)
    r.sync()
    t_csc = (time.perf_counter() - t0) / csc_sample
    t0 = time.perf_counter()
    r.update_best()
    r.sync()
    t_best = time.perf_counter() - t0
    t0 = time.perf_counter()
    r.dict_update(CDL_ITER)
    r.sync()
    t_dict = time.perf_counter() - t0
    peak = torch.cuda.max_memory_allocated() / 1e9
    del r
    torch.cuda.empty_cache()
    settle("freeing the reference-structured loop's %.0f GB" % peak)
    t_train = ITER_NUM * (CSC_ITER * t_csc + t_best + CDL_ITER * t_dict)
    log("bench: reference-structured loop at N=%d: CSC %.1f ms, update_best %.1f ms, dict %.1f ms per call "
        "(init %.2f s, peak %.0f GB) -> %.2f s per training" % (N, 1e3 * t_csc, 1e3 * t_best, 1e3 * t_dict, t_init,
                                                                peak, t_train))
    out = {"workload": "single-dataset configuration N=%d L=%d M=%d filter_len=%d, 10 x (20 CSC + update_best + "
                       "1 dict) per training" % (N, L_BASES, M_FILT, FLEN),
           "plain_loop": {"csc_ms": 1e3 * t_csc, "update_best_ms": 1e3 * t_best, "dict_ms": 1e3 * t_dict,
                          "init_s": t_init, "peak_device_GB": peak, "training_s": t_train,
                          "datasets_per_s": 1.0 / t_train,
                          "sample": "%d CSC iterations + update_best + %d dict timed after one untimed outer "
                                    "iteration; training time extrapolated, initialisation excluded"
                                    % (csc_sample, CDL_ITER)},
           "fused_single": {"training_s": t_fused_step, "datasets_per_s": 1.0 / t_fused_step,
                            "vs_plain_loop": t_train / t_fused_step}}
    if n_batch > 0:
        ctx = cdlm.default_context()
        sets = [synth.onehot(synth.c2_dataset(seed=100 + i, N=N, L=L_BASES)[0]) for i in range(n_batch)]
        t0 = time.perf_counter()
        probs = [cdlm.CDL(sets[i], fl, RHO, SIGMA, LAM, D0=D0, X0=None, seed=200 + i, ctx=ctx) for i in range(n_batch)]
        b = cdlm.Batch(probs)
        _tr, th, _cnt = b.train(ITER_NUM, CSC_ITER, CDL_ITER, 1.0)
        b.close()
        for p in probs:
            p.close()
        t_b = time.perf_counter() - t0
        log("bench: batched %d datasets of the same shape in %.2f s (%.3f datasets/s)" % (n_batch, t_b, n_batch / t_b))
        settle("the same-shape batch")
        out["batched"] = {"datasets": n_batch, "seconds": t_b, "datasets_per_s": n_batch / t_b,
                          "vs_plain_loop": (n_batch / t_b) * t_train}
    return out


def topology_leg(c):
    """Activation record and topology (record_training_result!, get_spatial_info!,
    src/SEARCH/store.jl) for two trainings at the single-dataset configuration: the timed
    problem's last training and a second training from another seed; codes over each
    training's std threshold (wrap.jl:12), filter_size 5 bases.  Kernel time from HIP events;
    the kernel's work is the pairs of active entries of each sequence."""
    import ctypes as Cc
    import cdlm
    from cdlm import _lib, synth
    ctx = cdlm.default_context()
    bases, _ = synth.c2_dataset(seed=2, N=c.N, L=c.L)
    c2 = cdlm.CDL(synth.onehot(bases), [FLEN] * M_FILT, RHO, SIGMA, LAM, seed=98765)
    c2.optimize(ITER_NUM, CSC_ITER, CDL_ITER)
    rec = cdlm.ActivationRecord(c.N, c.L, M_FILT, max_trainings=2)
    t0 = time.perf_counter()
    for prob in (c, c2):
        th, _ = prob.code_thresh(1.0)
        rec.add(prob, th, [FLEN // 4] * M_FILT)
    t_add = time.perf_counter() - t0
    c2.close()
    over, non, nxt = rec.topology()
    v = Cc.c_double()
    _lib.check(_lib.load().cdlm_topology_last_ms(ctx.h, Cc.byref(v)), ctx.h)
    K, T, C, N = rec.dims()
    act = int(rec.start_positions(0, min(N, 2000)).sum()) * N / min(N, 2000)
    data = synth.onehot(bases)
    t0 = time.perf_counter()
    pairs, spans, pfms = cdlm.get_pair_pfms(rec, over, 50, data, FLEN // 4)
    t_pfm = time.perf_counter() - t0
    rec.close()
    npairs = (act / N) ** 2 * N
    log("bench: topology K=%d (2 trainings), ~%.0f active entries, kernel %.2f ms (%.2e pair visits/s), record %.2f s;"
        " %d pair PFMs in %.3f s" % (K, act, v.value, npairs / (v.value * 1e-3), t_add, len(pairs), t_pfm))
    return {"workload": "activation topology of 2 trainings at N=%d L=%d M=%d (filter_size %d)" % (N, C, M_FILT, FLEN // 4),
            "K": K, "active_entries": act, "kernel_ms": v.value, "record_s": t_add,
            "pair_visits_per_s": npairs / (v.value * 1e-3), "reference_pair_visits": float(C * K) ** 2 * N,
            "counts": {"overlap": int(over.sum()), "non_overlap": int(non.sum()), "next_to": int(nxt.sum())},
            "pair_pfms": {"pairs": int(len(pairs)), "seconds": t_pfm,
                          "note": "wall time of the call, including the 64 MB data upload"}}


def ingest_leg(bases, s, fl, D0):
    """Problem creation from the host one-hot matrix (4 bytes x 4L per sequence over the bus)
    against creation from 1-byte base codes expanded on the device; both must give the same
    objective."""
    import cdlm
    out = {}
    for kind in ("onehot", "bases"):
        t0 = time.perf_counter()
        if kind == "onehot":
            c = cdlm.CDL(s, fl, RHO, SIGMA, LAM, D0=D0, seed=1234)
        else:
            c = cdlm.CDL(None, fl, RHO, SIGMA, LAM, D0=D0, seed=1234, bases=bases)
        dt = time.perf_counter() - t0
        out[kind] = {"create_ms": 1000.0 * dt, "host_bytes": int(s.nbytes if kind == "onehot" else bases.size),
                     "objective": c.objective()}
        c.close()
        del c
    out["identical"] = out["onehot"]["objective"] == out["bases"]["objective"]
    log("bench: ingest: create from one-hot %.1f ms (%d B), from base codes %.1f ms (%d B), objectives equal %s"
        % (out["onehot"]["create_ms"], out["onehot"]["host_bytes"], out["bases"]["create_ms"],
           out["bases"]["host_bytes"], out["identical"]))
    return out


def pwm_leg(n_seq: int):
    """PWM scan: n_seq sequences of L=500 iid bases (GC 0.42) against 512 log-odds PWMs of
    length 8-24, both strands, threshold 0.8 x best score, full hit list; then the argmax and
    best-hit modes on the same workload.  Kernel time from HIP events around the MFMA kernels;
    algorithmic flops of the hit list = sum_k N (L - m_k + 1) x 2 strands x 2 x 4 m_k."""
    import ctypes as Cc
    import numpy as np
    import cdlm
    from cdlm import _lib, synth
    from cdlm import scan as S
    t0 = time.perf_counter()
    bases, pwms, lens, thresh = synth.pwm_workload(5, n_seq, PWM_L, PWM_K)
    log("bench: PWM workload %d x %d generated in %.1f s" % (n_seq, PWM_L, time.perf_counter() - t0))
    ctx = cdlm.default_context()
    hits = cdlm.scan_hits(bases, pwms, lens, thresh, both_strands=True, sort=False)   # warm-up, sizes the buffer
    ms = []
    for _ in range(3):
        t0 = time.perf_counter()
        h = cdlm.scan_hits(bases, pwms, lens, thresh, both_strands=True, cap=max(len(hits), 1024), sort=False)
        wall = time.perf_counter() - t0
        v = Cc.c_double()
        _lib.check(_lib.load().cdlm_pwm_last_ms(ctx.h, Cc.byref(v)), ctx.h)
        ms.append(v.value)
    k_ms = float(np.median(ms))
    flops = float(sum(n_seq * (PWM_L - int(m) + 1) * 2 * 2 * 4 * int(m) for m in lens))
    windows = float(sum(n_seq * (PWM_L - int(m) + 1) * 2 for m in lens))
    tf = flops / (k_ms * 1e-3) / 1e12
    n_hits = int(len(h))
    log("bench: PWM scan %d x %d, %d PWMs: kernel %.2f ms (%.1f TFLOP/s), %d hits, call %.2f s"
        % (n_seq, PWM_L, PWM_K, k_ms, tf, n_hits, wall))
    del hits, h
    # the best-window scans on the same workload: argmax of the thresholded position-major
    # scores, forward strand (scan_w_gpu!), and best hit per strand pair with nucleotide-major
    # sums (overlapping_scan!); algorithmic flops as for the hit list, one product per base and
    # position: 2 x 4 m_k per window and strand
    modes = {}
    for name, fn, strands in (("argmax", S.scan_w_gpu, 1), ("best_hit", S.overlapping_scan, 2)):
        mms = []
        for _ in range(3):
            fn(bases, pwms, lens, thresh)
            v = Cc.c_double()
            _lib.check(_lib.load().cdlm_pwm_last_ms(ctx.h, Cc.byref(v)), ctx.h)
            mms.append(v.value)
        mm = float(np.median(mms))
        fl = float(sum(n_seq * (PWM_L - int(m) + 1) * strands * 2 * 4 * int(m) for m in lens))
        wn = float(sum(n_seq * (PWM_L - int(m) + 1) * strands for m in lens))
        modes[name] = {"kernel_ms": mm, "windows_per_s": wn / (mm * 1e-3), "mfma_tflops": fl / (mm * 1e-3) / 1e12}
        log("bench: PWM %s: kernels %.2f ms (%.2e windows/s, %.1f TFLOP/s on MFMA)"
            % (name, mm, wn / (mm * 1e-3), fl / (mm * 1e-3) / 1e12))
    pv = float(np.float32(2.7e-4))
    th_ms, th_wall = [], []
    for _ in range(3):
        t0 = time.perf_counter()
        th = cdlm.get_thresh((pwms, lens), pv)
        th_wall.append(1000.0 * (time.perf_counter() - t0))
        th_ms.append(cdlm.pvalue_last_ms())
    log("bench: p-value thresholds of %d PWMs: kernels %.3f ms, call %.2f ms" % (PWM_K, float(np.median(th_ms)),
                                                                               float(np.median(th_wall))))
    return {"workload": "PWM scan: %d sequences x L=%d, %d PWMs of length 8-24, both strands, hit list" % (n_seq, PWM_L, PWM_K),
            "kernel_ms": k_ms, "windows_per_s": windows / (k_ms * 1e-3), "hits": n_hits,
            "best_window_modes": modes,
            "pvalue_thresholds": {"pwms": PWM_K, "pvalue": pv, "kernel_ms": float(np.median(th_ms)),
                                  "call_ms": float(np.median(th_wall))},
            "roofline": {"bound": "mfma", "achieved": tf, "peak": MFMA_FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": tf / MFMA_FP32_PEAK_TFLOPS, "algorithmic_flops": flops}}


# ------------------------------------------------------------------------- main
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=N_SEQ, help="sequences (default: the workload's 20000)")
    ap.add_argument("--budget-s", type=float, default=430.0,
                    help="seconds from the start of the process within which the legs after the headline run")
    ap.add_argument("--cpu-sample", type=int, default=384)
    ap.add_argument("--cpu-pwm-seqs", type=int, default=10000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-datasets", type=int, default=256, help="datasets in the batched leg (0: skip it)")
    ap.add_argument("--plain-datasets", type=int, default=64,
                    help="datasets of the batched leg also run through the plain loop (and served)")
    ap.add_argument("--no-topology", dest="topology", action="store_false", help="skip the activation-topology leg")
    ap.add_argument("--no-ingest", dest="ingest", action="store_false", help="skip the problem-creation (ingest) leg")
    ap.add_argument("--pwm-seqs", type=int, default=1000000, help="sequences in the PWM-scan leg (0: skip it)")
    ap.add_argument("--serve-rates", default="2", help="comma-separated job arrival rates of the served leg (jobs/s)")
    ap.add_argument("--serve-policies", default="fifo,shortest", help="server admission orders to run")
    ap.add_argument("--serve-jobs", type=int, default=64, help="jobs per served run (0: skip the served leg)")
    ap.add_argument("--small-datasets", type=int, default=64,
                    help="datasets with N in [200, 2000] in the small-dataset batched leg (0: skip it)")
    ap.add_argument("--wave-frac", type=float, default=0.25,
                    help="batched legs: device memory per wave, as a fraction of the free memory")
    ap.add_argument("--reference-loop", type=int, default=2, metavar="CSC_ITERS",
                    help="CSC iterations timed in the reference-structured plain-loop leg (0: skip it)")
    ap.add_argument("--reference-batch", type=int, default=4,
                    help="datasets of the single-dataset shape in that leg's batched comparison (0: skip it)")
    ap.add_argument("--outer", type=int, default=ITER_NUM, help="outer iterations per step (profiling runs only)")
    ap.add_argument("--calibrate-copy", type=float, default=0.0, metavar="GIB",
                    help="only run a plain device copy of GIB GiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X and has no multi-GPU path; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    budget = Budget(args.budget_s)

    import numpy as np
    import cdlm
    from cdlm import synth

    if args.calibrate_copy > 0:
        import ctypes as C
        from cdlm import _lib
        ctx = cdlm.default_context()
        nbytes = int(args.calibrate_copy * (1 << 30))
        ms = C.c_double()
        _lib.check(_lib.load().cdlm_copy_bench(ctx.h, nbytes, 5, C.byref(ms)), ctx.h)
        print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "ms_per_copy": ms.value,
                          "GBps": 2 * nbytes / (ms.value * 1e-3) / 1e9, "reps": 5, "warmup_launches": 1}))
        return

    if args.reference_loop > 0:
        # torch carries its own HIP runtime; it has to open the device before libcdlm's
        # runtime does, or it finds no device later in the process
        import torch
        torch.cuda.init()
    N = args.n
    Lg = 4 * L_BASES + 2 * FLEN
    fl = [FLEN] * M_FILT
    bases, _ = synth.c2_dataset(seed=2, N=N, L=L_BASES)
    s = synth.onehot(bases)
    D0 = synth.rand_probvec_filters(7, M_FILT, Lg, fl)
    log("bench: building problem N=%d L=%d M=%d filter_len=%d" % (N, L_BASES, M_FILT, FLEN))
    # launch choices for the shape (FFT stage orders, dictionary batch, CSC work-group shape),
    # measured once before any timing, as the reference builds its FFT plans up front
    t0 = time.perf_counter()
    tuning = cdlm.autotune(L_BASES, M_FILT, FLEN)
    log("bench: autotuned %s in %.2f s" % (tuning, time.perf_counter() - t0))
    c = cdlm.CDL(s, fl, RHO, SIGMA, LAM, D0=D0, X0=None, seed=1234)

    def step():
        c.reset()
        c.optimize(args.outer, CSC_ITER, CDL_ITER)
        return c.code_thresh(1.0)

    for i in range(args.warmup):
        t0 = time.perf_counter()
        step()
        log("bench: warmup step %d %.3f s" % (i, time.perf_counter() - t0))
    c.timing(True)
    t0 = time.perf_counter()
    for i in range(args.steps):
        th, cnt = step()
    dt = time.perf_counter() - t0
    csc_ms, csc_n = c.timing_read(0)
    dict_ms, dict_n = c.timing_read(1)
    best_ms, best_n = c.timing_read(2)
    red_ms, red_n = c.timing_read(3)
    c.timing(False)
    ms_step = 1000.0 * dt / args.steps
    iters = args.outer * CSC_ITER
    value = N * iters / (dt / args.steps)
    avg_csc = csc_ms / max(csc_n, 1)                 # per CSC iteration
    csc_launches = max(csc_n // CSC_ITER, 1)          # one launch per outer iteration
    launch_ms = csc_ms / csc_launches
    avg_dict = dict_ms / max(dict_n, 1)
    b_launch = compulsory_csc_bytes(N, Lg, M_FILT) * CSC_ITER
    b_carry = carry_csc_bytes(N, Lg, M_FILT, CSC_ITER)
    achieved = b_launch / (launch_ms * 1e-3) / 1e9
    achieved_carry = b_carry / (launch_ms * 1e-3) / 1e9
    b_step = args.outer * (CSC_ITER * compulsory_csc_bytes(N, Lg, M_FILT) + CDL_ITER * compulsory_dict_bytes(N, Lg, M_FILT))
    b_step_floor = args.outer * (b_carry + CDL_ITER * dict_floor_bytes(N, Lg, M_FILT))
    b_dict = compulsory_dict_bytes(N, Lg, M_FILT)
    b_dict_floor = dict_floor_bytes(N, Lg, M_FILT)
    traffic, tsrc = None, None
    tpath = os.path.join(ROOT, "profiles", "csc_c2_traffic.json")
    if os.path.exists(tpath) and N == N_SEQ:
        try:
            traffic = json.load(open(tpath)).get("bytes_per_launch")
            tsrc = ("committed figure, profiles/csc_c2_traffic.json: rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, "
                    "per k_csc dispatch, scaled by a plain 2 GiB copy's counters (not this run's counters)")
        except Exception:
            traffic = None
    log("bench: %d steps %.3f s/step; CSC %.3f ms x %d, dict %.3f ms x %d, best %.3f ms x %d, reduce %.3f ms x %d;"
        " code threshold %.5f (%d positive codes)" % (args.steps, dt / args.steps, avg_csc, csc_n, avg_dict, dict_n,
                                                      best_ms / max(best_n, 1), best_n, red_ms / max(red_n, 1), red_n,
                                                      th, cnt))
    out = {
        "metric": "CSC-ADMM seq*iters/s at N=20k,L=200,K=64",
        "value": value,
        "unit": "seq*iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic planted-motif sequences (5 Dirichlet(0.5) motifs, 20% occupancy), random init",
        "config": {"workload": "single-dataset CDL: N=%d L=%d M=%d filter_len=%d Lg=%d lambda=%g rho=%g sigma=%g, "
                               "%dx(%d CSC + update_best + %d dict) = %d ADMM iterations per step"
                               % (N, L_BASES, M_FILT, FLEN, Lg, LAM, RHO, SIGMA, ITER_NUM, CSC_ITER, CDL_ITER, iters),
                   "N": N, "L": L_BASES, "M": M_FILT, "filter_len": FLEN, "iterations": iters,
                   "launch_choices": tuning},
        "roofline": {"bound": "hbm", "kernel": "k_csc (%d CSC iterations per launch)" % CSC_ITER,
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "traffic_source": tsrc,
                     "algorithmic_bytes_per_launch": b_launch,
                     "algorithmic_basis": "4 N Lg (6M + 7) bytes per CSC iteration (X, y2, y3 read and written, y1, "
                                          "Y1, DX read and written, the data read) x %d iterations" % CSC_ITER,
                     "note": "frac above 1 means the kernel moves fewer bytes than that compulsory count: it carries "
                             "its state between the iterations of a launch as (X + y2, X + y3) and keeps y1, Y1, DX "
                             "on chip; carry_floor and traffic give the bytes it does move",
                     "avg_launch_ms": launch_ms, "launches": csc_launches, "iterations_per_launch": CSC_ITER,
                     "ms_per_iteration": avg_csc,
                     "carry_floor": {"bytes_per_launch": b_carry, "achieved": achieved_carry,
                                     "frac": achieved_carry / HBM_PEAK_GBS,
                                     "basis": "the kernel's own floor: the state carried between the iterations of "
                                              "a launch as (X + y2, X + y3), 4 N Lg ((5.5 + 4 (I - 1)) M + 9)"},
                     "step": {"algorithmic_bytes": b_step, "seconds": dt / args.steps,
                              "achieved": b_step / (dt / args.steps) / 1e9,
                              "frac": b_step / (dt / args.steps) / 1e9 / HBM_PEAK_GBS,
                              "floor_bytes": b_step_floor,
                              "floor_frac": b_step_floor / (dt / args.steps) / 1e9 / HBM_PEAK_GBS,
                              "basis": "%d x (%d CSC iterations + %d dictionary iterations) of compulsory bytes "
                                       "(4 N Lg (6M + 7) and 4 N Lg (8M + 4)), and of the kernels' own floors; the "
                                       "whole step's wall time, initialisation and update_best included"
                                       % (args.outer, CSC_ITER, CDL_ITER)},
                     "k_dict": {"ms": avg_dict, "algorithmic_bytes": b_dict,
                                "frac": b_dict / (avg_dict * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                "floor_bytes": b_dict_floor,
                                "floor_frac": b_dict_floor / (avg_dict * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                "basis": "4 N Lg (8M + 4) compulsory bytes; floor: X read, the g1 and stored Y3 half "
                                         "spectra read and written, per-sequence G2, g2, data, Td and kappa"}},
        "kernels_ms": {"csc": avg_csc, "dict": avg_dict, "best": best_ms / max(best_n, 1),
                       "dict_reduce": red_ms / max(red_n, 1)},
    }

    # ---- the other legs, in priority order within the time budget
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_sample, args.cpu_pwm_seqs)
    else:
        out["cpu_baseline"] = None
    if args.reference_loop > 0 and budget.fits("reference_loop", 40):
        out["reference_loop"] = reference_loop_leg(s, fl, D0, dt / args.steps, args.reference_batch,
                                                   args.reference_loop)
    if args.topology and budget.fits("topology", 15):
        out["topology"] = topology_leg(c)
    c.close()
    del c
    if args.ingest and budget.fits("ingest", 5):
        out["ingest"] = ingest_leg(bases, s, fl, D0)
    del s
    if args.pwm_seqs > 0 and budget.fits("pwm", 25 + 20e-6 * args.pwm_seqs):
        out["pwm"] = pwm_leg(args.pwm_seqs)
    if args.batch_datasets > 0 and budget.fits("batched", 60):
        n_plain = min(args.plain_datasets, args.batch_datasets)
        cs = C3Set(args.batch_datasets)
        out["batched"] = batched_leg(cs, n_plain, budget, wave_frac=args.wave_frac,
                                     keep_plain=n_plain if args.serve_jobs > 0 else 0)
        if args.serve_jobs > 0:
            _, t_each, _, _ = batched_leg.cache
            est = 1.2 * sum(t_each[:args.serve_jobs]) + 15
            served = []
            for pol in [p for p in args.serve_policies.split(",") if p]:
                for rate in [float(r) for r in args.serve_rates.split(",") if r]:
                    if budget.fits("served_%s_%g" % (pol, rate), est):
                        settle("the previous leg")
                        served.append(served_leg(rate, args.serve_jobs, pol))
            if served:
                out["served"] = served
        del cs
    if args.small_datasets > 0 and budget.fits("batched_small", 25):
        settle("the previous leg")
        out["batched_small"] = batched_leg(C3Set(args.small_datasets, N_range=(200, 2000)), args.small_datasets, budget,
                                           wave_frac=args.wave_frac)
    out["skipped"] = budget.skipped
    out["bench_seconds"] = time.perf_counter() - T_START
    print(json.dumps(out))


if __name__ == "__main__":
    main()
