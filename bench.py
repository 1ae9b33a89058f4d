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
fused CSC iteration, HBM-bound; per-launch average from HIP events on the library's
stream over the timed steps) and a cpu_baseline object (the CPU oracle timed on a bounded
slice of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cdlmotif.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

N_SEQ, L_BASES, M_FILT, FLEN = 20000, 200, 64, 20
ITER_NUM, CSC_ITER, CDL_ITER = 10, 20, 1
LAM, RHO, SIGMA = 0.05, 1.0, 100.0
HBM_PEAK_GBS = 8000.0


def csc_bytes(N, Lg, M, iters):
    """Algorithmic HBM bytes of one CSC launch of `iters` iterations, per (l, n), 4 bytes each:
    the state enters as Xt and y2 (2M) and y3 off the stride-4 support (0.75M; on the support
    y3 is identically zero) and leaves the same way; between two iterations of the launch it
    is carried as the two shrink/projection inputs u2 = X + y2 and u3 = X + y3 (2M written,
    2M read).  So 5.5M + 4M (iters - 1) in all; once per launch y1, Y1, DX, h and the data
    read and h, DX, y1, Y1 written (9), which stay on chip between iterations."""
    return 4.0 * N * Lg * ((5.5 + 4.0 * (iters - 1)) * M + 9)


def cpu_baseline(sample_n: int):
    """The oracle (reference-structured FP32 restatement, OpenMP over columns) on a
    sample_n-sequence slice of the workload for one outer iteration (20 CSC iterations,
    update_best, 1 dictionary iteration): the same work per sequence-iteration as a step."""
    import numpy as np
    import oracle as O
    from cdlm import synth
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
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": sample_n * CSC_ITER / dt, "unit": "seq*iters/s", "cores": threads, "kind": "port",
            "host_cpus": os.cpu_count(), "cpu_model": model,
            "sample": "oracle f32 reference-structured CDL on %d sequences of the same shape (L=%d, M=%d, "
                      "filter_len=%d), one outer iteration (%d CSC + update_best + %d dict) in %.1f s"
                      % (sample_n, L_BASES, M_FILT, FLEN, CSC_ITER, CDL_ITER, dt)}


C3_ITER_NUM, C3_CSC_ITER, C3_CDL_ITER = 10, 15, 1

# Freeing device memory starts work in the driver that slows the device for a few seconds
# afterwards (measured: about 1 s per 50 GB freed; the next batch ran 2-4 s slower after 228 GB
# were freed, and at its normal speed after a 6 s pause).  The legs pause this long, outside
# their timed regions, after steps that free tens of GB, so that one leg's frees do not land in
# the next leg's timing.
SETTLE_S = 6.0


def settle(log, what):
    log("bench: %.0f s pause after %s" % (SETTLE_S, what))
    time.sleep(SETTLE_S)


def batched_leg(n_datasets: int, log, N_range=(200, 20000), cache=True, wave_frac=0.25):
    """Batched workload sample: n_datasets datasets drawn as in the batched configuration
    (N in [200, 20000], L in {100..500}, M in {32, 64}, filter_len in {16, 20, 24}), each
    trained once (10 x (15 CSC + update_best + 1 dict) = 150 ADMM iterations, code
    threshold).  Plain: one dataset at a time through the single-problem entry points.
    Batched: datasets packed by device footprint into waves of at most wave_frac of the free
    device memory; one batch training (a captured HIP graph) per wave.  Both include problem
    creation (data upload) and teardown."""
    import numpy as np
    import cdlm
    from cdlm import synth
    specs = synth.c3_specs(n_datasets, N_range=N_range)
    data = [synth.c3_onehot(sp) for sp in specs]
    ctx = cdlm.default_context()

    def make(i):
        sp = specs[i]
        return cdlm.CDL(data[i], [sp["filter_len"]] * sp["M"], RHO, SIGMA, LAM, seed=sp["seed"], ctx=ctx)

    # launch choices per shape (radix order, dictionary batch, CSC slots), outside the timing
    shapes = sorted({(sp["L"], sp["M"], sp["filter_len"]) for sp in specs})
    t0 = time.perf_counter()
    tunings = {shape: cdlm.autotune(*shape, ctx=ctx) for shape in shapes}
    log("bench: autotuned %d shapes in %.2f s" % (len(shapes), time.perf_counter() - t0))
    # first use of every kernel variant outside the timed regions
    warm = [cdlm.CDL(np.zeros((4, 4 * 8), np.float32), [f] * 2, RHO, SIGMA, LAM, seed=1, ctx=ctx) for f in (16, 20, 24)]
    wb = cdlm.Batch(warm)
    wb.train(1, 1, 1)
    wb.close()
    for w in warm:
        w.optimize(1, 1, 1)
        w.close()

    t0 = time.perf_counter()
    th_plain, t_each = [], []
    for i in range(len(specs)):
        t1 = time.perf_counter()
        c = make(i)
        c.optimize(C3_ITER_NUM, C3_CSC_ITER, C3_CDL_ITER)
        th_plain.append(c.code_thresh(1.0)[0])
        c.close()
        t_each.append(time.perf_counter() - t1)
    t_plain = time.perf_counter() - t0
    log("bench: batched workload, plain loop: %d datasets in %.2f s" % (len(specs), t_plain))
    settle(log, "the plain loop")

    free, _total = ctx.mem_info()
    budget = int(free * wave_frac)
    sizes = [cdlm.footprint_bytes(sp["L"], sp["N"], sp["M"], sp["filter_len"]) for sp in specs]
    waves = cdlm.pack_by_footprint(sizes, budget)
    th_batch = [None] * len(specs)
    t0 = time.perf_counter()
    for w in waves:
        probs = [make(i) for i in w]
        b = cdlm.Batch(probs)
        _tr, th, _cnt = b.train(C3_ITER_NUM, C3_CSC_ITER, C3_CDL_ITER, 1.0)
        b.close()
        for p in probs:
            p.close()
        for j, i in enumerate(w):
            th_batch[i] = float(th[j])
    t_batch = time.perf_counter() - t0
    log("bench: batched workload, batched: %d waves, %.2f s" % (len(waves), t_batch))
    seq_iters = sum(sp["N"] for sp in specs) * C3_ITER_NUM * C3_CSC_ITER
    if cache:
        batched_leg.cache = (specs, data, t_each, th_plain, tunings)
    return {
        "workload": "%d datasets of the batched configuration (seeds 3..%d, N in [%d, %d]), 150 ADMM iterations each"
                    % (len(specs), 2 + len(specs), N_range[0], N_range[1]),
        "datasets": len(specs), "waves": len(waves), "wave_budget_bytes": budget, "wave_fraction_of_free": wave_frac,
        "plain_datasets_per_s": len(specs) / t_plain, "batched_datasets_per_s": len(specs) / t_batch,
        "speedup": t_plain / t_batch,
        "plain_seq_iters_per_s": seq_iters / t_plain, "batched_seq_iters_per_s": seq_iters / t_batch,
        "thresholds_identical": bool(np.array_equal(np.array(th_plain), np.array(th_batch), equal_nan=True)),
    }


def served_leg(rate: float, n_jobs: int, log, policy: str = "fifo"):
    """Served workload sample: the first n_jobs datasets of the batched leg arrive as a
    Poisson process of the given rate (seeded) and are submitted to the library's server
    (continuous batching, admission by device footprint, one outer iteration per tick).
    Latency = arrival -> results in host memory.  The plain per-job loop's latencies follow
    from the measured alone times of the same datasets (FIFO, one job at a time)."""
    import threading
    import numpy as np
    import cdlm
    specs, data, t_each, th_plain, tunings = batched_leg.cache
    n = min(n_jobs, len(specs))
    # the launch choices the plain results were computed with (a later leg may have re-tuned
    # a shape; the FFT stage order changes rounding, so another choice changes result bits)
    for shape, choice in tunings.items():
        cdlm.tuning_set(*shape, choice)
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
            sp = specs[i]
            ids[i] = srv.submit(data[i], [sp["filter_len"]] * sp["M"], RHO, SIGMA, LAM, C3_ITER_NUM, seed=sp["seed"])

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
    # plain per-job loop, FIFO from the same arrivals with the measured alone times
    done, plat = 0.0, []
    for i in range(n):
        done = max(done, arrive[i]) + t_each[i]
        plat.append(done - arrive[i])
    lat, plat = np.array(lat), np.array(plat)
    q = lambda v, p: float(np.sort(v)[max(0, int(np.ceil(p * len(v))) - 1)])
    log("bench: served (%s) %d jobs at %.1f jobs/s: p50 %.2f s p99 %.2f s, %.2f datasets/s (plain loop p50 %.2f s "
        "p99 %.2f s, %.2f datasets/s); service time sum %.2f s vs alone %.2f s"
        % (policy, n, rate, q(lat, 0.5), q(lat, 0.99), n / (t_end - t0), q(plat, 0.5), q(plat, 0.99), n / float(done),
           sum(svc), sum(t_each[:n])))
    return {
        "workload": "%d batched-configuration datasets arriving as Poisson(%.1f jobs/s), 150 ADMM iterations each"
                    % (n, rate), "policy": policy,
        "jobs": n, "arrival_rate": rate, "p50_latency_s": q(lat, 0.5), "p99_latency_s": q(lat, 0.99),
        "datasets_per_s": n / (t_end - t0), "ticks": st["ticks"],
        "plain_loop_p50_latency_s": q(plat, 0.5), "plain_loop_p99_latency_s": q(plat, 0.99),
        "plain_loop_datasets_per_s": n / float(done),
        "results_identical_to_alone": bool(same),
    }


def reference_loop_leg(s, fl, D0, t_fused_step, n_batch, csc_sample, log):
    """The plain per-dataset GPU loop, as the reference runs it, at the single-dataset
    configuration, against the batched form on the same shape.

    Plain loop: cdlm.refloop (the reference's code_update!/dict_update!/update_best!
    statement for statement on torch + hipFFT, a device synchronisation at each CUDA.@sync)
    on the full N=20000 problem.  Bounded sample: after one untimed outer iteration (FFT
    plan creation), csc_sample CSC iterations, update_best and one dictionary iteration are
    timed; a training is 10 x (20 CSC + update_best + 1 dict), extrapolated from those
    per-call times.  Initialisation and the code threshold are left out, which favours the
    plain loop.  Batched: n_batch datasets of the same shape (other seeds) trained in one
    batch (a captured HIP graph), problem creation and teardown included."""
    import torch
    import cdlm
    from cdlm import synth
    from cdlm.refloop import ReferenceStructuredCDL
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
    settle(log, "freeing the reference-structured loop's %.0f GB" % peak)
    t_train = ITER_NUM * (CSC_ITER * t_csc + t_best + CDL_ITER * t_dict)
    log("bench: reference-structured loop at N=%d: CSC %.1f ms, update_best %.1f ms, dict %.1f ms per call "
        "(init %.2f s, peak %.0f GB) -> %.2f s per training"
        % (N, 1e3 * t_csc, 1e3 * t_best, 1e3 * t_dict, t_init, peak, t_train))
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
        settle(log, "the same-shape batch")
        out["batched"] = {"datasets": n_batch, "seconds": t_b, "datasets_per_s": n_batch / t_b,
                          "vs_plain_loop": (n_batch / t_b) * t_train}
    return out


def topology_leg(c, log):
    """Activation record and topology (record_training_result!, get_spatial_info!,
    src/SEARCH/store.jl) for two trainings at the single-dataset configuration: the timed
    problem's last training and a second training from another seed; codes over each
    training's std threshold (wrap.jl:12), filter_size 5 bases.  Kernel time from HIP events;
    the kernel's work is the pairs of active entries of each sequence."""
    import ctypes as Cc
    import numpy as np
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
    t0 = time.perf_counter()
    over, non, nxt = rec.topology()
    t_topo = time.perf_counter() - t0
    v = Cc.c_double()
    _lib.check(_lib.load().cdlm_topology_last_ms(ctx.h, Cc.byref(v)), ctx.h)
    K, T, C, N = rec.dims()
    act = int(rec.start_positions(0, min(N, 2000)).sum()) * N / min(N, 2000)
    # pair PFMs of the pairs overlapping in more than count_threshold = 50 sequences
    # (extractions.jl:27; WRAPPER_setup.jl default)
    data = synth.onehot(bases)
    t0 = time.perf_counter()
    pairs, spans, pfms = cdlm.get_pair_pfms(rec, over, 50, data, FLEN // 4)
    t_pfm = time.perf_counter() - t0
    rec.close()
    # pair visits: the active entries of each sequence against each other (estimated from
    # the mean per sequence); the reference's kernel visits (C K)^2 per sequence
    npairs = (act / N) ** 2 * N
    log("bench: topology K=%d (2 trainings), ~%.0f active entries, kernel %.2f ms (%.2e pair visits/s), record %.2f s;"
        " %d pair PFMs in %.3f s" % (K, act, v.value, npairs / (v.value * 1e-3), t_add, len(pairs), t_pfm))
    return {"workload": "activation topology of 2 trainings at N=%d L=%d M=%d (filter_size %d)" % (N, C, M_FILT, FLEN // 4),
            "K": K, "active_entries": act, "kernel_ms": v.value, "record_s": t_add,
            "pair_visits_per_s": npairs / (v.value * 1e-3), "reference_pair_visits": float(C * K) ** 2 * N,
            "counts": {"overlap": int(over.sum()), "non_overlap": int(non.sum()), "next_to": int(nxt.sum())},
            "pair_pfms": {"pairs": int(len(pairs)), "seconds": t_pfm,
                          "note": "wall time of the call, including the 64 MB data upload"}}


PWM_L, PWM_K = 500, 512
MFMA_FP32_PEAK_TFLOPS = 157.3


def ingest_leg(bases, s, fl, D0, log):
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


def pwm_leg(n_seq: int, log):
    """PWM scan sample: n_seq sequences of L=500 iid bases (GC 0.42) against 512 log-odds
    PWMs of length 8-24 (the PWM-scan configuration at a bounded N), both strands,
    threshold 0.8 x best score, full hit list; then the argmax and best-hit modes on the same
    workload.  Kernel time from HIP events around the MFMA kernels; algorithmic flops of the
    hit list = sum_k N (L - m_k + 1) x 2 strands x 2 x 4 m_k."""
    import ctypes as Cc
    import numpy as np
    import cdlm
    from cdlm import _lib, synth
    bases, pwms, lens, thresh = synth.pwm_workload(5, n_seq, PWM_L, PWM_K)
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
    log("bench: PWM scan %d x %d, %d PWMs: kernel %.2f ms (%.1f TFLOP/s), %d hits, call %.2f s"
        % (n_seq, PWM_L, PWM_K, k_ms, tf, len(h), wall))
    # the best-window scans on the same workload: argmax of the thresholded position-major
    # scores, forward strand (scan_w_gpu!), and best hit per strand pair with nucleotide-major
    # sums (overlapping_scan!, the workload's best-hit mode); algorithmic flops as for the hit
    # list, one product per base and position: 2 x 4 m_k per window and strand
    from cdlm import scan as S
    modes = {}
    for name, fn, passes, strands in (("argmax", S.scan_w_gpu, 1, 1), ("best_hit", S.overlapping_scan, 1, 2)):
        mms = []
        for _ in range(3):
            fn(bases, pwms, lens, thresh)
            v = Cc.c_double()
            _lib.check(_lib.load().cdlm_pwm_last_ms(ctx.h, Cc.byref(v)), ctx.h)
            mms.append(v.value)
        mm = float(np.median(mms))
        fl = float(sum(n_seq * (PWM_L - int(m) + 1) * strands * 2 * 4 * int(m) * passes for m in lens))
        wn = float(sum(n_seq * (PWM_L - int(m) + 1) * strands for m in lens))
        modes[name] = {"kernel_ms": mm, "windows_per_s": wn / (mm * 1e-3), "mfma_tflops": fl / (mm * 1e-3) / 1e12}
        log("bench: PWM %s: kernels %.2f ms (%.2e windows/s, %.1f TFLOP/s on MFMA)"
            % (name, mm, wn / (mm * 1e-3), fl / (mm * 1e-3) / 1e12))
    # score thresholds of the same PWMs for the reference's p-value (get_thresh, pvalue 2.7e-4
    # as Float32): exact score distributions, one work-group per PWM
    pv = float(np.float32(2.7e-4))
    th_ms, th_wall = [], []
    for _ in range(3):
        t0 = time.perf_counter()
        th = cdlm.get_thresh((pwms, lens), pv)
        th_wall.append(1000.0 * (time.perf_counter() - t0))
        th_ms.append(cdlm.pvalue_last_ms())
    log("bench: p-value thresholds of %d PWMs: kernels %.3f ms, call %.2f ms, mean threshold %.3f"
        % (PWM_K, float(np.median(th_ms)), float(np.median(th_wall)), float(th.mean())))
    return {"workload": "PWM scan: %d sequences x L=%d, %d PWMs of length 8-24, both strands, hit list" % (n_seq, PWM_L, PWM_K),
            "kernel_ms": k_ms, "windows_per_s": windows / (k_ms * 1e-3), "hits": int(len(h)),
            "best_window_modes": modes,
            "pvalue_thresholds": {"pwms": PWM_K, "pvalue": pv, "kernel_ms": float(np.median(th_ms)),
                                  "call_ms": float(np.median(th_wall))},
            "roofline": {"bound": "mfma", "achieved": tf, "peak": MFMA_FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": tf / MFMA_FP32_PEAK_TFLOPS, "algorithmic_flops": flops}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=N_SEQ, help="sequences (default: the workload's 20000)")
    ap.add_argument("--cpu-sample", type=int, default=384)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-datasets", type=int, default=16, help="datasets in the batched leg (0: skip it)")
    ap.add_argument("--no-topology", dest="topology", action="store_false", help="skip the activation-topology leg")
    ap.add_argument("--no-ingest", dest="ingest", action="store_false", help="skip the problem-creation (ingest) leg")
    ap.add_argument("--pwm-seqs", type=int, default=100000, help="sequences in the PWM-scan leg (0: skip it)")
    ap.add_argument("--serve-rate", type=float, default=2.0, help="job arrival rate of the served leg (jobs/s)")
    ap.add_argument("--small-datasets", type=int, default=64,
                    help="datasets with N in [200, 2000] in the small-dataset batched leg (0: skip it)")
    ap.add_argument("--wave-frac", type=float, default=0.25,
                    help="batched legs: device memory per wave, as a fraction of the free memory")
    ap.add_argument("--serve-policy", default="fifo", choices=["fifo", "shortest"], help="server admission order")
    ap.add_argument("--serve-jobs", type=int, default=16, help="jobs in the served leg (0: skip it)")
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
    print("bench: building problem N=%d L=%d M=%d filter_len=%d" % (N, L_BASES, M_FILT, FLEN), file=sys.stderr,
          flush=True)
    # launch choices for the shape (FFT stage orders, dictionary batch, CSC work-group shape),
    # measured once before any timing, as the reference builds its FFT plans up front
    t0 = time.perf_counter()
    tuning = cdlm.autotune(L_BASES, M_FILT, FLEN)
    print("bench: autotuned %s in %.2f s" % (tuning, time.perf_counter() - t0), file=sys.stderr, flush=True)
    c = cdlm.CDL(s, fl, RHO, SIGMA, LAM, D0=D0, X0=None, seed=1234)

    def step():
        c.reset()
        c.optimize(args.outer, CSC_ITER, CDL_ITER)
        return c.code_thresh(1.0)

    for i in range(args.warmup):
        t0 = time.perf_counter()
        step()
        print("bench: warmup step %d %.3f s" % (i, time.perf_counter() - t0), file=sys.stderr, flush=True)
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
    achieved = csc_bytes(N, Lg, M_FILT, CSC_ITER) / (launch_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "csc_c2_traffic.json")
    if os.path.exists(tpath) and N == N_SEQ:
        try:
            traffic = json.load(open(tpath)).get("bytes_per_launch")
        except Exception:
            traffic = None
    print("bench: %d steps %.3f s/step; CSC %.3f ms x %d, dict %.3f ms x %d, best %.3f ms x %d, reduce %.3f ms x %d;"
          " code threshold %.5f (%d positive codes)"
          % (args.steps, dt / args.steps, avg_csc, csc_n, dict_ms / max(dict_n, 1), dict_n,
             best_ms / max(best_n, 1), best_n, red_ms / max(red_n, 1), red_n, th, cnt), file=sys.stderr, flush=True)
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
                     "traffic": traffic, "algorithmic_bytes_per_launch": csc_bytes(N, Lg, M_FILT, CSC_ITER),
                     "avg_launch_ms": launch_ms, "launches": csc_launches, "iterations_per_launch": CSC_ITER,
                     "ms_per_iteration": avg_csc},
        "kernels_ms": {"csc": avg_csc, "dict": dict_ms / max(dict_n, 1), "best": best_ms / max(best_n, 1),
                       "dict_reduce": red_ms / max(red_n, 1)},
    }
    if args.topology:
        out["topology"] = topology_leg(c, lambda m: print(m, file=sys.stderr, flush=True))
    c.close()
    del c
    if args.reference_loop > 0:
        out["reference_loop"] = reference_loop_leg(s, fl, D0, dt / args.steps, args.reference_batch,
                                                   args.reference_loop,
                                                   lambda m: print(m, file=sys.stderr, flush=True))
    if args.ingest:
        out["ingest"] = ingest_leg(bases, s, fl, D0, lambda m: print(m, file=sys.stderr, flush=True))
    if args.batch_datasets > 0:
        log = lambda m: print(m, file=sys.stderr, flush=True)
        out["batched"] = batched_leg(args.batch_datasets, log, wave_frac=args.wave_frac)
        if args.small_datasets > 0:
            settle(log, "the batched waves")
            out["batched_small"] = batched_leg(args.small_datasets, log, N_range=(200, 2000), cache=False,
                                               wave_frac=args.wave_frac)
        if args.serve_jobs > 0:
            out["served"] = served_leg(args.serve_rate, args.serve_jobs,
                                       lambda m: print(m, file=sys.stderr, flush=True), args.serve_policy)
    if args.pwm_seqs > 0:
        out["pwm"] = pwm_leg(args.pwm_seqs, lambda m: print(m, file=sys.stderr, flush=True))
    if not args.no_cpu:
        print("bench: CPU baseline on a %d-sequence slice" % args.cpu_sample, file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(args.cpu_sample)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
