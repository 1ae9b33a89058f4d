"""Writes tests/golden/c1.npz: the CPU-runnable workload exactly as configured, through
thresholding, the activation topology and the pair PFMs, by the oracle in FP32 and FP64.

Workload: cdlm.synth.c1_dataset(seed=1): N=500 sequences of L=100 iid ACGT bases, 2 planted
motifs of length 10 and 12 (Dirichlet(0.5) columns), each in 30% of the sequences at a uniform
start, reverse-complemented with probability 0.5; M=10 filters of filter_size 4 (flattened
length 16, Lg=432), lambda 0.05, rho 1, sigma 100 (the reference's default, WRAPPER_setup.jl:5),
10 x (10 CSC + update_best + 1 dict) = 100 ADMM iterations, num_train = 2 trainings (the
reference's default, WRAPPER_setup.jl:12) from seed-fixed inits: training t has
D0 = synth.rand_probvec_filters(11 + t, ...) and X0 ~ N(0, 1) from numpy's default_rng(21 + t).

Stored per precision and training: the objective trace, G1, G3, code_thresh and the positive
count, and the thresholded codes (thresh_summary.summary, the whole active set).  From the FP64
trainings also: the record's filters (uinds, training numbers, lengths; record_training_result!,
src/SEARCH/store.jl:1-40), the topology counts of get_spatial_info! (store.jl:64-144), the pairs
whose overlap count exceeds count_threshold = 50 (extractions.jl:27-29) and their pair PFMs
(extractions.jl:18-100, with the FP64 codes rounded to Float32).

Run:  python tests/golden/make_c1.py     (about a minute per precision on 8 host cores)
"""
from __future__ import annotations

import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "cdlmotif.jl_amd"), HERE]

N, L, M, FILTER_SIZE = 500, 100, 10, 4
FLEN = 4 * FILTER_SIZE
ITER_NUM, CSC_ITER, CDL_ITER = 10, 10, 1
LAM, RHO, SIGMA = 0.05, 1.0, 100.0
TRAININGS = 2
COUNT_THRESHOLD = 50
BAND = 1e-3


def inputs():
    from cdlm import synth
    bases, _ = synth.c1_dataset(seed=1, N=N, L=L)
    s = synth.onehot(bases)
    Lg = 4 * L + 2 * FLEN
    inits = [(synth.rand_probvec_filters(11 + t, M, Lg, [FLEN] * M),
              np.random.default_rng(21 + t).standard_normal((N, M, Lg)).astype(np.float32)) for t in range(TRAININGS)]
    return s, inits


def digest(a) -> str:
    return hashlib.sha1(np.ascontiguousarray(a).tobytes()).hexdigest()


def record_of(actives, xss):
    """The activation record of the trainings' active sets (N, M, L): start positions (N, K, C),
    uinds, training numbers, lengths and the codes (T, N, M, C)."""
    sp, u, tn, fls = [], [], [], []
    for t, act in enumerate(actives):
        uinds = [m for m in range(M) if act[:, m, :].any()]
        sp.append(act[:, uinds, :])
        u += [m + 1 for m in uinds]
        tn += [t + 1] * len(uinds)
        fls += [FILTER_SIZE] * len(uinds)
    return (np.concatenate(sp, axis=1).astype(np.uint8), np.asarray(u, np.int32), np.asarray(tn, np.int32),
            np.asarray(fls, np.int32), np.stack(xss).astype(np.float32))


def main():
    import oracle as O
    from thresh_summary import summary, xbest_slice
    s, inits = inputs()
    out = {"digest_s": digest(s)}
    for t, (D0, X0) in enumerate(inits):
        out["digest_D0_%d" % t] = digest(D0)
        out["digest_X0_%d" % t] = digest(X0)
    acts64, xss64 = [], []
    for prec in ("f32", "f64"):
        for t, (D0, X0) in enumerate(inits):
            t0 = time.time()
            o = O.OracleCDL(s, D0, X0, [FLEN] * M, RHO, SIGMA, LAM, prec)
            trace = []
            for _ in range(ITER_NUM):
                o.code_update(CSC_ITER)
                trace.append(o.update_best()[0])
                o.dict_update(CDL_ITER)
            k = "%s_%d" % (prec, t)
            out["trace_" + k] = np.asarray(trace, np.float64)
            out["G1_" + k] = np.asarray(o.get("G1"))
            out["G3_" + k] = np.asarray(o.get("G3"))
            th, cnt = o.code_thresh(1.0)
            xb = o.get("Xbest")
            summ = summary(xb, o.gamma, 4 * L, th, band=BAND, n_xs=2, n_act=N)
            assert summ["count"] == cnt
            for kk, v in summ.items():
                out["%s_%s" % (kk, k)] = v
            out["thresh_" + k] = np.float64(th)
            if prec == "f64":
                xs = xbest_slice(xb, o.gamma, 4 * L)
                acts64.append(xs > th)
                xss64.append(xs)
            print("%s training %d: %.1f s, thresh %.9g, positive %d, active %d, uinds %s, band %d" % (
                prec, t, time.time() - t0, th, cnt, int(summ["actcnt"].sum()), summ["uinds"].tolist(),
                len(summ["band_idx"])), flush=True)
            del o, xb
    sp, u, tn, fls, codes = record_of(acts64, xss64)
    maxfl, apart = int(fls.max()), L - int(fls.min()) - 1
    over, non, nxt = O.topology(sp, tn, fls, np.ones((TRAININGS, N), np.int32), maxfl, apart)
    out.update(rec_uinds=u, rec_train=tn, rec_lens=fls, top_over=over, top_non=non, top_next=nxt)
    idx = np.argwhere(over > COUNT_THRESHOLD)              # (d, fc, fr) rows in the reference's order
    pairs = (np.stack([idx[:, 2] + 1, idx[:, 1] + 1, idx[:, 0] + 1], axis=1).astype(np.int32) if len(idx)
             else np.zeros((0, 3), np.int32))
    span, pfm = O.pair_pfms(sp, codes, u, tn, fls, s, pairs, 2 * FILTER_SIZE - 1) if len(pairs) else (
        np.zeros(0, np.int32), np.zeros((0, 4, 2 * FILTER_SIZE - 1), np.float32))
    out.update(pairs=pairs, pair_span=span, pair_pfms=pfm)
    print("record K=%d, active entries %d, topology counts %d/%d/%d, pairs over %d: %d" % (
        len(u), int(sp.sum()), int(over.sum()), int(non.sum()), int(nxt.sum()), COUNT_THRESHOLD, len(pairs)), flush=True)
    np.savez_compressed(os.path.join(HERE, "c1.npz"), **out)


if __name__ == "__main__":
    main()
