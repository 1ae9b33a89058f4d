"""Writes tests/golden/c2_slice.npz: the oracle (FP32 and FP64) on a 1000-sequence dataset of the
single-dataset configuration's shape for the whole schedule, 10 x (20 CSC + update_best + 1 dict)
= 200 ADMM iterations, so that the GPU parity test can compare a full-length run without
running the oracle at test time (about 25 minutes per precision on 8 host cores).

Inputs (rebuilt by the test from the same generators and seeds): the planted-motif dataset of
cdlm.synth.c2_dataset(seed=2, N=1000, L=200), 64 filters of length 20 from
synth.rand_probvec_filters(7, ...), X0 ~ N(0, 1) from numpy's default_rng(8); lambda 0.05,
rho 1, sigma 100.

Stored, for both precisions: SHA-1 digests of the inputs, the objective traces, the final G1,
G3, D_best and DX of the first 16 sequences, and what the codes give after thresholding
(cdl_optimize, src/WRAPPER/wrap.jl:12, and record_training_result!, src/SEARCH/store.jl:15-19):
code_thresh = std(X_best[X_best .> 0]) with the count, sum and sum of squares of the positive
entries, the per-(sequence, filter) number of active codes (X_best[gamma+1:4:gamma+4L] > thresh)
and uinds (the filters with any active code).  From the FP64 oracle also: the X_best slice of
the first N_XS sequences, the packed active set of the first N_ACT sequences, and every code of
the 1000 sequences within BAND (relative) of the threshold (index and value), which are the only
entries whose activity an FP32 computation can legitimately decide differently.

Run:  python tests/golden/make_c2_slice.py
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

N, L, M, FLEN = 1000, 200, 64, 20
ITER_NUM, CSC_ITER, CDL_ITER = 10, 20, 1
LAM, RHO, SIGMA = 0.05, 1.0, 100.0
N_DX = 16
N_XS = 2       # sequences whose X_best slice is stored (FP64)
N_ACT = 64     # sequences whose active set is stored bit for bit (FP64)
BAND = 1e-3    # |x - thresh| <= BAND * thresh: the near-threshold entries stored with their values


def inputs():
    from cdlm import synth
    bases, _ = synth.c2_dataset(seed=2, N=N, L=L)
    s = synth.onehot(bases)
    Lg = 4 * L + 2 * FLEN
    D0 = synth.rand_probvec_filters(7, M, Lg, [FLEN] * M)
    X0 = np.random.default_rng(8).standard_normal((N, M, Lg)).astype(np.float32)
    return s, D0, X0


def digest(a) -> str:
    return hashlib.sha1(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    import oracle as O
    from thresh_summary import summary
    s, D0, X0 = inputs()
    out = {"digest_s": digest(s), "digest_D0": digest(D0), "digest_X0": digest(X0)}
    for prec in ("f32", "f64"):
        t0 = time.time()
        o = O.OracleCDL(s, D0, X0, [FLEN] * M, RHO, SIGMA, LAM, prec)
        trace = []
        for _ in range(ITER_NUM):
            o.code_update(CSC_ITER)
            trace.append(o.update_best()[0])
            o.dict_update(CDL_ITER)
        out["trace_" + prec] = np.asarray(trace, np.float64)
        for f in ("G1", "G3", "Dbest"):
            out["%s_%s" % (f, prec)] = np.asarray(o.get(f))
        out["DX_" + prec] = np.asarray(o.get("DX"))[:N_DX]
        th, cnt = o.code_thresh(1.0)
        xb = o.get("Xbest")
        summ = summary(xb, o.gamma, 4 * L, th, band=BAND if prec == "f64" else None, n_xs=N_XS if prec == "f64" else 0,
                       n_act=N_ACT if prec == "f64" else 0)
        assert summ["count"] == cnt, (summ["count"], cnt)
        for k, v in summ.items():
            out["%s_%s" % (k, prec)] = v
        out["thresh_" + prec] = np.float64(th)
        del xb
        print("%s: %.0f s, trace %s, thresh %.9g, positive %d, uinds %d, band %s" % (
            prec, time.time() - t0, np.array2string(out["trace_" + prec], precision=6), th, cnt,
            len(summ["uinds"]), len(summ.get("band_idx", []))), flush=True)
        del o
    np.savez_compressed(os.path.join(HERE, "c2_slice.npz"), **out)


if __name__ == "__main__":
    main()
