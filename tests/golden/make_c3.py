"""Writes tests/golden/c3.npz: datasets of the batched workload at their real shapes (L, M,
filter length, hence padded length) but a reduced number of sequences, run by the oracle in
FP32 and FP64 for the reference's default schedule, 10 x (15 CSC + update_best + 1 dict) = 150
ADMM iterations (WRAPPER_setup.jl:8-10).

Datasets: cdlm.synth.c3_specs(SPECS) (seeds 3..9): padded lengths 432, 2048, 440, 848, 1240,
632 and 2032 = 16 x 127, M in {32, 64}, filter lengths 16, 20 and 24, each with its planted
motifs, drawn with N_SMALL sequences (synth.planted_bases with the spec's seed, motif lengths
and occupancies); D0 = synth.rand_probvec_filters(100 + seed, ...), X0 ~ N(0, 1) from numpy's
default_rng(200 + seed); lambda 0.05, rho 1, sigma 100.

Stored per dataset and precision: the objective trace, G1 and G3 on the filter support (both
vanish past the filter length), code_thresh with the positive count and the thresholded codes
(thresh_summary.summary; the active set of the first 8 sequences bit for bit).

Run:  python tests/golden/make_c3.py     (about 10 minutes on 8 host cores)
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

SPECS = 7
N_SMALL = 40
ITER_NUM, CSC_ITER, CDL_ITER = 10, 15, 1
LAM, RHO, SIGMA = 0.05, 1.0, 100.0
BAND = 1e-3


def specs():
    from cdlm import synth
    return synth.c3_specs(SPECS)


def inputs(spec, N=N_SMALL):
    """(s, D0, X0) of one dataset with N sequences (D0, X0 depend on the spec's seed only)."""
    from cdlm import synth
    bases, _ = synth.planted_bases(spec["seed"], N, spec["L"], spec["motif_lens"], spec["occupancy"], alpha=0.5)
    s = synth.onehot(bases)
    M, fl = spec["M"], spec["filter_len"]
    Lg = 4 * spec["L"] + 2 * fl
    D0 = synth.rand_probvec_filters(100 + spec["seed"], M, Lg, [fl] * M)
    X0 = np.random.default_rng(200 + spec["seed"]).standard_normal((N, M, Lg)).astype(np.float32)
    return s, D0, X0


def digest(a) -> str:
    return hashlib.sha1(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    import oracle as O
    from thresh_summary import summary
    out = {}
    for i, sp in enumerate(specs()):
        s, D0, X0 = inputs(sp)
        out["digest_%d" % i] = digest(s) + digest(D0) + digest(X0)
        fl = sp["filter_len"]
        for prec in ("f32", "f64"):
            t0 = time.time()
            o = O.OracleCDL(s, D0, X0, [fl] * sp["M"], RHO, SIGMA, LAM, prec)
            trace = []
            for _ in range(ITER_NUM):
                o.code_update(CSC_ITER)
                trace.append(o.update_best()[0])
                o.dict_update(CDL_ITER)
            k = "%s_%d" % (prec, i)
            out["trace_" + k] = np.asarray(trace, np.float64)
            out["G1_" + k] = np.asarray(o.get("G1"))[:, :fl]
            out["G3_" + k] = np.asarray(o.get("G3"))[:, :fl]
            th, cnt = o.code_thresh(1.0)
            summ = summary(o.get("Xbest"), o.gamma, 4 * sp["L"], th, band=BAND, n_xs=1, n_act=8)
            assert summ["count"] == cnt
            for kk, v in summ.items():
                out["%s_%s" % (kk, k)] = v
            out["thresh_" + k] = np.float64(th)
            print("dataset %d (seed %d, L=%d M=%d filter_len=%d Lg=%d) %s: %.1f s, thresh %.9g, active %d, band %d" % (
                i, sp["seed"], sp["L"], sp["M"], fl, o.Lg, prec, time.time() - t0, th, int(summ["actcnt"].sum()),
                len(summ["band_idx"])), flush=True)
            del o
    np.savez_compressed(os.path.join(HERE, "c3.npz"), **out)


if __name__ == "__main__":
    main()
