"""Seeded synthetic planted-motif datasets (the reference's simulator semantics, written
with numpy's generator; the Julia RNG streams themselves are not reproduced).

Planting follows sample_backgound_with_motif_single (src/LOAD/sample.jl:139-172): for
each motif and sequence, a Bernoulli(occupancy) draw plants one realisation of the motif
(each column drawn from its categorical) at a uniform start, reverse-complemented with
probability rc_prob (TYPE_setup.jl:16); later motifs overwrite earlier ones.  Motif
columns are Dirichlet(alpha) draws (SIM_setup.jl:14-35 uses Dirichlet columns).
"""
from __future__ import annotations

from typing import List, Sequence, Tuple

import numpy as np


def planted_bases(seed: int, N: int, L: int, motif_lens: Sequence[int], occupancy, alpha: float = 0.5,
                  bg=(0.25, 0.25, 0.25, 0.25), rc_prob: float = 0.5) -> Tuple[np.ndarray, List[np.ndarray]]:
    """(N, L) uint8 base codes and the list of motif PFMs (4, m)."""
    rng = np.random.default_rng(seed)
    bases = rng.choice(4, size=(N, L), p=np.asarray(bg, float)).astype(np.uint8)
    occ = np.broadcast_to(np.asarray(occupancy, float), (len(motif_lens),))
    pfms = []
    for i, m in enumerate(motif_lens):
        m = int(min(m, L))
        pfm = rng.dirichlet(alpha * np.ones(4), m)            # (m, 4)
        pfms.append(pfm.T.copy())
        hit = rng.random(N) < occ[i]
        idx = np.nonzero(hit)[0]
        if idx.size == 0:
            continue
        starts = rng.integers(0, L - m + 1, idx.size)
        cdf = np.cumsum(pfm, axis=1)
        u = rng.random((idx.size, m))
        real = (u[:, :, None] > cdf[None, :, :]).sum(axis=2).clip(0, 3).astype(np.uint8)   # (n, m)
        rc = rng.random(idx.size) < rc_prob
        real[rc] = 3 - real[rc, ::-1]
        cols = starts[:, None] + np.arange(m)[None, :]
        bases[idx[:, None], cols] = real
    return bases, pfms


def onehot(bases: np.ndarray) -> np.ndarray:
    """(N, L) codes -> (N, 4L) float32 one-hot rows = the reference's (4L, N) data matrix
    (src/LOAD/to_dummy.jl:2-38)."""
    N, L = bases.shape
    out = np.zeros((N, 4 * L), np.float32)
    ok = bases < 4
    n_idx, l_idx = np.nonzero(ok)
    out[n_idx, 4 * l_idx + bases[ok]] = 1.0
    return out


def rand_probvec_filters(seed: int, M: int, Lg: int, filter_len: Sequence[int]) -> np.ndarray:
    """(M, Lg) initial filters: each aligned 4-row group a uniform point on the simplex
    (rand_probvec, CDL_setup.jl:59-68), masked to the filter support (:183)."""
    rng = np.random.default_rng(seed)
    u = np.sort(rng.random((M, Lg // 4, 3)), axis=2)
    z = np.zeros((M, Lg // 4, 1))
    o = np.ones((M, Lg // 4, 1))
    D = np.diff(np.concatenate([z, u, o], axis=2), axis=2).reshape(M, Lg).astype(np.float32)
    for m in range(M):
        D[m, filter_len[m]:] = 0
    return D


def c2_dataset(seed: int = 2, N: int = 20000, L: int = 200):
    """The single-dataset workload: N sequences of L bases, 5 planted motifs of length
    8..15 (Dirichlet(0.5) columns) at 20% occupancy each, random strand."""
    rng = np.random.default_rng(seed + 1000)
    lens = rng.integers(8, 16, 5)
    return planted_bases(seed, N, L, lens, 0.2, alpha=0.5)


def c1_dataset(seed: int = 1, N: int = 500, L: int = 100):
    """CPU reference workload: 2 motifs of length 10 and 12 in 30% of sequences."""
    return planted_bases(seed, N, L, [10, 12], 0.3, alpha=0.5)


def pwm_workload(seed: int, N: int, L: int, K: int, len_range=(8, 24), gc: float = 0.42):
    """PWM-scan workload: N x L iid bases with the given GC fraction; K log-odds PWMs
    (columns Dirichlet(0.5), pseudocount (p+0.01)/1.04, background 0.25, computed in
    float64 and stored as float32); thresholds 0.8 x best_score (sequential FP32)."""
    rng = np.random.default_rng(seed)
    at = (1.0 - gc) / 2
    # bases by inverse CDF of float32 uniforms, 2^20 sequences at a time (10^6 x 500 in seconds)
    cdf = np.cumsum([at, gc / 2, gc / 2]).astype(np.float32)
    bases = np.empty((N, L), np.uint8)
    for a in range(0, N, 1 << 20):
        b = min(N, a + (1 << 20))
        bases[a:b] = np.searchsorted(cdf, rng.random((b - a, L), dtype=np.float32), side="right")
    lens = rng.integers(len_range[0], len_range[1] + 1, K).astype(np.int32)
    maxlen = int(lens.max())
    pwms = np.zeros((K, 4, maxlen), np.float32)
    thresh = np.zeros(K, np.float32)
    for k in range(K):
        p = (rng.dirichlet(0.5 * np.ones(4), lens[k]) + 0.01) / 1.04
        w = np.log2(p / 0.25).T.astype(np.float32)
        pwms[k, :, :lens[k]] = w
        acc = np.float32(0)
        for j in range(lens[k]):
            acc = np.float32(acc + w[:, j].max())
        thresh[k] = np.float32(np.float32(0.8) * acc)
    return bases, pwms, lens, thresh


def c3_specs(count: int, seed0: int = 3, N_range=(200, 20000)):
    """Dataset descriptions of the batched workload: dataset i uses seed seed0+i; N uniform
    in N_range, L from {100,150,200,300,500}, 1-6 planted motifs of length 8-15 at
    occupancy U[0.1,0.4], M from {32,64}, flattened filter length from {16,20,24}."""
    out = []
    for i in range(count):
        rng = np.random.default_rng(seed0 + i)
        nm = int(rng.integers(1, 7))
        out.append({
            "seed": seed0 + i,
            "N": int(rng.integers(N_range[0], N_range[1] + 1)),
            "L": int(rng.choice([100, 150, 200, 300, 500])),
            "M": int(rng.choice([32, 64])),
            "filter_len": int(rng.choice([16, 20, 24])),
            "motif_lens": [int(x) for x in rng.integers(8, 16, nm)],
            "occupancy": [float(x) for x in rng.uniform(0.1, 0.4, nm)],
        })
    return out


def c3_onehot(spec) -> np.ndarray:
    """(N, 4L) one-hot data of one batched-workload dataset."""
    bases, _ = planted_bases(spec["seed"], spec["N"], spec["L"], spec["motif_lens"], spec["occupancy"], alpha=0.5)
    return onehot(bases)
