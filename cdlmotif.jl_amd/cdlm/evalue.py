"""E-value filter of the discovered motifs (filter_using_evalue!, src/SEARCH/e_value_filter.jl:1-46).

The counts come from the device scans (the number of sequences with a hit, foreground and
shuffled background); the test itself is host arithmetic: Fisher's exact test on the table
[[a, N], [b, N]] with the two-sided p-value of HypothesisTests.jl's method :central,
min(1, 2 min(P(X <= a), P(X >= a))) for X hypergeometric with the table's margins, times the
number of motifs.  Motifs pass with e-value < threshold and length > 5 (e_value_filter.jl:31-36).
"""
from __future__ import annotations

import math
from typing import Sequence

import numpy as np


def _log_pmf(n1: int, n2: int, k: int, x: int, lnorm: float) -> float:
    return (math.lgamma(n1 + 1) - math.lgamma(x + 1) - math.lgamma(n1 - x + 1) + math.lgamma(n2 + 1)
            - math.lgamma(k - x + 1) - math.lgamma(n2 - k + x + 1) - lnorm)


def _log_sum(logs):
    m = max(logs)
    if m == -math.inf:
        return m
    return m + math.log(sum(math.exp(v - m) for v in logs))


def fisher_pvalue(a: int, b: int, c: int, d: int) -> float:
    """pvalue(FisherExactTest(a, b, c, d)) with tail=:both, method=:central."""
    if min(a, b, c, d) < 0:
        raise ValueError("table entries must be non-negative")
    n1, n2, k = a + b, c + d, a + c
    lo, hi = max(0, k - n2), min(k, n1)
    lnorm = math.lgamma(n1 + n2 + 1) - math.lgamma(k + 1) - math.lgamma(n1 + n2 - k + 1)
    left = _log_sum([_log_pmf(n1, n2, k, x, lnorm) for x in range(lo, a + 1)])
    right = _log_sum([_log_pmf(n1, n2, k, x, lnorm) for x in range(a, hi + 1)])
    return min(1.0, 2.0 * math.exp(min(left, right)))


def evalues(n_fg: Sequence[int], n_bg: Sequence[int], N: int) -> np.ndarray:
    """evalues[i] = pvalue(FisherExactTest(n_fg[i], N, n_bg[i], N)) * num_motifs
    (e_value_filter.jl:17-23)."""
    K = len(n_fg)
    return np.array([fisher_pvalue(int(a), N, int(b), N) * K for a, b in zip(n_fg, n_bg)], np.float64)


def filter_using_evalue(n_fg, n_bg, N: int, lens, e_value_thresh: float):
    """(evalues, pass mask): evalue < e_value_thresh and length > 5 (evalue_filtering!,
    e_value_filter.jl:30-36)."""
    e = evalues(n_fg, n_bg, N)
    return e, (e < e_value_thresh) & (np.asarray(lens) > 5)


def sequences_with_hits(positions: np.ndarray) -> np.ndarray:
    """length.(positions) of the reference's per-motif {sequence => position} dictionaries
    from a (K, N) position array of the device scans (0 = no hit)."""
    return (np.asarray(positions) > 0).sum(axis=1).astype(np.int64)
