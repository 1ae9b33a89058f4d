"""Exact-rational restatement of the reference's p-value -> score threshold (Touzet-Varre
score distribution) and of its Fisher e-value test.

TEST INFRASTRUCTURE ONLY.  tests/ use it as the checker of the HIP library; the product
never imports it.  Pure Python loops over small cases.

Why exact rationals reproduce the reference's Double64 (DoubleFloats.jl 106-bit) arithmetic
bit for bit: after round_pwm every entry is k * c with k an integer and c = Float64(0.1)
(the product of a small integer and a double is exact in 106 bits), so every key of the
score distribution is an exact sum K * c; with the uniform background every probability is
an integer count times 4^-m, exact in 106 bits for m <= 53.  The one inexact step of the
reference, floor(pwm / c) on a Double64 quotient, is exact too: a nonzero Float32 entry x is
never an integer multiple of c, and its distance to the nearest one (>= 2^-53 |x| relative)
is far above the quotient's 2^-104 error.  So for m <= 53 the reference's result equals
the exact-rational one computed here.

Reference (CDLmotif.jl):
  best_score / worst_score      src/SEARCH/PWM_Touzet.jl:9-21
  min_score_range               PWM_Touzet.jl:32-33
  round_pwm                     PWM_Touzet.jl:44
  create_Q, score_distribution  PWM_Touzet.jl:68-102
  find_largest_alpha            PWM_Touzet.jl:147-159
  pvalue2score                  PWM_Touzet.jl:199-217  (ϵ = 0.1, bg uniform: TYPE_setup.jl:52-54)
  get_thresh                    src/SEARCH/helpers.jl:99
  filter_using_evalue!          src/SEARCH/e_value_filter.jl:1-28 (FisherExactTest from
                                HypothesisTests.jl, p-value tail=:both method=:central)
"""
from __future__ import annotations

import math
from fractions import Fraction
from typing import Dict, List

import numpy as np

GRANULARITY = 0.1                   # _granularity_ (TYPE_setup.jl:52), a Float64


def _col_best(pwm, i):
    return max(pwm[j][i] for j in range(4))


def _col_worst(pwm, i):
    return min(pwm[j][i] for j in range(4))


def round_pwm(pwm_f32: np.ndarray, eps: float = GRANULARITY) -> List[List[Fraction]]:
    """floor.(pwm ./ eps) * eps on exact values (PWM_Touzet.jl:44)."""
    c = Fraction(eps)
    return [[math.floor(Fraction(float(x)) / c) * c for x in row] for row in np.asarray(pwm_f32, np.float32)]


def min_score_range(pwm: List[List[Fraction]]) -> List[List[Fraction]]:
    """Columns reordered by decreasing best - worst (PWM_Touzet.jl:32-33; Julia's sortperm is
    stable, so is Python's sort)."""
    m = len(pwm[0])
    rng = [_col_best(pwm, i) - _col_worst(pwm, i) for i in range(m)]
    perm = sorted(range(m), key=lambda i: -rng[i])
    return [[row[i] for i in perm] for row in pwm]


def score_distribution(pwm: List[List[Fraction]], alpha, beta, bg) -> List[Dict[Fraction, Fraction]]:
    """score_distribution (PWM_Touzet.jl:79-102): Q[i][t] = P(score of columns 1..i = t),
    pruned to alpha - best(rest) <= t <= beta - worst(rest)."""
    m = len(pwm[0])
    Q = [dict() for _ in range(m + 1)]
    Q[0][Fraction(0)] = Fraction(1)
    for i in range(1, m + 1):
        bs = sum((_col_best(pwm, c) for c in range(i, m)), Fraction(0))
        ws = sum((_col_worst(pwm, c) for c in range(i, m)), Fraction(0))
        for score in sorted(Q[i - 1]):
            for j in range(4):
                t = score + pwm[j][i - 1]
                if alpha - bs <= t and (beta is None or t <= beta - ws):
                    Q[i][t] = Q[i].get(t, Fraction(0)) + Q[i - 1][score] * bg[j]
    return Q


def find_largest_alpha(Qm: Dict[Fraction, Fraction], pval: Fraction):
    """find_largest_alpha (PWM_Touzet.jl:147-159), keys in ascending order."""
    q_sum = sum(Qm.values(), Fraction(0))
    largest = None
    for k in sorted(Qm):
        if q_sum >= pval:
            largest = k
        else:
            return k
        q_sum -= Qm[k]
    return largest


def pvalue2score(pwm_f32: np.ndarray, pval: float, eps: float = GRANULARITY) -> Fraction:
    """pvalue2score(pwm, pval) (PWM_Touzet.jl:199-217) with the uniform background, exact."""
    pwm_f32 = np.asarray(pwm_f32, np.float32)
    assert 0 <= pval <= 1 and pwm_f32.shape[0] == 4
    bg = [Fraction(1, 4)] * 4
    mp = min_score_range([[Fraction(float(x)) for x in row] for row in pwm_f32])
    c = Fraction(eps)
    pe = [[math.floor(x / c) * c for x in row] for row in mp]
    ws = sum((_col_worst(pe, i) for i in range(len(pe[0]))), Fraction(0))
    Q = score_distribution(pe, ws, None, bg)
    return find_largest_alpha(Q[len(pe[0])], Fraction(float(pval)))


def get_thresh(pwms, pval: float) -> np.ndarray:
    """get_thresh(pwms, pval, Float32) (helpers.jl:99): Float32 of each threshold; the
    Double64 -> Float32 conversion rounds the Float64 head, i.e. Float32(Float64(K * c))."""
    return np.array([np.float32(float(pvalue2score(p, pval))) for p in pwms], np.float32)


def brute_force_threshold(pwm_f32: np.ndarray, pval: float, eps: float = GRANULARITY) -> Fraction:
    """Independent restatement by enumeration of all 4^m sequences (m <= 8): the smallest
    attainable rounded score whose upper tail probability is below pval, else the largest
    attainable score."""
    pe = round_pwm(pwm_f32, eps)
    m = len(pe[0])
    scores = {}
    for idx in range(4 ** m):
        s, v = Fraction(0), idx
        for i in range(m):
            s += pe[v & 3][i]
            v >>= 2
        scores[s] = scores.get(s, 0) + 1
    keys = sorted(scores)
    tail, p = 4 ** m, Fraction(float(pval))
    for k in keys:
        if Fraction(tail, 4 ** m) < p:
            return k
        tail -= scores[k]
    return keys[-1]


def fisher_central_pvalue(a: int, b: int, c: int, d: int) -> Fraction:
    """Two-sided p-value of Fisher's exact test on [[a, b], [c, d]], method :central of
    HypothesisTests.jl: min(1, 2 min(P(X <= a), P(X >= a))) for X hypergeometric with the
    table's margins, as an exact rational."""
    n1, n2, k = a + b, c + d, a + c
    lo, hi = max(0, k - n2), min(k, n1)
    tot = math.comb(n1 + n2, k)
    pmf = {x: Fraction(math.comb(n1, x) * math.comb(n2, k - x), tot) for x in range(lo, hi + 1)}
    cdf = sum((pmf[x] for x in range(lo, a + 1)), Fraction(0))
    ccdf = sum((pmf[x] for x in range(a, hi + 1)), Fraction(0))
    return min(Fraction(1), 2 * min(cdf, ccdf))


def evalues(n_fg: np.ndarray, n_bg: np.ndarray, N: int) -> np.ndarray:
    """filter_using_evalue! (e_value_filter.jl:17-23): FisherExactTest(a, N, b, N) per motif,
    e-value = p * num_motifs, in Float64."""
    K = len(n_fg)
    return np.array([float(fisher_central_pvalue(int(a), N, int(b), N) * K) for a, b in zip(n_fg, n_bg)])
