"""The oracle's exact-rational restatement of pvalue2score (src/SEARCH/PWM_Touzet.jl:199-217)
and of the Fisher e-value test (src/SEARCH/e_value_filter.jl:17-23), pinned by:
  * a hand-worked one- and two-column case;
  * brute-force enumeration of all 4^m sequences for m <= 7 (an independent restatement);
  * invariance under column permutation (min_score_range only reorders columns);
  * the Float32 conversion of K * Float64(0.1): no K in [-2e5, 2e5] where rounding the
    Float64 head differs from rounding the exact product, so either reading of
    Float32(::Double64) gives the same thresholds;
  * Fisher: symmetric tables give p = 1; a table computed by hand."""
from fractions import Fraction

import numpy as np
import pytest

import touzet as T

C = Fraction(0.1)


def _rand_pwms(rng, n, m_lo, m_hi):
    out = []
    for _ in range(n):
        m = int(rng.integers(m_lo, m_hi + 1))
        p = rng.dirichlet(np.full(4, 0.5), m).T
        p = (p + 0.01) / 1.04
        out.append(np.log2(p / 0.25).astype(np.float32))
    return out


def test_hand_case():
    # one column [0.35, -0.05, 0.2, -1.0]: rounded to 3c, -1c, 2c, -10c (floor); keys
    # -10c, -1c, 2c, 3c each with probability 1/4.  Tails: 1, 3/4, 1/2, 1/4.
    pwm = np.array([[0.35], [-0.05], [0.2], [-1.0]], np.float32)
    assert T.round_pwm(pwm) == [[3 * C], [-1 * C], [2 * C], [-10 * C]]
    assert T.pvalue2score(pwm, 0.3) == 3 * C           # first key with tail < 0.3 is 3c (tail 1/4)
    assert T.pvalue2score(pwm, 0.25) == 3 * C          # tail(3c) = 1/4 >= 0.25: no key qualifies -> last key
    assert T.pvalue2score(pwm, 0.6) == 2 * C
    assert T.pvalue2score(pwm, 1.0) == -1 * C          # tail(-10c) = 1 >= 1, tail(-1c) = 3/4 < 1
    assert T.pvalue2score(pwm, 0.0) == 3 * C
    # two columns: second column all 0.1 -> every key shifted by c (Float32(0.1) >= Float64(0.1))
    pwm2 = np.array([[0.35, 0.1], [-0.05, 0.1], [0.2, 0.1], [-1.0, 0.1]], np.float32)
    assert T.pvalue2score(pwm2, 0.3) == 4 * C
    # Float32(0.1) < Float64(0.1)? then floor gives 0; check which way it goes exactly
    assert Fraction(float(np.float32(0.1))) > C


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_against_brute_force(seed):
    rng = np.random.default_rng(seed)
    for pwm in _rand_pwms(rng, 12, 1, 7):
        for pv in (1e-4, 2.7e-4, 0.01, 0.1, 0.5):
            pv32 = float(np.float32(pv))
            assert T.pvalue2score(pwm, pv32) == T.brute_force_threshold(pwm, pv32), (pwm.shape, pv)


def test_column_permutation_invariance():
    rng = np.random.default_rng(9)
    for pwm in _rand_pwms(rng, 6, 8, 14):
        perm = rng.permutation(pwm.shape[1])
        assert T.pvalue2score(pwm, 2.7e-4) == T.pvalue2score(pwm[:, perm], 2.7e-4)


def test_float32_of_k_times_c_is_unambiguous():
    c = 0.1
    ks = np.arange(-200000, 200001, dtype=np.int64)
    head = (ks.astype(np.float64) * c)                      # Float64 head of the Double64 K*c
    via_head = head.astype(np.float32)
    # correctly rounded Float32 of the exact product, by exact rational comparison only where
    # the head sits within 2^-40 relative of a Float32 rounding boundary
    f32_lo = np.nextafter(via_head, np.float32(-np.inf))
    f32_hi = np.nextafter(via_head, np.float32(np.inf))
    near = np.zeros(len(ks), bool)
    for nb in (f32_lo, f32_hi):
        mid = (via_head.astype(np.float64) + nb.astype(np.float64)) / 2
        near |= np.abs(head - mid) <= np.abs(head) * 2.0 ** -40
    for i in np.nonzero(near)[0]:
        exact = int(ks[i]) * C
        cand = [via_head[i], f32_lo[i], f32_hi[i]]
        best = min(cand, key=lambda v: (abs(Fraction(float(v)) - exact), int(np.float32(v).view(np.int32)) & 1))
        assert best == via_head[i], int(ks[i])


def test_fisher_central():
    assert T.fisher_central_pvalue(5, 10, 5, 10) == 1
    # [[1, 3], [3, 1]]: X ~ Hypergeometric(4, 4, 4), P(X=x) = C(4,x)C(4,4-x)/70:
    # 1/70, 16/70, 36/70, 16/70, 1/70; P(X<=1) = 17/70, P(X>=1) = 69/70 -> p = 34/70
    assert T.fisher_central_pvalue(1, 3, 3, 1) == Fraction(34, 70)
    e = T.evalues(np.array([1, 5]), np.array([3, 5]), 4)
    np.testing.assert_allclose(e, [float(T.fisher_central_pvalue(1, 4, 3, 4)) * 2, 2.0])
