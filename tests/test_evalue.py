"""Host e-value arithmetic (filter_using_evalue!, src/SEARCH/e_value_filter.jl:17-36) against
the oracle's exact rationals (Fisher's exact test, central two-sided p-value); floating
point, compared to 1e-9 relative."""
import numpy as np
import pytest

import touzet as T
from cdlm import evalue as E


@pytest.mark.parametrize("table", [(5, 10, 5, 10), (1, 3, 3, 1), (0, 20, 7, 20), (20, 20, 0, 20),
                                   (130, 400, 60, 400), (900, 1000, 850, 1000), (3, 2000, 40, 2000)])
def test_fisher_central_matches_exact(table):
    got = E.fisher_pvalue(*table)
    ref = float(T.fisher_central_pvalue(*table))
    assert got == pytest.approx(ref, rel=1e-9, abs=1e-300)


# Known answers worked by hand from the hypergeometric law (X = top-left cell with the table's
# margins; central two-sided p = min(1, 2 min(P(X <= a), P(X >= a)))):
#  * (3, 1, 1, 3), Fisher's tea-tasting table: P(X >= 3) = (16 + 1) / C(8, 4) = 17/70 -> 34/70;
#  * (1, 9, 11, 3), the dieting table of the Fisher's exact test article: n1 = 10, n2 = 14, k = 12,
#    P(X = 0) = C(14,12)/C(24,12) = 91/2704156, P(X = 1) = 10 C(14,11)/C(24,12) = 3640/2704156,
#    so 2 (91 + 3640) / 2704156 = 7462/2704156 (= 0.0027594..., the article's 0.0028);
#  * (0, 5, 5, 0): P(X <= 0) = 1/C(10,5) = 1/252 -> 2/252;
#  * (1, 1, 1, 1): P(X <= 1) = P(X >= 1) = 5/6 -> min(1, 10/6) = 1;
#  * (2, 0, 0, 2): P(X >= 2) = 1/C(4,2) = 1/6 -> 1/3.
FISHER_KAT = [((3, 1, 1, 3), 34 / 70), ((1, 9, 11, 3), 7462 / 2704156), ((0, 5, 5, 0), 2 / 252),
              ((1, 1, 1, 1), 1.0), ((2, 0, 0, 2), 1 / 3)]


@pytest.mark.parametrize("table,p", FISHER_KAT)
def test_fisher_known_answers(table, p):
    assert E.fisher_pvalue(*table) == pytest.approx(p, rel=1e-12)
    assert float(T.fisher_central_pvalue(*table)) == pytest.approx(p, rel=1e-15)


def test_evalues_and_filter():
    n_fg, n_bg, N = np.array([120, 40, 300, 55]), np.array([30, 38, 90, 20]), 400
    e, ok = E.filter_using_evalue(n_fg, n_bg, N, [8, 12, 5, 10], 1e-3)
    ref = T.evalues(n_fg, n_bg, N)
    np.testing.assert_allclose(e, ref, rtol=1e-9)
    assert ok.tolist() == [bool(ref[0] < 1e-3), bool(ref[1] < 1e-3), False, bool(ref[3] < 1e-3)]
    assert ok[0] and not ok[1]
    np.testing.assert_array_equal(E.sequences_with_hits(np.array([[0, 3, 1], [0, 0, 0]])), [2, 0])
