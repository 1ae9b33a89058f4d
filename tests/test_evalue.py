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


def test_evalues_and_filter():
    n_fg, n_bg, N = np.array([120, 40, 300, 55]), np.array([30, 38, 90, 20]), 400
    e, ok = E.filter_using_evalue(n_fg, n_bg, N, [8, 12, 5, 10], 1e-3)
    ref = T.evalues(n_fg, n_bg, N)
    np.testing.assert_allclose(e, ref, rtol=1e-9)
    assert ok.tolist() == [bool(ref[0] < 1e-3), bool(ref[1] < 1e-3), False, bool(ref[3] < 1e-3)]
    assert ok[0] and not ok[1]
    np.testing.assert_array_equal(E.sequences_with_hits(np.array([[0, 3, 1], [0, 0, 0]])), [2, 0])
