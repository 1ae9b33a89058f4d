"""The reference-structured GPU loop used as the plain per-dataset baseline of the bench
(refloop.py at the repository root: the reference's code_update!/dict_update!/update_best! statement for statement
on torch/hipFFT) computes the same iteration as the oracle: relative error <= 1e-4 after a
short schedule, so the baseline it times is the reference's algorithm."""
import os

import numpy as np
import pytest

import oracle as O
from gpu_util import flush, rel

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_reference_structured_loop_matches_oracle():
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    from make_golden import make_cdl_inputs
    from refloop import ReferenceStructuredCDL
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    fl = [12, 8, 12, 16, 4]
    s, D0, X0, _ = make_cdl_inputs(31, 14, 23, len(fl), fl)
    r = ReferenceStructuredCDL(s, fl, 1.0, 100.0, 0.05, D0, X0)
    o = O.OracleCDL(s, D0, X0, fl, 1.0, 100.0, 0.05, "f32")
    assert abs(r.obj_best - o.objective()) <= 1e-4 * abs(o.objective())
    for _ in range(2):
        r.code_update(5); o.code_update(5)
        ob_r, imp_r = r.update_best(); ob_o, imp_o = o.update_best()
        flush("objective %.9g / %.9g" % (ob_r, ob_o))
        assert abs(ob_r - ob_o) <= 1e-4 * abs(ob_o) and imp_r == imp_o
        r.dict_update(1); o.dict_update(1)
    for f, g in (("X", "X"), ("DX", "DX"), ("d", "d"), ("G1", "G1"), ("G3", "G3"), ("Td", "Td")):
        e = rel(r.get(f), o.get(g))
        flush("%s: %.2e" % (f, e))
        assert e <= 1e-4, (f, e)
