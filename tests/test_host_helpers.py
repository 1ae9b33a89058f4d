"""Host helpers of the scan path pinned by hand-worked cases (pfm2pwm, pwm_complement,
best_score: src/SEARCH/helpers.jl:12-16, src/SEARCH/PWM_Touzet.jl:9), the dataset form of
init_cdl (src/SETUP/CDL_setup.jl:273-283 reads dna_dataset's data matrix) and the refusal of a
Sherman-Morrison constant other than the problem's (train.jl:1-4,46-49)."""
import numpy as np
import pytest

import oracle as O
from cdlm import scan as S
from cdlm import cdl as CD
from cdlm import io as IO
from cdlm._lib import ERR_ARG, CdlmError


def test_pfm2pwm_hand_cases():
    # log2(p / 0.25): 0.5 -> 1, 0.25 -> 0, 0.125 -> -1, 1 -> 2 (all exact in float32)
    pfm = np.array([[0.5, 0.25, 1.0], [0.25, 0.25, 0.0], [0.125, 0.25, 0.0], [0.125, 0.25, 0.0]])
    w = S.pfm2pwm(pfm)
    assert w.dtype == np.float32
    np.testing.assert_array_equal(w[:, 0], [1, 0, -1, -1])
    np.testing.assert_array_equal(w[:, 1], [0, 0, 0, 0])
    assert w[0, 2] == 2 and np.all(np.isneginf(w[1:, 2]))
    # a per-row background (bg vector, helpers.jl:12 broadcasts pfm ./ bg over columns)
    w2 = S.pfm2pwm(np.full((4, 1), 0.25), bg=[0.5, 0.25, 0.125, 0.125])
    np.testing.assert_array_equal(w2[:, 0], [-1, 0, 1, 1])


def test_pwm_complement_hand_case():
    # reverse(pwm) reverses both dimensions: rows A,C,G,T -> T,G,C,A and columns back to front
    pwm = np.arange(12, dtype=np.float32).reshape(4, 3)      # pwm[a, j] = 3a + j
    c = S.pwm_complement(pwm)
    expect = np.array([[11, 10, 9], [8, 7, 6], [5, 4, 3], [2, 1, 0]], np.float32)
    np.testing.assert_array_equal(c, expect)
    # the complement of a window's bases scores the reverse-complement strand: the forward
    # window ACG under pwm scores pwm[0,0]+pwm[1,1]+pwm[2,2]; its reverse complement CGT under c
    # scores c[1,0]+c[2,1]+c[3,2] = pwm[2,2]+pwm[1,1]+pwm[0,0]
    assert c[1, 0] + c[2, 1] + c[3, 2] == pwm[0, 0] + pwm[1, 1] + pwm[2, 2]
    np.testing.assert_array_equal(S.pwm_complement(c), pwm)


def test_best_score_hand_cases():
    pwm = np.array([[1.0, -1.0], [0.5, 2.0], [-3.0, 0.0], [0.25, 1.5]], np.float32)
    assert S.best_score(pwm) == np.float32(3.0)
    # sequential FP32: column maxima 1e8, 1, -1e8 sum to 0 in float32 ((1e8 + 1) rounds to 1e8),
    # to 1 in float64 -- the reference sums Float32 values one by one
    col = np.array([1e8, 1.0, -1e8], np.float32)
    pwm2 = np.full((4, 3), -1e9, np.float32)
    pwm2[2] = col
    assert S.best_score(pwm2) == np.float32(0.0)
    assert float(np.sum(col.astype(np.float64))) == 1.0
    assert O.best_score(pwm2) == 0.0 and O.best_score(pwm) == 3.0


def test_dataset_form_of_init_cdl_reads_data_matrix(tmp_path):
    fa = tmp_path / "r.fa"
    fa.write_text(">a\nACGT\nAC\n>b\nTTGCAA\n>c\nACNTTT\n")
    ds = CD.DNADataset.from_fasta(str(fa))
    assert (ds.N, ds.L) == (2, 6) and ds.raw_data == ["ACGTAC", "TTGCAA"]
    np.testing.assert_array_equal(ds.data_matrix, IO.data_2_dummy(["ACGTAC", "TTGCAA"]))
    # data_matrix is read from any object that has one, as the reference's init_cdl does
    seen = {}

    class FakeCDL:
        def __init__(self, data, fl, rho, sigma, lam, **kw):
            seen.update(data=data, fl=list(fl), rho=rho, sigma=sigma, lam=lam)

    orig = CD.CDL
    CD.CDL = FakeCDL
    try:
        CD.init_cdl(CD.CDL_setup(M=3, filter_size=5, rho=1.0, sigma=50.0, lam=0.05), ds)
    finally:
        CD.CDL = orig
    assert seen["data"] is ds.data_matrix and seen["fl"] == [20, 20, 20]
    assert (seen["rho"], seen["sigma"], seen["lam"]) == (1.0, 50.0, 0.05)


def test_other_sherman_constant_is_refused():
    c = CD.CDL.__new__(CD.CDL)              # no device needed: the check precedes the library call
    c.sherman_csc, c.sherman_cdl = 2.0, 2.0
    for fn, kw in ((c.code_update, "sherman_const_csc"), (c.dict_update, "sherman_const_cdl")):
        with pytest.raises(CdlmError) as e:
            fn(1, **{kw: 1.5})
        assert e.value.status == ERR_ARG and not isinstance(e.value, ValueError)
    for fn, kw in ((CD.code_update, "sherman_const_csc"), (CD.dict_update, "sherman_const_cdl")):
        with pytest.raises(CdlmError):
            fn(c, 1, np.complex64, **{kw: 3.0})
