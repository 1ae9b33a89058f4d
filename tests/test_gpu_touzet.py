"""Score thresholds for a p-value on the device (get_thresh / pvalue2score,
src/SEARCH/helpers.jl:99, src/SEARCH/PWM_Touzet.jl:199-217) against the oracle's exact
restatement: Float32 thresholds, compared exactly."""
import numpy as np
import pytest

import touzet as T
from gpu_util import flush

pytestmark = pytest.mark.gpu
cdlm = pytest.importorskip("cdlm")

PV32 = float(np.float32(2.7e-4))      # the reference's dat_t(2.7e-4) (wrap.jl:125)


def _pwms(rng, n, m_lo, m_hi):
    out = []
    for _ in range(n):
        m = int(rng.integers(m_lo, m_hi + 1))
        p = rng.dirichlet(np.full(4, 0.5), m).T
        out.append(cdlm.pfm2pwm((p + 0.01) / 1.04))
    return out


@pytest.mark.parametrize("pval", [0.0, PV32, 1e-3, 0.05, 0.5, 1.0])
def test_thresholds_match_oracle(pval):
    rng = np.random.default_rng(int(pval * 1e6) + 7)
    pwms = _pwms(rng, 40, 1, 14)
    got = cdlm.get_thresh(pwms, pval)
    ref = T.get_thresh(pwms, pval)
    flush("pval %g: %d PWMs, %.3f ms" % (pval, len(pwms), cdlm.pvalue_last_ms()))
    np.testing.assert_array_equal(got, ref)


def test_long_pwms_and_wide_counts():
    rng = np.random.default_rng(5)
    pwms = _pwms(rng, 3, 22, 26)
    # 53 columns of small integer multiples: counts up to 4^53 > 2^64
    q = rng.integers(-3, 4, (4, 53))
    pwms.append((q * np.float32(0.1)).astype(np.float32))
    pwms.append(np.zeros((4, 53), np.float32))          # one key: every p-value returns it
    got = cdlm.get_thresh(pwms, PV32)
    ref = T.get_thresh(pwms, PV32)
    flush("long PWMs %s: %s" % ([p.shape[1] for p in pwms], got))
    np.testing.assert_array_equal(got, ref)


def test_reference_pwms_of_the_scan_workload():
    """PWMs shaped like the scan workload (lengths 8-24 from Dirichlet(0.5) columns with the
    0.01 pseudocount): a sample of 24 against the oracle, then all 512 on the device."""
    rng = np.random.default_rng(11)
    pwms = _pwms(rng, 512, 8, 24)
    got = cdlm.get_thresh(pwms, PV32)
    flush("512 PWMs: %.3f ms on the device" % cdlm.pvalue_last_ms())
    idx = rng.choice(512, 24, replace=False)
    np.testing.assert_array_equal(got[idx], T.get_thresh([pwms[i] for i in idx], PV32))
    best = np.array([cdlm.best_score(p) for p in pwms])
    assert (got <= best + 1e-4).all()


def test_argument_errors():
    p = [np.zeros((4, 5), np.float32)]
    with pytest.raises(cdlm.CdlmError):
        cdlm.get_thresh(p, 1.5)                          # the reference asserts 0 <= pval <= 1
    with pytest.raises(cdlm.CdlmError):
        cdlm.get_thresh([np.zeros((4, 54), np.float32)], PV32)
    bad = np.zeros((4, 5), np.float32)
    bad[2, 3] = np.nan
    with pytest.raises(cdlm.CdlmError):
        cdlm.get_thresh([bad], PV32)
    assert cdlm.get_thresh(p, PV32).tolist() == [0.0]
