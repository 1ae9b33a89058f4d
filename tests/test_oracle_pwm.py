"""PWM-scan oracle against the numpy restatement's fixture and hand-computed answers. CPU only."""
import os

import numpy as np

import oracle as O


def _load(golden_dir):
    return np.load(os.path.join(golden_dir, "pwm_small.npz"))


def test_hits_match_fixture(golden_dir):
    z = _load(golden_dir)
    k, n, p, st, sc = O.pwm_hits(z["bases"], z["pwms"], z["lens"], z["thresh"], both_strands=True)
    got = np.stack([k, n, p, st], 1).astype(np.int64)
    exp = z["hits"]
    # both are emitted in (k, n, pos, strand) order
    np.testing.assert_array_equal(got, exp)
    np.testing.assert_array_equal(sc.view(np.uint32), z["hit_scores"].view(np.uint32))


def test_gpu_argmax_matches_fixture(golden_dir):
    z = _load(golden_dir)
    pos = O.pwm_gpu_argmax(z["bases"], z["pwms"], z["lens"], z["thresh"])
    np.testing.assert_array_equal(pos, z["argmax_pos"])


def test_best_matches_fixture(golden_dir):
    z = _load(golden_dir)
    pos, comp, sc = O.pwm_best(z["bases"], z["pwms"], z["lens"], z["thresh"])
    np.testing.assert_array_equal(pos, z["best_pos"])
    np.testing.assert_array_equal(comp, z["best_comp"])
    np.testing.assert_array_equal(sc.view(np.uint32), z["best_score"].view(np.uint32))


def _one(pwm_4m, bases, thresh):
    m = pwm_4m.shape[1]
    pw = np.asarray(pwm_4m, np.float32)[None]
    return np.asarray([bases], np.uint8), pw, np.array([m], np.int32), np.array([thresh], np.float32)


def test_known_answers_by_hand():
    # PWM of length 2: column 0 favours A (row 0), column 1 favours T (row 3)
    pwm = np.array([[1.0, -1.0], [-1.0, -1.0], [-1.0, -1.0], [-1.0, 1.5]], np.float32)
    # sequence A T C A T  -> forward windows: AT=2.5, TC=-2, CA=-2, AT=2.5
    # reverse complement PWM = reverse in both dims: rc[a][j] = pwm[3-a][1-j]
    # rc: col0 = pwm[:,1] reversed rows = [1.5,-1,-1,-1]; col1 = pwm[:,0] reversed = [-1,-1,-1,1]
    # rc windows: AT = 1.5+1 = 2.5, TC=-2, CA=-1+(-1)... C then A: rc[1][0]=-1, rc[0][1]=-1 -> -2, AT=2.5
    seq, pw, ln, th = _one(pwm, [0, 3, 1, 0, 3], 2.0)
    k, n, p, st, sc = O.pwm_hits(seq, pw, ln, th)
    assert list(zip(p.tolist(), st.tolist())) == [(0, 0), (0, 1), (3, 0), (3, 1)]
    assert np.all(sc == np.float32(2.5))
    # GPU argmax semantic: forward max at l=1 (1-based) and l=4; first max is position 1 -> no hit (Q6)
    assert O.pwm_gpu_argmax(seq, pw, ln, th)[0, 0] == 0
    # best-hit semantic: forward 2.5 at 1, rc 2.5 at 1, tie -> forward wins, position 1 reported
    pos, comp, s = O.pwm_best(seq, pw, ln, th)
    assert (pos[0, 0], comp[0, 0], s[0, 0]) == (1, 0, np.float32(2.5))
    # a sequence where the best forward window is the second one
    seq2, *_ = _one(pwm, [1, 0, 3, 1, 1], 2.0)
    assert O.pwm_gpu_argmax(seq2, pw, ln, th)[0, 0] == 2
    # below threshold everywhere -> no hit in every mode
    seq3, *_ = _one(pwm, [1, 1, 1, 1, 1], 2.0)
    assert O.pwm_gpu_argmax(seq3, pw, ln, th)[0, 0] == 0
    assert O.pwm_best(seq3, pw, ln, th)[0][0, 0] == 0
    assert len(O.pwm_hits(seq3, pw, ln, th)[0]) == 0


def test_negative_threshold_argmax_quirk():
    """scan.jl:189-197: windows past L-m+1 stay 0; with all in-range scores negative but
    above a negative threshold, the argmax lands on the first zero past the range."""
    pwm = np.full((4, 3), -1.0, np.float32)
    seq, pw, ln, th = _one(pwm, [0, 1, 2, 3, 0, 1], -5.0)
    # L=6, m=3: valid windows 1..4 all score -3 > -5 -> stored -3; positions 5,6 are 0
    assert O.pwm_gpu_argmax(seq, pw, ln, th)[0, 0] == 5


def test_summation_order_matters_and_is_pinned():
    """Position-major (GPU kernel, A19) and nucleotide-major (scan_n, A21) orders can round
    differently; the oracle keeps both.  Values chosen so that the two orders differ."""
    a, b, c = np.float32(1e8), np.float32(1.0), np.float32(-1e8)
    pwm = np.zeros((4, 3), np.float32)
    pwm[0, 0] = a      # A at position 0
    pwm[1, 1] = b      # C at position 1
    pwm[0, 2] = c      # A at position 2
    seq, pw, ln, th = _one(pwm, [0, 1, 0], -1e9)
    # position-major: (1e8 + 1) + -1e8 = 0 (1e8+1 rounds to 1e8); nucleotide-major: A first:
    # 1e8 + -1e8 = 0, then C: 0 + 1 = 1
    _, _, _, st, sc = O.pwm_hits(seq, pw, ln, th, both_strands=False)
    assert sc[0] == np.float32(0.0)
    pos, comp, s = O.pwm_best(seq, pw, ln, th)
    assert s[0, 0] == np.float32(1.0)


def test_best_score_is_sum_of_column_max():
    pwm = np.array([[0.5, -2, 1], [1.5, -1, 0], [0, -3, 2], [-1, -0.5, 0]], np.float32)
    assert O.best_score(pwm) == np.float32(1.5) + np.float32(-0.5) + np.float32(2.0)
