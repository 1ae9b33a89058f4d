"""GPU PWM scans: bit-exact against the CPU oracle and the golden fixture."""
import os

import numpy as np
import pytest

import oracle as O
from gpu_util import flush

pytestmark = pytest.mark.gpu
cdlm = pytest.importorskip("cdlm")


def _fixture(golden_dir):
    return np.load(os.path.join(golden_dir, "pwm_small.npz"))


def test_hits_golden(golden_dir):
    z = _fixture(golden_dir)
    h = cdlm.scan_hits(z["bases"], z["pwms"], z["lens"], z["thresh"], both_strands=True)
    np.testing.assert_array_equal(h, z["hits"])


def test_argmax_golden(golden_dir):
    z = _fixture(golden_dir)
    np.testing.assert_array_equal(cdlm.scan_w_gpu(z["bases"], z["pwms"], z["lens"], z["thresh"]), z["argmax_pos"])


def test_best_golden(golden_dir):
    z = _fixture(golden_dir)
    pos, comp, sc = cdlm.overlapping_scan(z["bases"], z["pwms"], z["lens"], z["thresh"])
    np.testing.assert_array_equal(pos, z["best_pos"])
    np.testing.assert_array_equal(comp.astype(np.int8), z["best_comp"])
    np.testing.assert_array_equal(sc.view(np.uint32), z["best_score"].view(np.uint32))


def _random_case(seed, N, L, K, lmin=8, lmax=24):
    rng = np.random.default_rng(seed)
    lens = rng.integers(lmin, lmax + 1, K).astype(np.int32)
    maxlen = int(lens.max())
    pwms = np.zeros((K, 4, maxlen), np.float32)
    for k in range(K):
        p = rng.dirichlet(0.5 * np.ones(4), lens[k])
        p = (p + 0.01) / 1.04
        pwms[k, :, :lens[k]] = np.log2(p / 0.25).T.astype(np.float32)
    bases = rng.choice(4, size=(N, L), p=[0.29, 0.21, 0.21, 0.29]).astype(np.uint8)
    thresh = np.array([np.float32(0.8) * O.best_score(pwms[k, :, :lens[k]]) for k in range(K)], np.float32)
    return bases, pwms, lens, thresh


def test_hits_random_vs_oracle():
    bases, pwms, lens, thresh = _random_case(5, 300, 500, 24)
    k, n, p, st, _ = O.pwm_hits(bases, pwms, lens, thresh, both_strands=True)
    exp = np.stack([k, n, p, st], 1).astype(np.int64)
    got = cdlm.scan_hits(bases, pwms, lens, thresh, both_strands=True)
    flush("random hits: %d" % len(exp))
    assert len(exp) > 0
    np.testing.assert_array_equal(got, exp)


def test_argmax_and_best_random_vs_oracle():
    bases, pwms, lens, thresh = _random_case(6, 200, 300, 16)
    np.testing.assert_array_equal(cdlm.scan_w_gpu(bases, pwms, lens, thresh),
                                  O.pwm_gpu_argmax(bases, pwms, lens, thresh))
    pos, comp, sc = cdlm.overlapping_scan(bases, pwms, lens, thresh)
    opos, ocomp, osc = O.pwm_best(bases, pwms, lens, thresh)
    np.testing.assert_array_equal(pos, opos)
    np.testing.assert_array_equal(comp.astype(np.int8), ocomp)
    np.testing.assert_array_equal(sc.view(np.uint32), osc.view(np.uint32))


def test_rare_and_negative_thresholds_vs_oracle():
    """Thresholds in the upper tail of strongly negative scores: windows above the threshold are
    rare (most 32-window chunks of a tile have none, the others some), and many of them score
    below 0, so the argmax's running best can be negative and a later window of value 0 (below
    its threshold) replaces it; hit list, argmax and best hit bit for bit against the oracle."""
    rng = np.random.default_rng(31)
    K, N, L = 20, 120, 333
    lens = rng.integers(6, 17, K).astype(np.int32)
    pwms = np.zeros((K, 4, int(lens.max())), np.float32)
    for k in range(K):
        pwms[k, :, :lens[k]] = (rng.standard_normal((4, lens[k])) * 0.5 - 1.0).astype(np.float32)
    bases = rng.integers(0, 4, (N, L)).astype(np.uint8)
    # per PWM, the score exceeded by about 0.3% of random windows (mean -m, sd ~0.5 sqrt(m))
    thresh = np.array([np.float32(-lens[k] + 2.75 * 0.5 * np.sqrt(lens[k])) for k in range(K)], np.float32)
    k_, n_, p_, st_, sc_ = O.pwm_hits(bases, pwms, lens, thresh, both_strands=True)
    flush("rare thresholds: %d hits, %d with a negative score" % (len(k_), int((np.asarray(sc_) < 0).sum())))
    assert 0 < len(k_) < N * L * K * 2 // 50
    assert (np.asarray(sc_) < 0).any()
    exp = np.stack([k_, n_, p_, st_], 1).astype(np.int64)
    exp = exp[np.lexsort((exp[:, 3], exp[:, 2], exp[:, 1], exp[:, 0]))]
    np.testing.assert_array_equal(cdlm.scan_hits(bases, pwms, lens, thresh, both_strands=True), exp)
    np.testing.assert_array_equal(cdlm.scan_w_gpu(bases, pwms, lens, thresh), O.pwm_gpu_argmax(bases, pwms, lens, thresh))
    pos, comp, sc = cdlm.overlapping_scan(bases, pwms, lens, thresh)
    opos, ocomp, osc = O.pwm_best(bases, pwms, lens, thresh)
    np.testing.assert_array_equal(pos, opos)
    np.testing.assert_array_equal(comp.astype(np.int8), ocomp)
    np.testing.assert_array_equal(sc.view(np.uint32), osc.view(np.uint32))


def test_edge_cases():
    rng = np.random.default_rng(9)
    pwms = np.zeros((2, 4, 12), np.float32)
    pwms[0, :, :12] = rng.standard_normal((4, 12))
    pwms[1, :, :3] = rng.standard_normal((4, 3))
    lens = np.array([12, 3], np.int32)
    thresh = np.array([-100.0, -100.0], np.float32)
    # sequences shorter than the PWM: no window for k=0
    bases = rng.integers(0, 4, (5, 8)).astype(np.uint8)
    h = cdlm.scan_hits(bases, pwms, lens, thresh)
    assert np.all(h[:, 0] == 1) and len(h) == 5 * 6 * 2
    pos, comp, sc = cdlm.overlapping_scan(bases, pwms, lens, thresh)
    assert np.all(pos[0] == 0)
    np.testing.assert_array_equal(cdlm.scan_w_gpu(bases, pwms, lens, thresh), O.pwm_gpu_argmax(bases, pwms, lens, thresh))
    # an unknown base (code 4, an all-zero one-hot column) contributes 0
    bases2 = bases.copy(); bases2[:, 2] = 4
    k, n, p, st, _ = O.pwm_hits(bases2, pwms, lens, thresh)
    got = cdlm.scan_hits(bases2, pwms, lens, thresh)
    np.testing.assert_array_equal(got, np.stack([k, n, p, st], 1).astype(np.int64))


def test_onehot_to_bases_roundtrip():
    rng = np.random.default_rng(3)
    b = rng.integers(0, 4, (7, 33)).astype(np.uint8)
    oh = np.zeros((7, 4 * 33), np.float32)
    oh[np.arange(7)[:, None], 4 * np.arange(33)[None, :] + b] = 1
    oh[2, 4 * 5:4 * 6] = 0
    got = cdlm.onehot_to_bases(oh)
    exp = b.copy(); exp[2, 5] = 4
    np.testing.assert_array_equal(got, exp)


def test_mfma_layout_and_rounding_probe():
    """The operand layout the hit-list kernel assumes for v_mfma_f32_16x16x4_f32, and that
    an accumulation with one non-zero product per K-block rounds like one FP32 addition."""
    import ctypes as C
    from cdlm import _lib
    ctx = cdlm.default_context()
    rng = np.random.default_rng(11)
    a = rng.standard_normal(64).astype(np.float32)
    b = rng.standard_normal(64).astype(np.float32)
    c = rng.standard_normal(256).astype(np.float32)
    d = np.empty(256, np.float32)
    _lib.check(_lib.load().cdlm_selftest_mfma(ctx.h, a.ctypes.data, b.ctypes.data, c.ctypes.data, d.ctypes.data), ctx.h)
    A = np.zeros((16, 4)); B = np.zeros((4, 16)); Cm = np.zeros((16, 16))
    for l in range(64):
        A[l % 16, l // 16] = a[l]
        B[l // 16, l % 16] = b[l]
        for r in range(4):
            Cm[4 * (l // 16) + r, l % 16] = c[4 * l + r]
    ref = A @ B + Cm
    got = np.zeros((16, 16))
    for l in range(64):
        for r in range(4):
            got[4 * (l // 16) + r, l % 16] = d[4 * l + r]
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-5)
    # one-hot A: exactly one non-zero product per output -> result must be fl32(c + p)
    a1 = np.zeros(64, np.float32)
    for l in range(64):
        a1[l] = 1.0 if (l // 16) == (l % 16) % 4 else 0.0     # row i selects k = i % 4
    b1 = (rng.standard_normal(64) * 3).astype(np.float32)
    c1 = (rng.standard_normal(256) * 1e3).astype(np.float32)
    _lib.check(_lib.load().cdlm_selftest_mfma(ctx.h, a1.ctypes.data, b1.ctypes.data, c1.ctypes.data, d.ctypes.data), ctx.h)
    for l in range(64):
        for r in range(4):
            i, j = 4 * (l // 16) + r, l % 16
            p = b1[(i % 4) * 16 + j]                  # B[k=i%4][j]
            exp = np.float32(np.float32(c1[4 * l + r]) + p)
            assert d[4 * l + r].view(np.uint32) == exp.view(np.uint32), (l, r)


@pytest.mark.parametrize("L,lmin,lmax,frac", [(300, 30, 64, 0.0), (77, 8, 33, 0.35), (64, 1, 16, 0.8),
                                              (960, 8, 24, 0.8), (961, 8, 24, 0.8), (1500, 8, 24, 0.8),
                                              (1200, 30, 64, 0.2)])
def test_hits_long_pwms_and_ragged_lengths_vs_oracle(L, lmin, lmax, frac):
    """PWMs longer than 32 columns (the 64-column kernel), sequence lengths that are not a
    multiple of the kernel's window chunk, single-column PWMs; sequences up to 960 bases with
    PWMs of at most 32 columns are held in LDS as one-hot floats, from 961 on (and 1200 with
    64-column PWMs) as bytes; thresholds frac x best score.  Hit list, argmax and best hit."""
    bases, pwms, lens, _ = _random_case(L + lmax, 90, L, 20, lmin, lmax)
    thresh = np.array([np.float32(frac) * O.best_score(pwms[k, :, :lens[k]]) for k in range(len(lens))], np.float32)
    k, n, p, st, _ = O.pwm_hits(bases, pwms, lens, thresh, both_strands=True)
    exp = np.stack([k, n, p, st], 1).astype(np.int64)
    exp = exp[np.lexsort((exp[:, 3], exp[:, 2], exp[:, 1], exp[:, 0]))]
    got = cdlm.scan_hits(bases, pwms, lens, thresh, both_strands=True)
    flush("L=%d lens %d-%d: %d hits" % (L, lmin, lmax, len(exp)))
    assert len(exp) > 0
    np.testing.assert_array_equal(got, exp)
    # the best-window scans on the same tiles: argmax (position-major) and best hit
    # (nucleotide-major, both strands)
    np.testing.assert_array_equal(cdlm.scan_w_gpu(bases, pwms, lens, thresh),
                                  O.pwm_gpu_argmax(bases, pwms, lens, thresh))
    pos, comp, sc = cdlm.overlapping_scan(bases, pwms, lens, thresh)
    opos, ocomp, osc = O.pwm_best(bases, pwms, lens, thresh)
    np.testing.assert_array_equal(pos, opos)
    np.testing.assert_array_equal(comp.astype(np.int8), ocomp)
    np.testing.assert_array_equal(sc.view(np.uint32), osc.view(np.uint32))


def _oracle_by_pwm_chunks(bases, pwms, lens, thresh, workers=16):
    """The oracle's three scans over chunks of PWMs in parallel threads (the C calls release
    the interpreter lock); results reassembled in PWM order."""
    from concurrent.futures import ThreadPoolExecutor
    K = len(lens)
    bounds = np.linspace(0, K, workers + 1).astype(int)
    chunks = [(a, b) for a, b in zip(bounds[:-1], bounds[1:]) if b > a]

    def one(ab):
        a, b = ab
        sl = (bases, pwms[a:b], lens[a:b], thresh[a:b])
        k, n, p, st, _ = O.pwm_hits(*sl, both_strands=True)
        hits = np.stack([k + a, n, p, st], 1).astype(np.int64)
        return hits, O.pwm_gpu_argmax(*sl), O.pwm_best(*sl)

    with ThreadPoolExecutor(len(chunks)) as ex:
        parts = list(ex.map(one, chunks))
    hits = np.concatenate([p[0] for p in parts])
    hits = hits[np.lexsort((hits[:, 3], hits[:, 2], hits[:, 1], hits[:, 0]))]
    argmax = np.concatenate([p[1] for p in parts])
    best = tuple(np.concatenate([p[2][i] for p in parts]) for i in range(3))
    return hits, argmax, best


def test_pwm_scan_workload_all_512_pwms_vs_oracle():
    """The PWM-scan workload as configured (512 PWMs of length 8-24 with Dirichlet(0.5) columns,
    background 0.25, thresholds 0.8 x best score; sequences of L=500 with GC fraction 0.42, both
    strands) on 1000 sequences: hit list, argmax and best hit bit for bit against the oracle, so
    every one of the 64 tiles of 16 strand-PWMs (and 32 tiles of 32 for the best hit) is covered."""
    from cdlm import synth
    bases, pwms, lens, thresh = synth.pwm_workload(5, 1000, 500, 512)
    flush("PWM workload: 1000 x 500, 512 PWMs, oracle in parallel chunks")
    hits, argmax, (opos, ocomp, osc) = _oracle_by_pwm_chunks(bases, pwms, lens, thresh)
    got = cdlm.scan_hits(bases, pwms, lens, thresh, both_strands=True)
    flush("  %d hits (oracle %d), %d PWMs with a hit" % (len(got), len(hits), len(np.unique(hits[:, 0]))))
    assert len(hits) > 1000
    np.testing.assert_array_equal(got, hits)
    np.testing.assert_array_equal(cdlm.scan_w_gpu(bases, pwms, lens, thresh), argmax)
    pos, comp, sc = cdlm.overlapping_scan(bases, pwms, lens, thresh)
    np.testing.assert_array_equal(pos, opos)
    np.testing.assert_array_equal(comp.astype(np.int8), ocomp)
    np.testing.assert_array_equal(sc.view(np.uint32), osc.view(np.uint32))
