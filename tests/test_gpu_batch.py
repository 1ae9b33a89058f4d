"""Batched CDL: several problems of different N, L, M and filter lengths run with one launch
per kernel (and from a captured HIP graph).  Each problem's results in the batch must be
bit-identical to the results it gets alone: the batch only changes which block runs which
sequence, the zero taps past a problem's own filter length add exact zeros, and every
accumulation keeps its order.  Problems alone are themselves checked against the oracle in
test_gpu_cdl.py.  The reference trains one dataset per cdl_optimize call
(src/WRAPPER/wrap.jl:4-12), so "alone" is the reference's behaviour."""
import numpy as np
import pytest

from gpu_util import flush

pytestmark = pytest.mark.gpu
cdlm = pytest.importorskip("cdlm")

# (seed, N, L, filter lengths): padded lengths 20 .. 1056, taps 4 .. 20; the last one needs
# the 8-slot CSC variant (Lg/4 > 256), the 20-sample ones wrap their halos more than once
# when the batch runs 20 taps.
CASES = [
    (101, 40, 30, [8] * 5),
    (102, 25, 50, [20] * 8),
    (103, 60, 12, [4, 8, 12]),
    (104, 10, 3, [4, 4]),
    (105, 6, 260, [8, 8]),
]
RHO, SIGMA, LAM = 1.0, 100.0, 0.05
FIELDS = ("X", "d", "G1", "G3", "DX", "Xbest_slice", "Dbest", "y2", "g2")


def _data(seed, N, L):
    from cdlm import synth
    bases, _ = synth.planted_bases(seed, N, L, [6, 8], 0.3)
    return synth.onehot(bases)


def _make(cases, ctx=None):
    return [cdlm.CDL(_data(sd, N, L), fl, RHO, SIGMA, LAM, seed=sd, ctx=ctx) for sd, N, L, fl in cases]


def _state(c):
    return {f: c.get(f) for f in FIELDS}


def _same(tag, a, b):
    for f in FIELDS:
        assert np.array_equal(a[f], b[f]), (tag, f, float(np.abs(a[f] - b[f]).max()))


def test_batch_train_equals_alone():
    flush("batch train vs alone: %d problems" % len(CASES))
    alone = _make(CASES)
    ref = []
    for c in alone:
        tr = c.optimize(3, 5, 1)
        th, cnt = c.code_thresh(1.0)
        ref.append((tr, th, cnt, _state(c)))
    probs = _make(CASES)
    b = cdlm.Batch(probs)
    tr, th, cnt = b.train(3, 5, 1, 1.0)
    for i, c in enumerate(probs):
        flush("problem %d: trace %s" % (i, tr[i]))
        assert np.array_equal(tr[i], ref[i][0]), i
        assert th[i] == ref[i][1] and cnt[i] == ref[i][2], i
        _same(i, _state(c), ref[i][3])
    # continuing the schedule on the batch (no reset) equals continuing alone
    t2 = b.optimize(2, 4, 2)
    for i, c in enumerate(alone):
        assert np.array_equal(t2[i], c.optimize(2, 4, 2)), i
        _same(i, _state(probs[i]), _state(c))
    b.close()


def test_graph_replays_equal_eager():
    """Every replay of the captured graph equals the eager launches, with other eager work
    and synchronisations between replays."""
    cases = CASES[:4]
    probs = _make(cases)
    b = cdlm.Batch(probs, graph=False)
    tr0, th0, cnt0 = b.train(2, 6, 1, 1.0)
    st0 = [_state(c) for c in probs]
    b.set_graph(True)
    other = _make(CASES[1:2])[0]
    for rep in range(3):
        flush("graph replay %d" % rep)
        tr, th, cnt = b.train(2, 6, 1, 1.0)
        assert np.array_equal(tr, tr0) and np.array_equal(th, th0) and np.array_equal(cnt, cnt0), rep
        for i, c in enumerate(probs):
            _same((rep, i), _state(c), st0[i])
        other.code_update(3)        # eager work on the same stream between replays
        other.update_best()
    b.close()


def test_batch_single_problem_matches_problem_api():
    (c,) = _make(CASES[1:2])
    (d,) = _make(CASES[1:2])
    b = cdlm.Batch([d])
    t1 = c.optimize(2, 5, 1)
    t2 = b.optimize(2, 5, 1)
    assert np.array_equal(t1, t2[0])
    _same("single", _state(c), _state(d))
    b.close()


def test_batch_lifetime_and_arguments():
    probs = _make(CASES[2:4])
    b = cdlm.Batch(probs)
    with pytest.raises(cdlm.CdlmError):
        probs[0].close()                      # still part of a batch
    with pytest.raises(cdlm.CdlmError):
        cdlm.Batch([probs[0], probs[0]])      # duplicate
    b.close()
    probs[0].close()
    probs[1].close()


def test_footprint_is_what_a_problem_holds():
    from cdlm import _lib
    L = _lib.load()
    for sd, N, Lb, fl in CASES:
        c = cdlm.CDL(_data(sd, N, Lb), fl, RHO, SIGMA, LAM, seed=sd)
        assert c.device_bytes == L.cdlm_footprint_bytes(4 * Lb, N, len(fl), max(fl), 0)
        c.close()
    X0 = np.zeros((6, 2, 12 + 8), np.float32)
    c = cdlm.CDL(_data(1, 6, 3), [4, 4], RHO, SIGMA, LAM, X0=X0)
    assert c.device_bytes == L.cdlm_footprint_bytes(12, 6, 2, 4, 1)


def test_tuned_launch_choices_keep_parity():
    """Every stage order, dictionary work-group width (2, 4 or 8 waves, one column each) and
    CSC shape (slots per lane, waves per work-group) meets the oracle tolerance; the autotuner's
    choice is recorded and used by problems created afterwards; a batch of problems with
    different tunings still equals each problem alone."""
    import oracle as O
    from gpu_util import rel
    L, M, fl = 31, 4, [20, 12, 20, 8]            # Lg = 124 + 40 = 164 = 4 * 41
    s = _data(77, 12, L)
    D0 = cdlm.synth.rand_probvec_filters(9, M, 4 * L + 40, fl)
    X0 = np.random.default_rng(10).standard_normal((12, M, 4 * L + 40)).astype(np.float32)
    o = O.OracleCDL(s, D0, X0, fl, RHO, SIGMA, LAM, "f32")
    o.code_update(6)
    o.update_best()
    o.dict_update(1)
    o.code_update(3)
    choices = [dict(fft_order=order, fft4_order=order4, dict_batch=B, csc_slots=8, csc_waves=4, csc_mfma=1)
               for order in (0, 1) for order4 in (0, 1) for B in (2, 4, 8)]
    choices += [dict(fft_order=0, fft4_order=0, dict_batch=2, csc_slots=4, csc_waves=wv, csc_mfma=1) for wv in (4, 8, 16)]
    choices += [dict(fft_order=o, fft4_order=0, dict_batch=4, csc_slots=0, csc_waves=wv, csc_mfma=2)
                for o in (0, 1) for wv in (8, 16)]
    try:
        for ch in choices:
            cdlm.tuning_set(L, M, 20, ch)
            assert cdlm.tuning_get(L, M, 20) == ch
            c = cdlm.CDL(s, fl, RHO, SIGMA, LAM, D0=D0, X0=X0)
            c.code_update(6)
            c.update_best()
            c.dict_update(1)
            c.code_update(3)
            for f in ("X", "DX", "d", "G1"):
                e = rel(c.get(f), o.get(f))
                flush("%s %s: %.2e" % (ch, f, e))
                assert e <= 1e-4, (ch, f, e)
            c.close()
        got = cdlm.autotune(L, M, 20, n_seq=64)
        flush("autotune: %s" % got)
        assert got == cdlm.tuning_get(L, M, 20)
        a = cdlm.CDL(s, fl, RHO, SIGMA, LAM, D0=D0, X0=X0)
        cdlm.tuning_set(L, M, 20, None)
        b_alone = cdlm.CDL(s, fl, RHO, SIGMA, LAM, D0=D0, X0=X0)
        ref = [a.optimize(2, 3, 1), b_alone.optimize(2, 3, 1)]
        refx = [a.get("X"), b_alone.get("X")]
        a2 = cdlm.CDL(s, fl, RHO, SIGMA, LAM, D0=D0, X0=X0)         # untuned (cleared)
        cdlm.tuning_set(L, M, 20, got)
        b2 = cdlm.CDL(s, fl, RHO, SIGMA, LAM, D0=D0, X0=X0)         # tuned
        bt = cdlm.Batch([b2, a2])
        tr = bt.optimize(2, 3, 1)
        assert np.array_equal(tr[0], ref[0]) and np.array_equal(tr[1], ref[1])
        assert np.array_equal(b2.get("X"), refx[0]) and np.array_equal(a2.get("X"), refx[1])
        bt.close()
    finally:
        cdlm.tuning_set(L, M, 20, None)
