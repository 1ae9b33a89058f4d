"""GPU parity of the CDL ADMM against the CPU oracle (reference-structured FP32 and FP64)
and the committed golden fixtures.

Errors are relative norms |a-b|_2 / |b|_2 per array.  Tolerances:
  * short horizons (golden fixtures, single steps): GPU vs oracle <= 1e-4;
  * the objective seen by update_best at every outer iteration: <= 1e-4 relative to the
    FP32 oracle and to the FP64 oracle;
  * arrays after long schedules (100-200 iterations): GPU vs FP64 oracle
    <= max(1e-4, 1.5 x (FP32 oracle vs FP64 oracle)).  The iteration amplifies rounding:
    the reference-structured FP32 computation itself ends ~1e-3 away from FP64 on the
    single-dataset shape (L=200, M=64, filter_len=20), so two FP32 implementations with different operation orders
    cannot agree to 1e-4 there; the gate requires the GPU to be as accurate as the
    reference's own FP32 arithmetic.  All three errors are printed."""
import os

import numpy as np
import pytest

import oracle as O
from gpu_util import flush, rel

pytestmark = pytest.mark.gpu
TOL = 1e-4

cdlm = pytest.importorskip("cdlm")


def _synth(seed, N, L, M, fl):
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    from make_golden import make_cdl_inputs
    return make_cdl_inputs(seed, N, L, M, fl)


def envelope_check(tag, g, o32, o64, fields):
    for f in fields:
        a, b32, b64 = g.get(f), o32.get(f), o64.get(f)
        e_g64, e_3264, e_g32 = rel(a, b64), rel(b32, b64), rel(a, b32)
        flush("%s %s: gpu-f64 %.2e  f32oracle-f64 %.2e  gpu-f32oracle %.2e" % (tag, f, e_g64, e_3264, e_g32))
        assert e_g64 <= max(TOL, 1.5 * e_3264), (f, e_g64, e_3264)


def _both(s, D0, X0, fl, rho, sigma, lam):
    g = cdlm.CDL(s, fl, rho, sigma, lam, D0=D0, X0=X0)
    o32 = O.OracleCDL(s, D0, X0, fl, rho, sigma, lam, "f32")
    return g, o32


@pytest.mark.parametrize("case", ["cdl_tiny", "cdl_small", "cdl_mixedfl", "cdl_long"])
def test_golden_schedule(case, golden_dir):
    flush("case %s" % case)
    z = np.load(os.path.join(golden_dir, case + ".npz"))
    rho, sigma, lam = [float(x) for x in z["params"]]
    g = cdlm.CDL(z["s"], z["fl"], rho, sigma, lam, D0=z["D0"], X0=z["X0"])
    outer, ncsc, ndict = [int(x) for x in z["sched"]]
    objs = []
    for _ in range(outer):
        g.code_update(ncsc)
        o, _imp = g.update_best()
        objs.append(o)
        g.dict_update(ndict)
    np.testing.assert_allclose(objs, z["obj"], rtol=TOL)
    for f in ("X", "DX", "d", "G1", "G3", "y1", "y2", "y3", "Y1", "g1", "g2", "g3", "Td", "Dbest"):
        e = rel(g.get(f), z[f])
        assert e < TOL, (f, e)
    L = g.L
    xs = g.get("Xbest_slice")
    np.testing.assert_allclose(xs, z["Xbest"][:, :, g.gamma:g.gamma + 4 * L:4], rtol=0, atol=TOL * np.abs(z["Xbest"]).max())
    th, cnt = g.code_thresh(1.0)
    assert abs(th - float(z["thresh"])) <= TOL * float(z["thresh"])


def test_single_steps_match_oracle():
    fl = [16] * 5
    s, D0, X0, _ = _synth(21, 12, 25, 5, fl)
    g, o = _both(s, D0, X0, fl, 1.0, 100.0, 0.05)
    assert abs(g.objective() - o.objective()) <= TOL * abs(o.objective())
    for f in ("X", "DX", "d", "G1"):
        assert rel(g.get(f), o.get(f)) < 1e-6, f
    g.code_update(1); o.code_update(1)
    for f in ("X", "DX", "y1", "y2", "y3", "Y1"):
        assert rel(g.get(f), o.get(f)) < TOL, f
    ob_g, imp_g = g.update_best(); ob_o, imp_o = o.update_best()
    assert abs(ob_g - ob_o) <= TOL * abs(ob_o) and imp_g == imp_o
    g.dict_update(1); o.dict_update(1)
    for f in ("d", "g1", "g2", "g3", "G1", "G3", "G2", "Td", "Dbest"):
        assert rel(g.get(f), o.get(f)) < TOL, f


def test_c1_like_full_schedule_vs_oracle():
    """The small CPU-runnable workload's shape (L=100, M=10, filter_len=16, lambda=0.05, rho=1) with N=60,
    10 outer x (10 CSC + best + 1 dict) = 100 ADMM iterations."""
    fl = [16] * 10
    s, D0, X0, _ = _synth(1, 60, 100, 10, fl)
    g, o = _both(s, D0, X0, fl, 1.0, 100.0, 0.05)
    o64 = O.OracleCDL(s, D0, X0, fl, 1.0, 100.0, 0.05, "f64")
    tg, to, t64 = [], [], []
    for _ in range(10):
        g.code_update(10); o.code_update(10); o64.code_update(10)
        tg.append(g.update_best()[0]); to.append(o.update_best()[0]); t64.append(o64.update_best()[0])
        g.dict_update(1); o.dict_update(1); o64.dict_update(1)
    np.testing.assert_allclose(tg, to, rtol=TOL)
    np.testing.assert_allclose(tg, t64, rtol=TOL)
    envelope_check("C1-like", g, o, o64, ("X", "DX", "d", "G1", "G3", "Dbest"))
    tg_, cg = g.code_thresh(1.0)
    t64_, c64 = o64.code_thresh(1.0)
    to_, co = o.code_thresh(1.0)
    flush("C1-like code threshold gpu %.6f f32 %.6f f64 %.6f; positive codes %d %d %d" % (tg_, to_, t64_, cg, co, c64))
    assert abs(tg_ - t64_) <= max(TOL * t64_, 1.5 * abs(to_ - t64_))


def test_c2_shape_slice_vs_oracle():
    """Single-dataset shape per sequence (L=200, M=64, filter_len=20, Lg=840) on a
    24-sequence slice, 10 outer x (20 CSC + best + 1 dict) = 200 ADMM iterations."""
    fl = [20] * 64
    s, D0, X0, _ = _synth(2, 24, 200, 64, fl)
    g, o = _both(s, D0, X0, fl, 1.0, 100.0, 0.05)
    o64 = O.OracleCDL(s, D0, X0, fl, 1.0, 100.0, 0.05, "f64")
    tg = g.optimize(10, 20, 1)
    to, t64 = [], []
    for _ in range(10):
        for oo, tr in ((o, to), (o64, t64)):
            oo.code_update(20)
            tr.append(oo.update_best()[0])
            oo.dict_update(1)
    flush("C2-slice obj trace gpu %s" % np.array2string(np.asarray(tg), precision=6))
    np.testing.assert_allclose(tg, to, rtol=TOL)
    np.testing.assert_allclose(tg, t64, rtol=TOL)
    envelope_check("C2-slice", g, o, o64, ("X", "DX", "d", "G1", "G3", "Dbest"))


C2_SLICE = os.path.join(os.path.dirname(__file__), "golden", "c2_slice.npz")


@pytest.mark.skipif(not os.path.exists(C2_SLICE), reason="c2_slice.npz not generated")
@pytest.mark.parametrize("reps", [1, 20])
def test_c2_slice_1000_vs_oracle_fixture(reps):
    """A 1000-sequence dataset of the single-dataset shape (L=200, 64 filters of length 20,
    Lg=840) for the whole schedule, 10 x (20 CSC + best + 1 dict) = 200 ADMM iterations,
    against the FP32 and FP64 oracle results committed by tests/golden/make_c2_slice.py (the
    oracle takes ~45 min on 8 host cores for this size).

    reps=1 runs the slice as its own problem.  reps=20 runs it tiled 20 times, N=20000 at the
    full single-dataset size: every mean over n is then the slice's mean and the objective is
    exactly 20 times the slice's (checked on the oracle to 1e-15), so the full-size run must
    track the same fixture, and the copies of a sequence, computed independently by the
    per-sequence kernels with the same filters, must stay equal bit for bit."""
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import make_c2_slice as mk
    ref = np.load(C2_SLICE)
    s, D0, X0 = mk.inputs()
    for nm, a in (("s", s), ("D0", D0), ("X0", X0)):
        assert mk.digest(a) == str(ref["digest_" + nm]), "input %s differs from the fixture's" % nm
    if reps > 1:
        s, X0 = np.tile(s, (reps, 1)), np.tile(X0, (reps, 1, 1))
    tag = "C2-%d" % (mk.N * reps)
    flush("%s: N=%d" % (tag, s.shape[0]))
    fl = [mk.FLEN] * mk.M
    g = cdlm.CDL(s, fl, mk.RHO, mk.SIGMA, mk.LAM, D0=D0, X0=X0)
    del X0
    tg = np.asarray(g.optimize(mk.ITER_NUM, mk.CSC_ITER, mk.CDL_ITER)) / reps
    flush("%s obj trace gpu / %d %s" % (tag, reps, np.array2string(tg, precision=6)))
    np.testing.assert_allclose(tg, ref["trace_f32"], rtol=TOL)
    np.testing.assert_allclose(tg, ref["trace_f64"], rtol=TOL)
    for f in ("G1", "G3", "Dbest", "DX"):
        a = g.get(f)
        if f == "DX":
            for r in range(1, reps):
                np.testing.assert_array_equal(a[r * mk.N:r * mk.N + mk.N_DX], a[:mk.N_DX], err_msg="copy %d" % r)
            a = a[:mk.N_DX]
        b32, b64 = ref[f + "_f32"], ref[f + "_f64"]
        e_g64, e_3264, e_g32 = rel(a, b64), rel(b32, b64), rel(a, b32)
        flush("%s %s: gpu-f64 %.2e  f32oracle-f64 %.2e  gpu-f32oracle %.2e" % (tag, f, e_g64, e_3264, e_g32))
        # at this shape the GPU holds 1e-4 against FP64 outright (measured ~1e-5), tighter than the
        # FP32 oracle's own distance from FP64
        assert e_g64 <= TOL, (f, e_g64, e_3264)
    if "thresh_f64" not in ref.files:
        return
    # after thresholding (cdl_optimize, wrap.jl:12; record_training_result!, store.jl:15-19): the
    # code threshold within 1e-4 of FP64, the active sets equal FP64's except at codes within
    # 1e-3 of the threshold (counted and bounded), identical uinds, the X_best slice within 1e-4
    # of FP64.  The tiled run's copies are equal, its positive count is reps times the slice's and
    # its corrected std differs from the slice's by a factor 1 - O(1/count).
    import thresh_summary as TS
    th, cnt = g.code_thresh(1.0)
    xs = g.get("Xbest_slice")
    for r in range(1, reps):
        np.testing.assert_array_equal(xs[r * mk.N:(r + 1) * mk.N], xs[:mk.N], err_msg="X_best copy %d" % r)
    assert cnt % reps == 0
    TS.check(tag, xs[:mk.N], th, cnt // reps, ref, log=flush)


def test_csc_is_per_sequence_exact():
    """Size-independent property: a CSC iteration touches each sequence alone, so sequences
    5..8 of a 40-sequence problem equal the same 4 sequences run as their own problem, bit
    for bit (the dictionary is not updated)."""
    fl = [20] * 8
    s, D0, X0, _ = _synth(3, 40, 50, 8, fl)
    a = cdlm.CDL(s, fl, 1.0, 100.0, 0.05, D0=D0, X0=X0)
    b = cdlm.CDL(s[5:9], fl, 1.0, 100.0, 0.05, D0=D0, X0=X0[5:9])
    a.code_update(7); b.code_update(7)
    for f in ("X", "DX", "y2", "y3", "Y1", "y1"):
        np.testing.assert_array_equal(a.get(f)[5:9], b.get(f), err_msg=f)


@pytest.mark.parametrize("mfma,waves", [(1, 4), (1, 8), (1, 16), (2, 8), (2, 16)])
def test_csc_launch_split_is_exact(mfma, waves):
    """One launch of 12 CSC iterations carries the state between its iterations as
    (X + y2, X + y3) and rebuilds Xt, y2, y3 from it; 12 launches of one iteration each store
    (Xt, y2, y3) after every iteration.  Both run the same float operations, so every field
    must agree bit for bit, for each CSC kernel and work-group shape, and so must the next
    dictionary step."""
    fl = [20] * 12
    s, D0, X0, _ = _synth(6, 24, 200, 12, fl)
    L, M = 200, len(fl)
    cdlm.tuning_set(L, M, 20, {"fft_order": 0, "fft4_order": 0, "dict_batch": 2, "csc_slots": 4, "csc_waves": waves,
                               "csc_mfma": mfma})
    try:
        a = cdlm.CDL(s, fl, 1.0, 100.0, 0.05, D0=D0, X0=X0)
        b = cdlm.CDL(s, fl, 1.0, 100.0, 0.05, D0=D0, X0=X0)
    finally:
        cdlm.tuning_set(L, M, 20, None)
    a.code_update(12)
    for _ in range(12):
        b.code_update(1)
    for f in ("X", "DX", "y1", "Y1", "y2", "y3"):
        np.testing.assert_array_equal(a.get(f), b.get(f), err_msg=f)
    assert a.update_best() == b.update_best()
    a.dict_update(1); b.dict_update(1)
    for f in ("d", "G1", "G3", "g1", "g3"):
        np.testing.assert_array_equal(a.get(f), b.get(f), err_msg=f)


def test_projection_invariants():
    """y3 is exactly 0 on the stride-4 support after a CSC iteration (y3 += X - Pi4(X + y3)),
    G1 and G3 vanish off the filter support, G3 lies in [0, 1]."""
    fl = [12, 8, 16]
    s, D0, X0, _ = _synth(4, 10, 30, 3, fl)
    g = cdlm.CDL(s, fl, 1.0, 100.0, 0.05, D0=D0, X0=X0)
    g.code_update(3)
    y3 = g.get("y3")
    assert np.all(y3[:, :, 0::4] == 0)
    g.update_best(); g.dict_update(2)
    G1, G3 = g.get("G1"), g.get("G3")
    for m, f in enumerate(fl):
        assert np.all(G1[m, f:] == 0) and np.all(G3[m, f:] == 0)
    assert G3.min() >= 0 and G3.max() <= 1


def test_reset_reproduces():
    fl = [16] * 4
    s, D0, X0, _ = _synth(5, 8, 20, 4, fl)
    g = cdlm.CDL(s, fl, 1.0, 100.0, 0.05, D0=D0, X0=X0)
    t1 = g.optimize(3, 4, 1)
    x1 = g.get("X")
    g.reset()
    t2 = g.optimize(3, 4, 1)
    np.testing.assert_array_equal(t1, t2)
    np.testing.assert_array_equal(x1, g.get("X"))


def test_bad_arguments_raise():
    with pytest.raises(cdlm.CdlmError):
        cdlm.CDL(np.zeros((3, 10), np.float32), [16], 1.0, 1.0, 0.1)     # 4L not a multiple of 4
    with pytest.raises(cdlm.CdlmError):
        cdlm.CDL(np.zeros((3, 16), np.float32), [18], 1.0, 1.0, 0.1)     # filter_len not a multiple of 4
    with pytest.raises(cdlm.CdlmError):
        cdlm.CDL(np.zeros((3, 16), np.float32), [36], 1.0, 1.0, 0.1)     # beyond 32 taps


def test_x_materialisation_paths_agree():
    """X is written back materialised either by update_best's pass or, when no update_best
    ran since the last CSC iteration, by the dictionary step; both give the same state."""
    fl = [8, 12, 12]
    s, D0, X0, _ = _synth(41, 30, 25, 3, fl)
    a = cdlm.CDL(s, fl, 1.0, 100.0, 0.05, D0=D0, X0=X0)
    b = cdlm.CDL(s, fl, 1.0, 100.0, 0.05, D0=D0, X0=X0)
    for _ in range(2):
        a.code_update(4)
        a.update_best()
        a.dict_update(1)
        b.code_update(4)
        b.dict_update(1)
    for f in ("X", "d", "G1", "G3", "DX", "g1", "g2", "Td"):
        np.testing.assert_array_equal(a.get(f), b.get(f), err_msg=f)


@pytest.mark.parametrize("mfma", [0, 1, 2])
@pytest.mark.parametrize("shape", [(1, 2, 1, [4]), (3, 1, 2, [4, 4]), (2, 5, 3, [8, 4, 12]), (5, 3, 17, [8] * 17),
                                   (2, 252, 2, [8, 8]), (3, 260, 3, [8, 12, 8]), (1, 502, 2, [20, 20]),
                                   (2, 13, 40, [4] * 39 + [32])])
def test_tiny_shapes_vs_oracle(shape, mfma):
    """Edge shapes: one sequence, one base, one filter, fewer filters than the CSC kernel has
    waves (waves without columns), an odd filter count across the waves; padded lengths 1024
    (the first with more than 256 float4 slots of half spectrum, the dictionary kernel's wider
    variant, and the last with four 256-position blocks per column in the matrix-core CSC
    kernel), 1064 and 2048 (the largest), 116 with 32 taps (blocks mostly past the padded
    length); short schedule; each CSC kernel (1: vector ALUs, 2: matrix cores) and the default
    choice (0)."""
    N, L, M, fl = shape
    s, D0, X0, _ = _synth(40 + N + M, N, L, M, fl)
    if mfma:
        cdlm.tuning_set(L, M, max(fl), {"csc_mfma": mfma})
    try:
        g, o = _both(s, D0, X0, fl, 1.0, 100.0, 0.05)
    finally:
        cdlm.tuning_set(L, M, max(fl), None)
    for _ in range(2):
        g.code_update(4); o.code_update(4)
        ob_g, imp_g = g.update_best(); ob_o, imp_o = o.update_best()
        assert abs(ob_g - ob_o) <= TOL * abs(ob_o) and imp_g == imp_o, (ob_g, ob_o)
        g.dict_update(1); o.dict_update(1)
    for f in ("X", "DX", "y1", "y2", "y3", "Y1", "d", "g1", "g3", "G1", "G3"):
        e = rel(g.get(f), o.get(f))
        flush("tiny %s %s: %.2e" % (shape, f, e))
        assert e < TOL, (shape, f, e)
    g.close()
