"""The configured workloads on the GPU against committed oracle fixtures, through thresholding:

  * the CPU-runnable workload exactly as configured (N=500, L=100, 2 planted motifs, M=10,
    filter_len 16, 100 ADMM iterations, two trainings): objective traces, filters, code
    thresholds, the active sets, the activation record's filters, the topology counts and the
    pair PFMs of the pairs over the count threshold (tests/golden/make_c1.py);
  * datasets of the batched workload at their real L, M and filter lengths with a reduced
    number of sequences, 150 ADMM iterations each (tests/golden/make_c3.py), and the same
    datasets at their full sizes trained alone, as one batch and through the server with
    Poisson arrivals: batch and served results equal the results alone bit for bit.

Tolerances: objective at every update_best within 1e-4 relative of the FP32 and the FP64
oracle; filters against FP64 within max(1e-4, 1.5 x the FP32 oracle's own error); code
threshold within 1e-4 of FP64; active sets equal to FP64's except at codes within 1e-3 of the
threshold, which are counted (tests/golden/thresh_summary.py).  Integer outputs (uinds, record
filters, topology counts, pair lists) are compared exactly when the active sets are equal."""
import os
import sys
import threading
import time

import numpy as np
import pytest

import oracle as O
from gpu_util import flush, rel

pytestmark = pytest.mark.gpu
TOL = 1e-4

cdlm = pytest.importorskip("cdlm")
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
sys.path.insert(0, GOLDEN)
import thresh_summary as TS   # noqa: E402


def _envelope(tag, a, b32, b64):
    e_g64, e_3264 = rel(a, b64), rel(b32, b64)
    flush("%s: gpu-f64 %.2e  f32oracle-f64 %.2e  gpu-f32oracle %.2e" % (tag, e_g64, e_3264, rel(a, b32)))
    assert e_g64 <= max(TOL, 1.5 * e_3264), (tag, e_g64, e_3264)


@pytest.mark.skipif(not os.path.exists(os.path.join(GOLDEN, "c1.npz")), reason="c1.npz not generated")
def test_cpu_workload_as_configured_through_topology_and_pair_pfms():
    import make_c1 as mk
    z = np.load(os.path.join(GOLDEN, "c1.npz"))
    s, inits = mk.inputs()
    assert mk.digest(s) == str(z["digest_s"])
    fl = [mk.FLEN] * mk.M
    rec = cdlm.ActivationRecord(mk.N, mk.L, mk.M, max_trainings=mk.TRAININGS)
    flips = 0
    for t, (D0, X0) in enumerate(inits):
        assert mk.digest(D0) == str(z["digest_D0_%d" % t]) and mk.digest(X0) == str(z["digest_X0_%d" % t])
        flush("CPU workload training %d: N=%d L=%d M=%d filter_len=%d" % (t, mk.N, mk.L, mk.M, mk.FLEN))
        g = cdlm.CDL(s, fl, mk.RHO, mk.SIGMA, mk.LAM, D0=D0, X0=X0)
        tr = g.optimize(mk.ITER_NUM, mk.CSC_ITER, mk.CDL_ITER)
        np.testing.assert_allclose(tr, z["trace_f32_%d" % t], rtol=TOL)
        np.testing.assert_allclose(tr, z["trace_f64_%d" % t], rtol=TOL)
        for f in ("G1", "G3"):
            _envelope("training %d %s" % (t, f), g.get(f), z["%s_f32_%d" % (f, t)], z["%s_f64_%d" % (f, t)])
        th, cnt = g.code_thresh(1.0)
        sub = {k[:-2]: z[k] for k in z.files if k.endswith("_%d" % t)}
        flips += TS.check("training %d" % t, g.get("Xbest_slice"), th, cnt, sub, log=flush)
        assert rec.add(g, th, [mk.FILTER_SIZE] * mk.M) == len(sub["uinds_f64"])
        g.close()
    u, tn, fls = rec.filters()
    np.testing.assert_array_equal(u, z["rec_uinds"])
    np.testing.assert_array_equal(tn, z["rec_train"])
    np.testing.assert_array_equal(fls, z["rec_lens"])
    over, non, nxt = rec.topology()
    ref = (z["top_over"], z["top_non"], z["top_next"])
    diff = sum(int(np.abs(a.astype(np.int64) - b).sum()) for a, b in zip((over, non, nxt), ref))
    flush("topology counts %d/%d/%d (fp64 oracle %d/%d/%d), |difference| %d, flipped activities %d" % (
        over.sum(), non.sum(), nxt.sum(), ref[0].sum(), ref[1].sum(), ref[2].sum(), diff, flips))
    if flips == 0:
        for a, b in zip((over, non, nxt), ref):
            np.testing.assert_array_equal(a, b)
    else:
        # the record's topology is exact for its own active sets; a flipped activity moves at most
        # the pairs it takes part in: 2 x K x C of them
        sp = rec.start_positions()
        own = O.topology(sp, tn, fls, np.ones((mk.TRAININGS, mk.N), np.int32), int(fls.max()), mk.L - int(fls.min()) - 1)
        for a, b in zip((over, non, nxt), own):
            np.testing.assert_array_equal(a, b)
        assert diff <= flips * 2 * len(u) * mk.L
    pairs = cdlm.overlap_pairs(over, mk.COUNT_THRESHOLD)
    flush("pairs over %d: %d (fp64 oracle %d)" % (mk.COUNT_THRESHOLD, len(pairs), len(z["pairs"])))
    if flips == 0:
        np.testing.assert_array_equal(pairs, z["pairs"])
        span, pfm = rec.pair_pfms(s, pairs, 2 * mk.FILTER_SIZE - 1)
        np.testing.assert_array_equal(span, z["pair_span"])
        e = rel(pfm, z["pair_pfms"])
        flush("pair PFMs: %d, rel err vs fp64 %.2e" % (len(pairs), e))
        assert e <= TOL
        np.testing.assert_allclose(pfm, z["pair_pfms"], rtol=TOL, atol=1e-6)
    rec.close()


C3 = os.path.join(GOLDEN, "c3.npz")


@pytest.mark.skipif(not os.path.exists(C3), reason="c3.npz not generated")
def test_batched_workload_shapes_vs_oracle_fixture():
    import make_c3 as mk
    z = np.load(C3)
    for i, sp in enumerate(mk.specs()):
        s, D0, X0 = mk.inputs(sp)
        assert mk.digest(s) + mk.digest(D0) + mk.digest(X0) == str(z["digest_%d" % i])
        fl = sp["filter_len"]
        tag = "dataset %d (L=%d M=%d filter_len=%d Lg=%d, N=%d)" % (i, sp["L"], sp["M"], fl, 4 * sp["L"] + 2 * fl,
                                                                   mk.N_SMALL)
        flush(tag)
        g = cdlm.CDL(s, [fl] * sp["M"], mk.RHO, mk.SIGMA, mk.LAM, D0=D0, X0=X0)
        tr = g.optimize(mk.ITER_NUM, mk.CSC_ITER, mk.CDL_ITER)
        np.testing.assert_allclose(tr, z["trace_f32_%d" % i], rtol=TOL)
        np.testing.assert_allclose(tr, z["trace_f64_%d" % i], rtol=TOL)
        for f in ("G1", "G3"):
            _envelope("%s %s" % (tag, f), g.get(f)[:, :fl], z["%s_f32_%d" % (f, i)], z["%s_f64_%d" % (f, i)])
        th, cnt = g.code_thresh(1.0)
        sub = {k[:-len("_%d" % i)]: z[k] for k in z.files if k.endswith("_%d" % i)}
        TS.check(tag, g.get("Xbest_slice"), th, cnt, sub, log=flush)
        g.close()


def _full_specs():
    import make_c3 as mk
    return mk, mk.specs()


def test_batched_workload_full_size_batch_and_served_equal_alone():
    """The batched workload's first datasets at their full sizes (N up to 18873, padded lengths
    432-2048, 123 GB of device state together): each trained alone, then all as one batch (one
    graph launch), then submitted to the server at Poisson(5 jobs/s) arrival times; every trace,
    threshold, positive count, best objective and X_best slice is the same bit for bit."""
    mk, specs = _full_specs()
    jobs = []
    for sp in specs:
        s, D0, _ = mk.inputs(sp, N=sp["N"])
        jobs.append((sp, s, D0))
    alone = []
    for sp, s, D0 in jobs:
        flush("alone: seed %d N=%d L=%d M=%d filter_len=%d" % (sp["seed"], sp["N"], sp["L"], sp["M"], sp["filter_len"]))
        g = cdlm.CDL(s, [sp["filter_len"]] * sp["M"], mk.RHO, mk.SIGMA, mk.LAM, D0=D0, seed=sp["seed"])
        tr = g.optimize(mk.ITER_NUM, mk.CSC_ITER, mk.CDL_ITER)
        th, cnt = g.code_thresh(1.0)
        alone.append((tr, th, cnt, g.obj_best(), g.get("Xbest_slice")))
        g.close()
    flush("batch of %d datasets" % len(jobs))
    ps = [cdlm.CDL(s, [sp["filter_len"]] * sp["M"], mk.RHO, mk.SIGMA, mk.LAM, D0=D0, seed=sp["seed"])
          for sp, s, D0 in jobs]
    b = cdlm.Batch(ps)
    tr, th, cnt = b.train(mk.ITER_NUM, mk.CSC_ITER, mk.CDL_ITER, 1.0)
    flush("  %d shape classes" % b.classes)
    for i, p in enumerate(ps):
        np.testing.assert_array_equal(tr[i], alone[i][0])
        assert th[i] == alone[i][1] and cnt[i] == alone[i][2]
        assert p.obj_best() == alone[i][3]
        np.testing.assert_array_equal(p.get("Xbest_slice"), alone[i][4])
    b.close()
    for p in ps:
        p.close()
    del ps
    flush("served at Poisson(5 jobs/s)")
    srv = cdlm.Server(csc_iter=mk.CSC_ITER, cdl_iter=mk.CDL_ITER)
    rng = np.random.default_rng(77)
    gaps = rng.exponential(1 / 5.0, len(jobs))
    ids = [None] * len(jobs)

    def submit_all():
        for i, (sp, s, D0) in enumerate(jobs):
            time.sleep(gaps[i])
            ids[i] = srv.submit(s, [sp["filter_len"]] * sp["M"], mk.RHO, mk.SIGMA, mk.LAM, D0=D0, seed=sp["seed"],
                                iter_num=mk.ITER_NUM, keep_codes=True)

    th_ = threading.Thread(target=submit_all)
    th_.start()
    th_.join()
    for i, jid in enumerate(ids):
        assert srv.wait(jid, timeout_s=600) == cdlm.serve.DONE
        res = srv.result(jid, codes=True)
        np.testing.assert_array_equal(res["trace"], alone[i][0])
        assert res["thresh"] == alone[i][1] and res["count"] == alone[i][2] and res["obj_best"] == alone[i][3]
        np.testing.assert_array_equal(res["codes"], alone[i][4])
        flush("  job %d: latency %.2f s (queue %.2f s)" % (i, res["latency_s"], res["queue_s"]))
    srv.close()
