"""Streamed problems (cdlm_problem_create_streamed): a problem whose footprint exceeds its device
budget keeps its per-sequence state in pinned host memory and moves the sequences through two
device windows in chunks (the next chunk's upload and the last one's download overlapping the
current chunk's kernels), or through one window when two do not fit.  The per-filter state and the fixed-order reductions stay on the
device, so every result must equal the resident problem's bit for bit: traces, objectives,
thresholds, every field, after single steps, whole schedules and a reset.  The server streams a
job larger than its budget the same way."""
import os

import numpy as np
import pytest

from gpu_util import flush

pytestmark = pytest.mark.gpu

cdlm = pytest.importorskip("cdlm")

FIELDS = ("X", "DX", "d", "G1", "G2", "G3", "Xbest_slice", "Dbest", "Y1", "y1", "y2", "y3", "g1", "g2", "g3", "Td")


def _synth(seed, N, L, M, fl):
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    from make_golden import make_cdl_inputs
    return make_cdl_inputs(seed, N, L, M, fl)


def _budget(L4, N, M, gamma, x0, chunks):
    """Device bytes for a window of about N / chunks sequences (the per-filter part and the
    per-sequence objective partials of all N come on top of the window)."""
    fp = cdlm.footprint_bytes
    # N and N - 1 sequences have the same dictionary-mean chunk count here (N % 8 != 1)
    per_seq = fp(L4 // 4, N, M, gamma, x0) - fp(L4 // 4, N - 1, M, gamma, x0)
    fixed = fp(L4 // 4, N, M, gamma, x0) - per_seq * N
    return int(fixed + per_seq * (-(-N // chunks) + 2) + 64 * 1024)


def _same(a, b, tag):
    for f in FIELDS:
        np.testing.assert_array_equal(a.get(f), b.get(f), err_msg="%s %s" % (tag, f))


@pytest.mark.parametrize("chunks,x0", [(2, True), (8, False)])
def test_streamed_equals_resident_bit_for_bit(chunks, x0):
    N, L, M, fl = 300, 100, 12, [8, 12, 16] * 4
    s, D0, X0, _ = _synth(91 + chunks, N, L, M, fl)
    kw = dict(D0=D0, X0=X0 if x0 else None, seed=1234 + chunks)
    gamma = max(fl)
    bud = _budget(4 * L, N, M, gamma, x0, chunks)
    flush("streamed N=%d L=%d M=%d window budget %d bytes (footprint %d), X0 %s" % (
        N, L, M, bud, cdlm.footprint_bytes(L, N, M, gamma, x0), x0))
    r = cdlm.CDL(s, fl, 1.0, 100.0, 0.05, **kw)
    g = cdlm.CDL(s, fl, 1.0, 100.0, 0.05, stream_budget=bud, **kw)
    assert r.chunks == 1
    flush("  chunks %d, device bytes %d (resident %d)" % (g.chunks, g.device_bytes, r.device_bytes))
    assert g.chunks >= chunks
    assert g.device_bytes < r.device_bytes
    _same(g, r, "after create")
    # single steps
    for _ in range(2):
        g.code_update(4); r.code_update(4)
        assert g.update_best() == r.update_best()
        g.dict_update(2); r.dict_update(2)
        assert g.objective() == r.objective()
    _same(g, r, "after steps")
    # the schedule of cdl_optimize, then the threshold
    tg = g.optimize(4, 6, 2)
    tr = r.optimize(4, 6, 2)
    np.testing.assert_array_equal(tg, tr)
    assert g.code_thresh(1.0) == r.code_thresh(1.0)
    assert g.obj_best() == r.obj_best()
    _same(g, r, "after optimize")
    # reset replays the initialisation from the same inputs
    g.reset(); r.reset()
    np.testing.assert_array_equal(g.optimize(2, 3, 1), r.optimize(2, 3, 1))
    _same(g, r, "after reset")
    with pytest.raises(cdlm.CdlmError):
        cdlm.Batch([g])
    g.close(); r.close()


def test_streamed_one_window_when_two_do_not_fit():
    """A budget that holds one window of a single sequence but not two: the problem streams
    through one window (its chunks one after another, no copy overlap), one sequence per chunk,
    and still gives the resident results bit for bit."""
    N, L, M, fl = 64, 100, 16, [8] * 16
    s, D0, X0, _ = _synth(17, N, L, M, fl)
    fp = cdlm.footprint_bytes
    per_seq = fp(L, N, M, 8, False) - fp(L, N - 1, M, 8, False)   # ~0.1 MB; 256-byte alignment aside
    fixed = fp(L, N, M, 8, False) - per_seq * N
    bud = int(fixed + 1.5 * per_seq)
    r = cdlm.CDL(s, fl, 1.0, 100.0, 0.05, D0=D0, seed=5)
    g = cdlm.CDL(s, fl, 1.0, 100.0, 0.05, D0=D0, seed=5, stream_budget=bud)
    flush("one-window stream: budget %d (one sequence %d), chunks %d, device bytes %d" % (
        bud, per_seq, g.chunks, g.device_bytes))
    assert g.chunks >= N // 2 and g.device_bytes <= bud < fixed + 2 * per_seq
    np.testing.assert_array_equal(g.optimize(2, 3, 1), r.optimize(2, 3, 1))
    assert g.code_thresh(1.0) == r.code_thresh(1.0)
    _same(g, r, "one window")
    g.close(); r.close()


def test_budget_that_holds_the_problem_stays_resident():
    N, L, M, fl = 40, 60, 4, [8] * 4
    s, D0, X0, _ = _synth(5, N, L, M, fl)
    g = cdlm.CDL(s, fl, 1.0, 100.0, 0.05, D0=D0, stream_budget=cdlm.footprint_bytes(L, N, M, 8, False) + 4096)
    assert g.chunks == 1
    g.close()


def test_budget_below_the_smallest_window_is_refused():
    N, L, M, fl = 64, 60, 4, [8] * 4
    s, D0, _, _ = _synth(6, N, L, M, fl)
    with pytest.raises(cdlm.CdlmOOM):
        cdlm.CDL(s, fl, 1.0, 100.0, 0.05, D0=D0, stream_budget=4096)


def test_server_streams_a_job_beyond_its_budget():
    """A server whose budget is half of one job's footprint (261 MB): that job is streamed, alone,
    in four chunks; a small job submitted beside it is resident; both give what they give
    trained alone."""
    big = (3000, 100, 8, [12] * 8, 11)
    small = (50, 80, 4, [8] * 4, 12)
    alone, data = [], []
    for N, L, M, fl, seed in (big, small):
        s, D0, _, _ = _synth(seed, N, L, M, fl)
        data.append((s, fl, D0, seed))
        g = cdlm.CDL(s, fl, 1.0, 100.0, 0.05, D0=D0, seed=seed)
        tr = g.optimize(3, 5, 1)
        th, cnt = g.code_thresh(1.0)
        alone.append((tr, th, cnt, g.obj_best(), g.get("Xbest_slice")))
        g.close()
    budget = cdlm.footprint_bytes(big[1], big[0], big[2], 12, False) // 2
    flush("server budget %d bytes, big job footprint %d" % (budget, cdlm.footprint_bytes(big[1], big[0], big[2], 12,
                                                                                         False)))
    srv = cdlm.Server(mem_budget=budget, csc_iter=5, cdl_iter=1)
    ids = [srv.submit(s, fl, 1.0, 100.0, 0.05, D0=D0, seed=seed, iter_num=3, keep_codes=True)
           for s, fl, D0, seed in data]
    for i, jid in enumerate(ids):
        assert srv.wait(jid, timeout_s=120) == cdlm.serve.DONE
        res = srv.result(jid, codes=True)
        np.testing.assert_array_equal(res["trace"], alone[i][0])
        assert res["thresh"] == alone[i][1] and res["count"] == alone[i][2] and res["obj_best"] == alone[i][3]
        np.testing.assert_array_equal(res["codes"], alone[i][4])
        flush("  job %d: latency %.2f s" % (i, res["latency_s"]))
    st = srv.stats()
    # what stays held is the small job's arena, kept for reuse
    assert st["completed"] == 2 and st["bytes_in_use"] <= budget
    srv.close()
