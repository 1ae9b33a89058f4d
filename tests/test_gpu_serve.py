"""Served CDL: jobs admitted by device footprint while others run, concurrently at their own
pace or advanced together one outer iteration per tick (lockstep).  Each job's result
(objective trace, code threshold, positive count, X_best slice) must be bit-identical to the
same training run alone through the problem API, whatever else shares the device (the
reference trains one dataset per cdl_optimize call, src/WRAPPER/wrap.jl:4-12)."""
import threading
import time

import numpy as np
import pytest

from gpu_util import flush

pytestmark = pytest.mark.gpu
cdlm = pytest.importorskip("cdlm")

RHO, SIGMA, LAM = 1.0, 100.0, 0.05
CSC, CDL_IT = 4, 1
# (seed, N, L, filter lengths, outer iterations)
JOBS = [
    (201, 50, 40, [12] * 6, 3),
    (202, 30, 25, [20] * 4, 5),
    (203, 70, 60, [8, 16, 16], 2),
    (204, 20, 12, [4, 4], 4),
    (205, 40, 80, [24] * 5, 3),
    (206, 15, 270, [8, 8, 8], 2),
]


def _data(seed, N, L):
    from cdlm import synth
    bases, _ = synth.planted_bases(seed, N, L, [6, 9], 0.3)
    return synth.onehot(bases)


def _alone(job):
    sd, N, L, fl, it = job
    c = cdlm.CDL(_data(sd, N, L), fl, RHO, SIGMA, LAM, seed=sd)
    tr = c.optimize(it, CSC, CDL_IT)
    th, cnt = c.code_thresh(1.0)
    codes = c.get("Xbest_slice")
    c.close()
    return tr, th, cnt, codes


def _same(r, ref, tag):
    tr, th, cnt, codes = ref
    assert np.array_equal(r["trace"], tr), (tag, r["trace"], tr)
    assert r["thresh"] == th and r["count"] == cnt, tag
    assert np.array_equal(r["codes"], codes), tag


@pytest.mark.parametrize("lockstep", [False, True])
def test_served_jobs_equal_alone_in_the_mix(lockstep):
    refs = [_alone(j) for j in JOBS]
    # alone through the server
    srv = cdlm.Server(csc_iter=CSC, cdl_iter=CDL_IT, lockstep=lockstep)
    sd, N, L, fl, it = JOBS[0]
    jid = srv.submit(_data(sd, N, L), fl, RHO, SIGMA, LAM, it, seed=sd, keep_codes=True)
    assert srv.wait(jid, 120) == cdlm.serve.DONE
    _same(srv.result(jid, codes=True), refs[0], "served alone")
    # in the mix: staggered arrivals while earlier jobs run
    ids = []
    for k, (sd, N, L, fl, it) in enumerate(JOBS):
        ids.append(srv.submit(_data(sd, N, L), fl, RHO, SIGMA, LAM, it, seed=sd, keep_codes=True))
        time.sleep(0.002 * k)
    for k, jid in enumerate(ids):
        assert srv.wait(jid, 120) == cdlm.serve.DONE, k
        r = srv.result(jid, codes=True)
        flush("job %d: latency %.4f s (queued %.4f s), thresh %.6g" % (k, r["latency_s"], r["queue_s"], r["thresh"]))
        _same(r, refs[k], k)
        assert r["latency_s"] >= r["queue_s"] >= 0
    st = srv.stats()
    flush("stats %s" % st)
    assert st["completed"] == len(JOBS) + 1 and st["active"] == 0 and st["queued"] == 0
    assert 0 < st["p50_s"] <= st["p99_s"]
    srv.close()


@pytest.mark.parametrize("lockstep", [False, True])
def test_admission_by_footprint(lockstep):
    """A budget that holds only the largest job: jobs queue, run as memory frees up, never
    exceed the budget, and still give their alone results; a job whose smallest streamed window
    (77 MB) is larger than the budget is refused at submission."""
    sizes = [cdlm.footprint_bytes(L, N, len(fl), max(fl)) for _, N, L, fl, _ in JOBS]
    budget = max(sizes) + min(sizes) // 2
    srv = cdlm.Server(mem_budget=budget, csc_iter=CSC, cdl_iter=CDL_IT, lockstep=lockstep)
    seen = []
    stop = threading.Event()

    def watch():
        while not stop.is_set():
            seen.append(srv.stats()["bytes_in_use"])
            time.sleep(0.0005)

    w = threading.Thread(target=watch)
    w.start()
    try:
        ids = [srv.submit(_data(sd, N, L), fl, RHO, SIGMA, LAM, it, seed=sd, keep_codes=True)
               for sd, N, L, fl, it in JOBS]
        for k, jid in enumerate(ids):
            assert srv.wait(jid, 120) == cdlm.serve.DONE, k
            _same(srv.result(jid, codes=True), _alone(JOBS[k]), ("budget", k))
    finally:
        stop.set()
        w.join()
    assert max(seen) <= budget
    big = _data(1, 400, 500)
    with pytest.raises(cdlm.CdlmOOM):
        srv.submit(big, [24] * 64, RHO, SIGMA, LAM, 2)
    srv.close()


@pytest.mark.parametrize("lockstep", [False, True])
def test_destroy_with_pending_jobs_fails_them_cleanly(lockstep):
    srv = cdlm.Server(csc_iter=CSC, cdl_iter=CDL_IT, max_active=1, lockstep=lockstep)
    ids = [srv.submit(_data(sd, N, L), fl, RHO, SIGMA, LAM, 50, seed=sd) for sd, N, L, fl, _ in JOBS[:3]]
    time.sleep(0.05)
    srv.close()   # must not hang; queued and resident jobs end as failed
    assert ids


@pytest.mark.parametrize("lockstep", [False, True])
@pytest.mark.parametrize("policy", ["fifo", "shortest"])
def test_admission_order_policy(policy, lockstep):
    """One resident job at a time: a long job runs while a big and then a small job queue.
    FIFO admits the big one first; shortest-work-first admits the small one first."""
    srv = cdlm.Server(csc_iter=CSC, cdl_iter=CDL_IT, max_active=1, policy=policy, lockstep=lockstep)
    a = srv.submit(_data(301, 60, 200), [16] * 8, RHO, SIGMA, LAM, 400, seed=1)
    time.sleep(0.01)
    db, ds = _data(302, 60, 200), _data(303, 10, 20)
    t0 = time.perf_counter()
    big = srv.submit(db, [16] * 8, RHO, SIGMA, LAM, 20, seed=2)
    small = srv.submit(ds, [8] * 2, RHO, SIGMA, LAM, 2, seed=3)
    gap = time.perf_counter() - t0          # >= small's submit time - big's
    assert srv.poll(a) == cdlm.serve.RUNNING
    for j in (a, big, small):
        assert srv.wait(j, 120) == cdlm.serve.DONE
    rb, rs = srv.result(big), srv.result(small)
    srv.close()
    # one job resident at a time, so the order shows in the server's own clock: a job that ran
    # first was done before the other was admitted (latency and queue time both count from the
    # job's own submission, and big was submitted at most `gap` before small)
    if policy == "fifo":
        assert rb["latency_s"] <= rs["queue_s"] + gap, (rb, rs, gap)
    else:
        assert rs["latency_s"] <= rb["queue_s"], (rb, rs)
