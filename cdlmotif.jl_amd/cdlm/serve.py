"""Host mirror of the library's continuous-batching server (cdlm_server_*).

The reference runs one motif-discovery dataset at a time (find_motifs_fasta ->
try_to_find_motif -> cdl_optimize, src/WRAPPER/wrap.jl:4-12,140-149).  A Server keeps
many CDL training jobs resident on one device: jobs are admitted by device footprint
as memory frees up and run concurrently, each at its own pace (or, with lockstep=True,
every tick advances all resident jobs by one outer iteration of the reference schedule in
one batch).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from ._lib import (CdlmError, check, cdlm_job_result, cdlm_job_spec, cdlm_params, cdlm_server_opts,
                   cdlm_server_stats_t, load, ERR_OOM, CdlmOOM)

QUEUED, RUNNING, DONE, FAILED = range(4)


class Server:
    def __init__(self, device: int = 0, mem_budget: int = 0, max_active: int = 0, csc_iter: int = 15,
                 cdl_iter: int = 1, graph: bool = True, fill_seqs: int = 0, policy: str = "fifo",
                 lockstep: bool = False):
        """fill_seqs: admit no job while the resident jobs hold this many sequences (0: twelve
        per CU, < 0: no limit besides memory).  policy: "fifo" (arrival order) or "shortest"
        (least work first); a job passed by 4 later arrivals is admitted first under either.
        lockstep: advance all resident jobs together, one batched outer iteration per tick
        (graph: replay the ticks from captured HIP graphs); else (default) up to four jobs, or
        max_active, run concurrently on their own streams, each at its own pace."""
        pol = {"fifo": 0, "shortest": 1}[policy]
        o = cdlm_server_opts(mem_budget, max_active, csc_iter, cdl_iter, 1 if graph else 0, fill_seqs, pol,
                             1 if lockstep else 0)
        h = C.c_void_p()
        st = load().cdlm_server_create(device, C.byref(o), C.byref(h))
        if st != 0:
            check(st)
        self.h = h
        self.csc_iter, self.cdl_iter = csc_iter, cdl_iter

    def _check(self, st):
        if st == 0:
            return
        m = load().cdlm_server_last_error(self.h)
        msg = m.decode() if m else ""
        if st == ERR_OOM:
            raise CdlmOOM(st, msg)
        raise CdlmError(st, msg)

    def close(self):
        if getattr(self, "h", None):
            load().cdlm_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, data: np.ndarray, filter_len: Sequence[int], rho: float, sigma: float, lam: float,
               iter_num: int, seed: int = 0, mult: float = 1.0, keep_codes: bool = False,
               D0: Optional[np.ndarray] = None, sherman_csc: float = 2.0, sherman_cdl: float = 2.0) -> int:
        """Queues one training (the data is copied); returns the job id."""
        data = np.ascontiguousarray(data, dtype=np.float32)
        N, L4 = data.shape
        fl = np.ascontiguousarray(filter_len, dtype=np.int32)
        d0 = None if D0 is None else np.ascontiguousarray(D0, dtype=np.float32)
        prm = cdlm_params(len(fl), fl.ctypes.data_as(C.POINTER(C.c_int32)), rho, sigma, lam, sherman_csc, sherman_cdl)
        spec = cdlm_job_spec(prm, data.ctypes.data, L4, N, None if d0 is None else d0.ctypes.data,
                             seed & 0xFFFFFFFFFFFFFFFF, iter_num, mult, 1 if keep_codes else 0)
        jid = C.c_uint64()
        self._check(load().cdlm_server_submit(self.h, C.byref(spec), C.byref(jid)))
        return jid.value

    def poll(self, jid: int) -> int:
        st = C.c_int32()
        self._check(load().cdlm_server_poll(self.h, jid, C.byref(st)))
        return st.value

    def wait(self, jid: int, timeout_s: Optional[float] = None) -> int:
        st = C.c_int32()
        ms = -1 if timeout_s is None else int(timeout_s * 1000)
        self._check(load().cdlm_server_wait(self.h, jid, ms, C.byref(st)))
        return st.value

    def result(self, jid: int, codes: bool = False) -> dict:
        r = cdlm_job_result()
        self._check(load().cdlm_server_result(self.h, jid, C.byref(r), None, None, 0))
        tr = np.zeros(max(r.iter_num, 1), np.float64)
        cd = np.empty((r.N, r.M, r.L), np.float32) if codes else None
        self._check(load().cdlm_server_result(self.h, jid, C.byref(r), tr.ctypes.data,
                                              None if cd is None else cd.ctypes.data, 0 if cd is None else cd.nbytes))
        out = {f: getattr(r, f) for f, _ in cdlm_job_result._fields_}
        out["trace"] = tr[:r.iter_num]
        if cd is not None:
            out["codes"] = cd
        return out

    def release(self, jid: int):
        self._check(load().cdlm_server_release(self.h, jid))

    def stats(self) -> dict:
        s = cdlm_server_stats_t()
        self._check(load().cdlm_server_stats(self.h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in cdlm_server_stats_t._fields_}
