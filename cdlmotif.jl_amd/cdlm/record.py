"""Host mirror of the activation record and topology (the consumer of X_best).

Reference (CDLmotif.jl):
  activation_record, activations_topology   src/SETUP/SEARCH_setup.jl:6-90
  record_training_result!                   src/SEARCH/store.jl:1-40
  topology_at_nth_pt!, get_spatial_info!    src/SEARCH/store.jl:64-144

Array conventions as elsewhere: a Julia column-major (C, K, N) array is numpy (N, K, C);
the topology counts (K, K, d) come back as numpy (d, K, K) indexed [d-1, fc-1, fr-1].
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from ._lib import check, load
from .device import Context, default_context


class ActivationRecord:
    """activation_record{T,S}(N, C, M) on the device (C = L), for up to max_trainings trainings."""

    def __init__(self, N: int, L: int, M: int, max_trainings: int = 2, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        h = C.c_void_p()
        check(load().cdlm_record_create(self.ctx.h, N, L, M, max_trainings, C.byref(h)), self.ctx.h)
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            load().cdlm_record_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, c, thresh: float, filter_spec: Sequence[int], cover_starts: Optional[np.ndarray] = None) -> int:
        """record_training_result! for problem c (its X_best); returns the filters added."""
        fs = np.ascontiguousarray(filter_spec, np.int32)
        cs = None if cover_starts is None else np.ascontiguousarray(cover_starts, np.int32)
        n = C.c_int32()
        check(load().cdlm_record_add(self.h, c.h, float(thresh), fs.ctypes.data, None if cs is None else cs.ctypes.data,
                                     C.byref(n)), self.ctx.h)
        return n.value

    def dims(self):
        v = [C.c_int32() for _ in range(4)]
        check(load().cdlm_record_dims(self.h, *[C.byref(x) for x in v]), self.ctx.h)
        return tuple(x.value for x in v)   # K, T, C, N

    def filters(self):
        """(uinds 1-based, training_num, filter_lengths), K each."""
        K = self.dims()[0]
        u, t, f = (np.zeros(K, np.int32) for _ in range(3))
        check(load().cdlm_record_filters(self.h, u.ctypes.data, t.ctypes.data, f.ctypes.data), self.ctx.h)
        return u, t, f

    def start_positions(self, n0: int = 0, count: Optional[int] = None) -> np.ndarray:
        """ar.start_positions[n0+1 : n0+count] as numpy (count, K, C) uint8."""
        K, _, Cc, N = self.dims()
        count = N - n0 if count is None else count
        out = np.zeros((count, K, Cc), np.uint8)
        check(load().cdlm_record_start_positions(self.h, n0, count, out.ctypes.data), self.ctx.h)
        return out

    def pair_pfms(self, data: np.ndarray, pairs: np.ndarray, max_span_len: int):
        """cdlm_record_pair_pfms: (span_len (P,), pfms (P, 4, max_span_len))."""
        pr = np.asfortranarray(np.asarray(pairs, np.int32).reshape(-1, 3))      # (P, 3) column-major
        P = pr.shape[0]
        dt = np.ascontiguousarray(data, np.float32)                               # (N, 4C) = (4C, N) column-major
        span = np.zeros(P, np.int32)
        pfm = np.zeros((max_span_len, 4, P), np.float32)                          # column-major (P, 4, S)
        check(load().cdlm_record_pair_pfms(self.h, dt.ctypes.data, P, pr.ctypes.data, max_span_len, span.ctypes.data,
                                           pfm.ctypes.data), self.ctx.h)
        return span, np.ascontiguousarray(pfm.transpose(2, 1, 0))

    def topology(self):
        """get_spatial_info!(ar): (overlap, non_overlap, next_to) as int32 numpy
        (max_fil_len, K, K), (apart, K, K), (K, K)."""
        K, mf, ap = C.c_int32(), C.c_int32(), C.c_int32()
        check(load().cdlm_record_topology_dims(self.h, C.byref(K), C.byref(mf), C.byref(ap)), self.ctx.h)
        K, mf, ap = K.value, mf.value, ap.value
        over = np.zeros((mf, K, K), np.int32)
        non = np.zeros((max(ap, 0), K, K), np.int32)
        nxt = np.zeros((K, K), np.int32)
        check(load().cdlm_record_topology(self.h, over.ctypes.data, non.ctypes.data, nxt.ctypes.data), self.ctx.h)
        return over, non, nxt


def overlap_pairs(overlap: np.ndarray, count_threshold: int) -> np.ndarray:
    """findall(at.overlap .> count_threshold) (extractions.jl:27-29) as (P, 3) 1-based
    (fil1, fil2, overlap) in the reference's column-major order; overlap is numpy (d, K, K)."""
    idx = np.argwhere(np.asarray(overlap) > count_threshold)      # rows (d, fc, fr), C order = Julia order
    return np.stack([idx[:, 2] + 1, idx[:, 1] + 1, idx[:, 0] + 1], axis=1).astype(np.int32) if len(idx) else \
        np.zeros((0, 3), np.int32)


def get_pair_pfms(ar: "ActivationRecord", overlap: np.ndarray, count_threshold: int, data: np.ndarray,
                  filter_size: int):
    """get_pair_pfms!(at, ar, g) up to the PFMs (extractions.jl:18-100): returns the pairs, the
    span lengths and the list of PFMs (4, span) of the pairs whose overlap count exceeds
    count_threshold.  data is the (N, 4L) one-hot matrix; max span = 2 filter_size - 1."""
    pairs = overlap_pairs(overlap, count_threshold)
    span, pfm = ar.pair_pfms(data, pairs, 2 * filter_size - 1)
    return pairs, span, [pfm[i, :, :span[i]] for i in range(len(pairs))]


def record_training_result(c, thresh, filter_spec, cover_starts, ar: ActivationRecord, train_num: int) -> int:
    """record_training_result!(c, thresh, filter_spec, cover_starts, ar, train_num) (store.jl:1-40);
    trainings are recorded in call order, so train_num must be the next one."""
    if train_num != ar.dims()[1] + 1:
        raise ValueError("train_num %d out of order (next is %d)" % (train_num, ar.dims()[1] + 1))
    return ar.add(c, thresh, filter_spec, cover_starts)


def get_spatial_info(ar: ActivationRecord):
    """get_spatial_info!(ar) (store.jl:121-144)."""
    return ar.topology()


def topology(start_pos, train_num, fil_lens, cover_starts, max_fil_len: int, apart: int, ctx: Optional[Context] = None):
    """topology_at_nth_pt! over all sequences for host arrays: start_pos numpy (N, K, C) uint8,
    train_num (K,) 1-based, fil_lens (K,), cover_starts numpy (T, N).  Returns int32 numpy
    (max_fil_len, K, K), (apart, K, K), (K, K)."""
    ctx = ctx or default_context()
    sp = np.ascontiguousarray(start_pos, np.uint8)
    N, K, Cc = sp.shape
    tn = np.ascontiguousarray(train_num, np.int32)
    fl = np.ascontiguousarray(fil_lens, np.int32)
    cs = np.ascontiguousarray(cover_starts, np.int32)
    T = cs.shape[0]
    over = np.zeros((max_fil_len, K, K), np.int32)
    non = np.zeros((max(apart, 0), K, K), np.int32)
    nxt = np.zeros((K, K), np.int32)
    check(load().cdlm_topology(ctx.h, sp.ctypes.data, tn.ctypes.data, fl.ctypes.data, cs.ctypes.data, Cc, K, N, T,
                               max_fil_len, apart, over.ctypes.data, non.ctypes.data, nxt.ctypes.data), ctx.h)
    return over, non, nxt
