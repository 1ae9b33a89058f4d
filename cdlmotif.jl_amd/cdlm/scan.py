"""Host mirror of the reference's PWM-scan entry points over libcdlm.

Reference (CDLmotif.jl):
  pfm2pwm          src/SEARCH/helpers.jl:12-13
  pwm_complement   src/SEARCH/helpers.jl:16
  best_score       src/SEARCH/PWM_Touzet.jl:9
  scan_w_gpu!      src/SEARCH/scan.jl:176-221 (+ greedy_search! :223-239)
  overlapping_scan! src/SEARCH/scan.jl:89-121 (+ scan_n :10-21)

PWMs are (4, m) float32 matrices (rows A, C, G, T); sequences are (N, L) uint8 base codes
0..3.  Positions returned by the scans are 1-based like the reference's Dicts.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence

import numpy as np

from ._lib import CdlmOOM, check, load
from .device import Context, default_context


def pfm2pwm(pfm: np.ndarray, bg=0.25) -> np.ndarray:
    """log2.(pfm ./ bg)  (helpers.jl:12-13); bg a scalar or a length-4 vector (per row)."""
    pfm = np.asarray(pfm, np.float64)
    bgv = np.asarray(bg, np.float64)
    if bgv.ndim == 1:
        bgv = bgv[:, None]
    return np.log2(pfm / bgv).astype(np.float32)


def pwm_complement(pwm: np.ndarray) -> np.ndarray:
    """reverse(pwm) in both dimensions (helpers.jl:16)."""
    return np.ascontiguousarray(np.asarray(pwm)[::-1, ::-1])


def best_score(pwm: np.ndarray) -> np.float32:
    """Sum over columns of the column maximum, sequential FP32 (PWM_Touzet.jl:9)."""
    acc = np.float32(0)
    for j in range(pwm.shape[1]):
        acc = np.float32(acc + np.float32(np.max(pwm[:, j])))
    return acc


def pack_pwms(pwms: Sequence[np.ndarray]):
    """(K, 4, maxlen) zero-padded float32 array and lengths, as scan_w_gpu! packs them
    (scan.jl:185-186)."""
    K = len(pwms)
    lens = np.array([p.shape[1] for p in pwms], np.int32)
    maxlen = int(lens.max())
    out = np.zeros((K, 4, maxlen), np.float32)
    for k, p in enumerate(pwms):
        out[k, :, :p.shape[1]] = p
    return out, lens


def _prep(bases, packed, lens, thresh):
    bases = np.ascontiguousarray(bases, np.uint8)
    packed = np.ascontiguousarray(packed, np.float32)
    lens = np.ascontiguousarray(lens, np.int32)
    thresh = np.ascontiguousarray(thresh, np.float32)
    N, L = bases.shape
    K, four, maxlen = packed.shape
    if four != 4 or lens.shape != (K,) or thresh.shape != (K,):
        raise ValueError("pwms must be (K,4,maxlen) with K lengths and K thresholds")
    return bases, packed, lens, thresh, N, L, K, maxlen


def onehot_to_bases(onehot: np.ndarray, ctx: Optional[Context] = None) -> np.ndarray:
    """(N, 4L) one-hot rows (the reference's (4L, N) data matrix) -> (N, L) base codes;
    a group with no 1 becomes 4."""
    ctx = ctx or default_context()
    oh = np.ascontiguousarray(onehot, np.float32)
    N, L4 = oh.shape
    out = np.empty((N, L4 // 4), np.uint8)
    check(load().cdlm_onehot_to_bases(ctx.h, oh.ctypes.data, L4, N, out.ctypes.data), ctx.h)
    return out


def scan_hits(bases, packed, lens, thresh, both_strands=True, cap: Optional[int] = None,
              ctx: Optional[Context] = None, sort=True):
    """Every window with score > thresh[k] (position-major FP32 score of greedy_search!),
    both strands.  Returns int64 array (n_hits, 4) of (k, n, pos0, strand), sorted by
    (k, n, pos, strand) when sort=True."""
    ctx = ctx or default_context()
    bases, packed, lens, thresh, N, L, K, maxlen = _prep(bases, packed, lens, thresh)
    cap = int(cap if cap is not None else max(1024, N * 8))
    while True:
        hk = np.empty(cap, np.int32); hn = np.empty(cap, np.int64); hp = np.empty(cap, np.int32)
        hs = np.empty(cap, np.int8)
        nh = C.c_int64()
        st = load().cdlm_pwm_hits(ctx.h, bases.ctypes.data, L, N, packed.ctypes.data, lens.ctypes.data,
                                  thresh.ctypes.data, K, maxlen, 1 if both_strands else 0, hk.ctypes.data,
                                  hn.ctypes.data, hp.ctypes.data, hs.ctypes.data, cap, C.byref(nh))
        if st == 3 and nh.value > cap:        # too many hits for the buffer: retry once sized
            cap = int(nh.value)
            continue
        check(st, ctx.h)
        n = nh.value
        out = np.stack([hk[:n].astype(np.int64), hn[:n], hp[:n].astype(np.int64), hs[:n].astype(np.int64)], 1)
        if sort and n:
            out = out[np.lexsort((out[:, 3], out[:, 2], out[:, 1], out[:, 0]))]
        return out


def scan_w_gpu(bases, packed, lens, thresh, ctx: Optional[Context] = None):
    """scan_w_gpu! positions: (K, N) int32, 1-based best forward position or 0."""
    ctx = ctx or default_context()
    bases, packed, lens, thresh, N, L, K, maxlen = _prep(bases, packed, lens, thresh)
    out = np.empty((K, N), np.int32)
    check(load().cdlm_pwm_argmax(ctx.h, bases.ctypes.data, L, N, packed.ctypes.data, lens.ctypes.data,
                                 thresh.ctypes.data, K, maxlen, out.ctypes.data), ctx.h)
    return out


def scan_w_gpu_dicts(bases, pwms: List[np.ndarray], thresh, ctx: Optional[Context] = None):
    """The Dict form of scan_w_gpu! (scan.jl:200-205, scan_bg=true): one {n: pos} per PWM,
    1-based sequence index and position."""
    packed, lens = pack_pwms(pwms)
    pos = scan_w_gpu(bases, packed, lens, thresh, ctx)
    return [{int(n) + 1: int(p) for n, p in enumerate(row) if p > 0} for row in pos]


def get_thresh(pwms, pval: float, eps: float = 0.1, ctx: Optional[Context] = None) -> np.ndarray:
    """get_thresh(pwms, pval, Float32) (helpers.jl:99) on the device: Float32 of
    pvalue2score(pwm, pval) per PWM (PWM_Touzet.jl:199-217).  pwms: a list of (4, m) arrays or
    a packed (K, 4, maxlen) array with lens.  pval is taken as given; the reference passes a
    Float32 (wrap.jl:125), so pass float(np.float32(p)) for its exact value."""
    ctx = ctx or default_context()
    if isinstance(pwms, tuple):
        packed, lens = pwms
    else:
        packed, lens = pack_pwms(pwms)
    packed = np.ascontiguousarray(packed, np.float32)
    lens = np.ascontiguousarray(lens, np.int32)
    K, maxlen = packed.shape[0], packed.shape[2]
    out = np.empty(K, np.float32)
    check(load().cdlm_pwm_pvalue_thresh(ctx.h, packed.ctypes.data, lens.ctypes.data, K, maxlen, float(pval),
                                        float(eps), out.ctypes.data), ctx.h)
    return out


def pvalue_last_ms(ctx: Optional[Context] = None) -> float:
    """Device time of the last get_thresh call's kernels."""
    ctx = ctx or default_context()
    ms = C.c_double()
    check(load().cdlm_pwm_pvalue_last_ms(ctx.h, C.byref(ms)), ctx.h)
    return ms.value


def overlapping_scan(bases, packed, lens, thresh, ctx: Optional[Context] = None):
    """overlapping_scan! best hits: (positions (K,N) 1-based or 0, use_comp (K,N) bool,
    scores (K,N) float32)."""
    ctx = ctx or default_context()
    bases, packed, lens, thresh, N, L, K, maxlen = _prep(bases, packed, lens, thresh)
    pos = np.empty((K, N), np.int32); comp = np.empty((K, N), np.int8); sc = np.empty((K, N), np.float32)
    check(load().cdlm_pwm_best(ctx.h, bases.ctypes.data, L, N, packed.ctypes.data, lens.ctypes.data,
                               thresh.ctypes.data, K, maxlen, pos.ctypes.data, comp.ctypes.data, sc.ctypes.data),
          ctx.h)
    return pos, comp.astype(bool), sc


__all__ = ["pfm2pwm", "pwm_complement", "best_score", "pack_pwms", "scan_hits", "scan_w_gpu", "scan_w_gpu_dicts",
           "overlapping_scan", "onehot_to_bases", "CdlmOOM"]
