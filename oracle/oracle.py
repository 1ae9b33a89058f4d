"""ctypes front end of the CPU oracle (liboracle.so built from oracle.c).

TEST INFRASTRUCTURE ONLY.  tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` use it as the checker of the HIP library; the product never imports
it.  The restated reference functions and their file:line sources are listed in the
header of oracle.c and oracle_impl.h.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

# field ids of orc_cdl_get (oracle_impl.h)
FIELDS = {"X": 0, "DX": 1, "d": 2, "G1": 3, "G2": 4, "G3": 5, "Xbest": 6, "Dbest": 7, "Y1": 8, "Y2": 9,
          "Y3": 10, "y1": 11, "y2": 12, "y3": 13, "g1": 14, "g2": 15, "g3": 16, "Td": 17}


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "liboracle.so"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    fp, dp, ip = C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_int)
    for sfx in ("_f32", "_f64"):
        f = getattr(L, "orc_cdl_create" + sfx)
        f.restype = C.c_void_p
        f.argtypes = [C.c_int, C.c_int, C.c_int, ip, C.c_double, C.c_double, C.c_double, fp, fp, fp]
        getattr(L, "orc_cdl_destroy" + sfx).argtypes = [C.c_void_p]
        for nm in ("orc_cdl_code_update", "orc_cdl_dict_update"):
            g = getattr(L, nm + sfx)
            g.argtypes = [C.c_void_p, C.c_int, C.c_double]
        g = getattr(L, "orc_cdl_update_best" + sfx)
        g.argtypes = [C.c_void_p, dp]
        g.restype = C.c_int
        for nm in ("orc_cdl_objective", "orc_cdl_obj_best"):
            g = getattr(L, nm + sfx)
            g.argtypes = [C.c_void_p]
            g.restype = C.c_double
        g = getattr(L, "orc_cdl_code_thresh" + sfx)
        g.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.c_long)]
        g.restype = C.c_double
        g = getattr(L, "orc_cdl_get" + sfx)
        g.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        g.restype = C.c_long
        for nm in ("orc_box", "orc_prox_simplex"):
            g = getattr(L, nm + sfx)
            g.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        g = getattr(L, "orc_fft" + sfx)
        g.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
    u8 = C.POINTER(C.c_uint8)
    L.orc_pwm_hits.restype = C.c_long
    L.orc_pwm_hits.argtypes = [u8, C.c_int, C.c_long, fp, ip, fp, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                               C.c_void_p, C.c_void_p, C.c_void_p, C.c_long]
    L.orc_pwm_gpu_argmax.argtypes = [u8, C.c_int, C.c_long, fp, ip, fp, C.c_int, C.c_int, C.c_void_p]
    L.orc_pwm_best.argtypes = [u8, C.c_int, C.c_long, fp, ip, fp, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                               C.c_void_p]
    L.orc_best_score.restype = C.c_float
    L.orc_best_score.argtypes = [fp, C.c_int, C.c_int]
    L.orc_topology.argtypes = [u8, ip, ip, ip, C.c_int, C.c_int, C.c_long, C.c_int, C.c_int, C.c_int, ip, ip, ip]
    L.orc_pair_pfms.argtypes = [u8, fp, ip, ip, ip, fp, C.c_int, C.c_int, C.c_int, C.c_long, C.c_int, ip, C.c_int, ip, fp]


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a, t=C.c_float):
    return a.ctypes.data_as(C.POINTER(t))


class OracleCDL:
    """Reference-structured CDL state on the CPU (precision "f32" or "f64").

    Arrays use the reference's column-major shapes flattened with l fastest:
    s is (4L, N) -> numpy (N, 4L); D0 is (Lg, M) -> (M, Lg); X0 is (Lg, M, N) -> (N, M, Lg).
    """

    def __init__(self, s, D0, X0, fl, rho, sigma, lam, precision="f32"):
        self.p = precision
        self.sfx = "_f32" if precision == "f32" else "_f64"
        self.s, self.D0, self.X0 = _f32(s), _f32(D0), _f32(X0)
        self.N, self.L4 = self.s.shape
        self.M = len(fl)
        self.fl = np.ascontiguousarray(fl, dtype=np.int32)
        self.gamma = int(self.fl.max())
        self.Lg = self.L4 + 2 * self.gamma
        assert self.D0.shape == (self.M, self.Lg), (self.D0.shape, (self.M, self.Lg))
        assert self.X0.shape == (self.N, self.M, self.Lg)
        L = lib()
        self.h = getattr(L, "orc_cdl_create" + self.sfx)(
            self.L4, self.N, self.M, _ptr(self.fl, C.c_int), rho, sigma, lam, _ptr(self.s), _ptr(self.D0),
            _ptr(self.X0))

    def _fn(self, nm):
        return getattr(lib(), nm + self.sfx)

    def close(self):
        if self.h:
            self._fn("orc_cdl_destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def code_update(self, n, sherman=2.0):
        self._fn("orc_cdl_code_update")(self.h, n, sherman)

    def dict_update(self, n, sherman=2.0):
        self._fn("orc_cdl_dict_update")(self.h, n, sherman)

    def update_best(self):
        obj = C.c_double()
        imp = self._fn("orc_cdl_update_best")(self.h, C.byref(obj))
        return obj.value, bool(imp)

    def objective(self):
        return self._fn("orc_cdl_objective")(self.h)

    def obj_best(self):
        return self._fn("orc_cdl_obj_best")(self.h)

    def code_thresh(self, mult=1.0):
        cnt = C.c_long()
        v = self._fn("orc_cdl_code_thresh")(self.h, mult, C.byref(cnt))
        return v, cnt.value

    def get(self, name):
        fid = FIELDS[name]
        cnt = self._fn("orc_cdl_get")(self.h, fid, None)
        dt = np.float32 if self.p == "f32" else np.float64
        out = np.empty(cnt, dtype=dt)
        self._fn("orc_cdl_get")(self.h, fid, out.ctypes.data_as(C.c_void_p))
        Lg, M, N = self.Lg, self.M, self.N
        if cnt == Lg * M * N:
            return out.reshape(N, M, Lg)
        if cnt == Lg * N:
            return out.reshape(N, Lg)
        return out.reshape(M, Lg)


def fft(x, inverse=False, precision="f64"):
    x = np.asarray(x)
    n = x.shape[-1]
    dt = np.float32 if precision == "f32" else np.float64
    re = np.ascontiguousarray(x.real, dtype=dt)
    im = np.ascontiguousarray(x.imag, dtype=dt)
    ore, oim = np.empty_like(re), np.empty_like(im)
    f = getattr(lib(), "orc_fft" + ("_f32" if precision == "f32" else "_f64"))
    f(n, re.ctypes.data_as(C.c_void_p), im.ctypes.data_as(C.c_void_p), ore.ctypes.data_as(C.c_void_p),
      oim.ctypes.data_as(C.c_void_p), 1 if inverse else 0)
    return ore + 1j * oim


def _proj(name, v, P, precision):
    dt = np.float32 if precision == "f32" else np.float64
    v = np.ascontiguousarray(v, dtype=dt)
    P = np.ascontiguousarray(P, dtype=dt)
    M, Lg = v.shape
    out = np.empty_like(v)
    getattr(lib(), name + ("_f32" if precision == "f32" else "_f64"))(
        Lg, M, v.ctypes.data_as(C.c_void_p), P.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p))
    return out


def box(v, P, precision="f32"):
    """projection2box! on (M, Lg) arrays."""
    return _proj("orc_box", v, P, precision)


def prox_simplex(v, P, precision="f32"):
    """prox_simplex_zeros! on (M, Lg) arrays."""
    return _proj("orc_prox_simplex", v, P, precision)


def _pwm_args(seqs, pwms, lens, thresh):
    seqs = np.ascontiguousarray(seqs, dtype=np.uint8)
    pwms = _f32(pwms)
    lens = np.ascontiguousarray(lens, dtype=np.int32)
    thresh = _f32(thresh)
    N, L = seqs.shape
    K, four, maxlen = pwms.shape
    assert four == 4 and len(lens) == K and len(thresh) == K
    return seqs, pwms, lens, thresh, N, L, K, maxlen


def pwm_hits(seqs, pwms, lens, thresh, both_strands=True, cap=None):
    """All windows with score > thresh; returns structured (k, n, pos0, strand, score)."""
    seqs, pwms, lens, thresh, N, L, K, maxlen = _pwm_args(seqs, pwms, lens, thresh)
    args = (_ptr(seqs, C.c_uint8), L, N, _ptr(pwms), _ptr(lens, C.c_int), _ptr(thresh), K, maxlen,
            1 if both_strands else 0)
    if cap is None:
        cap = lib().orc_pwm_hits(*args, None, None, None, None, None, 0)
    k = np.empty(cap, np.int32); n = np.empty(cap, np.int64); p = np.empty(cap, np.int32)
    st = np.empty(cap, np.int8); sc = np.empty(cap, np.float32)
    cnt = lib().orc_pwm_hits(*args, k.ctypes.data_as(C.c_void_p), n.ctypes.data_as(C.c_void_p),
                             p.ctypes.data_as(C.c_void_p), st.ctypes.data_as(C.c_void_p),
                             sc.ctypes.data_as(C.c_void_p), cap)
    cnt = min(cnt, cap)
    return k[:cnt], n[:cnt], p[:cnt], st[:cnt], sc[:cnt]


def pwm_gpu_argmax(seqs, pwms, lens, thresh):
    seqs, pwms, lens, thresh, N, L, K, maxlen = _pwm_args(seqs, pwms, lens, thresh)
    out = np.empty((K, N), np.int32)
    lib().orc_pwm_gpu_argmax(_ptr(seqs, C.c_uint8), L, N, _ptr(pwms), _ptr(lens, C.c_int), _ptr(thresh), K, maxlen,
                             out.ctypes.data_as(C.c_void_p))
    return out


def pwm_best(seqs, pwms, lens, thresh):
    seqs, pwms, lens, thresh, N, L, K, maxlen = _pwm_args(seqs, pwms, lens, thresh)
    pos = np.empty((K, N), np.int32); comp = np.empty((K, N), np.int8); sc = np.empty((K, N), np.float32)
    lib().orc_pwm_best(_ptr(seqs, C.c_uint8), L, N, _ptr(pwms), _ptr(lens, C.c_int), _ptr(thresh), K, maxlen,
                       pos.ctypes.data_as(C.c_void_p), comp.ctypes.data_as(C.c_void_p), sc.ctypes.data_as(C.c_void_p))
    return pos, comp, sc


def best_score(pwm_km):
    """pwm_km: (4, m) float32."""
    a = _f32(pwm_km)
    return float(lib().orc_best_score(_ptr(a), a.shape[1], a.shape[1]))


def topology(start_pos, train_num, fil_lens, cover_starts, maxfl, apart):
    """topology_at_nth_pt! (src/SEARCH/store.jl:64-107) over all sequences.  start_pos is the
    reference's (C, K, N) Bool array given as numpy (N, K, C) uint8; train_num (K,) 1-based,
    fil_lens (K,), cover_starts the reference's (N, T) matrix given as numpy (T, N).  Returns
    (over, non_over, next_to) as numpy arrays (maxfl, K, K), (apart, K, K), (K, K) of int32
    (the reference's column-major (K, K, d) with axes reversed: [d, fc, fr])."""
    sp = np.ascontiguousarray(start_pos, np.uint8)
    N, K, Cc = sp.shape
    tn = np.ascontiguousarray(train_num, np.int32)
    fl = np.ascontiguousarray(fil_lens, np.int32)
    cs = np.ascontiguousarray(cover_starts, np.int32)
    T = cs.shape[0]
    over = np.zeros((maxfl, K, K), np.int32)
    non = np.zeros((max(apart, 1), K, K), np.int32)
    nxt = np.zeros((K, K), np.int32)
    u8 = C.POINTER(C.c_uint8)
    ip = C.POINTER(C.c_int)
    lib().orc_topology(sp.ctypes.data_as(u8), tn.ctypes.data_as(ip), fl.ctypes.data_as(ip), cs.ctypes.data_as(ip),
                       Cc, K, N, T, maxfl, apart, over.ctypes.data_as(ip), non.ctypes.data_as(ip), nxt.ctypes.data_as(ip))
    return over, non[:apart], nxt


def pair_pfms(start_pos, codes, uinds, train_num, fil_lens, data, pairs, max_span_len):
    """get_pair_pfms! up to the PFMs (src/SEARCH/extractions.jl:18-100).  start_pos numpy
    (N, K, C) uint8; codes numpy (T, N, M, C) float32 (the reference's (C, M, N, T) reversed);
    uinds, train_num, fil_lens (K,) (1-based indices); data numpy (N, 4C) float32; pairs
    (P, 3) 1-based (fil1, fil2, overlap).  Returns (span_len (P,), pfms (P, 4, S))."""
    sp = np.ascontiguousarray(start_pos, np.uint8)
    N, K, Cc = sp.shape
    cd = np.ascontiguousarray(codes, np.float32)
    M = cd.shape[2]
    pr = np.ascontiguousarray(pairs, np.int32).reshape(-1, 3)
    P = pr.shape[0]
    dt = np.ascontiguousarray(data, np.float32)
    span = np.zeros(P, np.int32)
    pfm = np.zeros((P, 4, max_span_len), np.float32)
    ip = C.POINTER(C.c_int)
    fp = C.POINTER(C.c_float)
    u = np.ascontiguousarray(uinds, np.int32); t = np.ascontiguousarray(train_num, np.int32)
    f = np.ascontiguousarray(fil_lens, np.int32)
    lib().orc_pair_pfms(sp.ctypes.data_as(C.POINTER(C.c_uint8)), cd.ctypes.data_as(fp), u.ctypes.data_as(ip),
                        t.ctypes.data_as(ip), f.ctypes.data_as(ip), dt.ctypes.data_as(fp), Cc, K, M, N, P,
                        pr.ctypes.data_as(ip), max_span_len, span.ctypes.data_as(ip), pfm.ctypes.data_as(fp))
    return span, pfm
