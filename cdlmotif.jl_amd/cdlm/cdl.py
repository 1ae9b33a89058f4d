"""Host mirror of the reference's CDL entry points over libcdlm.

Reference (CDLmotif.jl):
  CDL_setup            src/SETUP/WRAPPER_setup.jl:1-13
  cdl / init_cdl       src/SETUP/CDL_setup.jl:70-283
  code_update!         src/CDL/train.jl:1-44
  dict_update!         src/CDL/train.jl:46-89
  update_best!         src/CDL/train.jl:91-98
  objective            src/SETUP/CDL_setup.jl:286
  cdl_optimize         src/WRAPPER/wrap.jl:1-32 (one training; the code threshold of :12)

Array conventions: the device keeps the reference's Julia column-major layouts; numpy
views use the same memory with reversed axes, so a Julia (Lg, M, N) array is a numpy
(N, M, Lg) array, and the data matrix (4L, N) is numpy (N, 4L).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from ._lib import ERR_ARG, CdlmError, check, cdlm_params, load
from .device import Context, default_context

FIELD_IDS = {"X": 0, "DX": 1, "d": 2, "G1": 3, "G2": 4, "G3": 5, "Xbest_slice": 6, "Dbest": 7, "Y1": 8, "y1": 11,
             "y2": 12, "y3": 13, "g1": 14, "g2": 15, "g3": 16, "Td": 17}


@dataclass
class CDL_setup:
    """Defaults of CDL_setup (src/SETUP/WRAPPER_setup.jl:1-13)."""
    M: int = 10
    filter_size: int = 4
    rho: float = 100.0
    sigma: float = 100.0
    lam: float = 0.01
    count_threshold: int = 50
    csc_iter_num: int = 15
    cdl_iter_num: int = 1
    iter_num: int = 10
    std_threshold_multiplier: float = 1.0
    num_train: int = 2
    filter_len: Optional[Sequence[int]] = field(default=None)   # per-filter override (flattened lengths)

    def filter_lens(self):
        if self.filter_len is not None:
            return [int(x) for x in self.filter_len]
        return [4 * self.filter_size] * self.M


def _f32c(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class CDL:
    """Device-resident state of one CDL problem (the reference's mutable struct `cdl`)."""

    def __init__(self, data: Optional[np.ndarray], filter_len: Sequence[int], rho: float, sigma: float, lam: float,
                 D0: Optional[np.ndarray] = None, X0: Optional[np.ndarray] = None, seed: int = 0,
                 sherman_csc: float = 2.0, sherman_cdl: float = 2.0, ctx: Optional[Context] = None,
                 bases: Optional[np.ndarray] = None, stream_budget: Optional[int] = None):
        """data: (N, 4L) one-hot float32; or bases: (N, L) uint8 codes 0..3 (one-hot built on the
        device, cdlm_problem_create_from_bases).  stream_budget: device bytes the problem may use
        (0: 80% of the free device memory); a problem whose footprint exceeds it keeps its state
        in pinned host memory and moves its sequences through the device in chunks
        (cdlm_problem_create_streamed; data only), with results equal to the resident problem's."""
        self.ctx = ctx or default_context()
        L = load()
        if bases is not None:
            bases = np.ascontiguousarray(bases, np.uint8)
            N, L4 = bases.shape[0], 4 * bases.shape[1]
        else:
            data = _f32c(data)                 # (N, 4L)
            N, L4 = data.shape
        fl = np.ascontiguousarray(filter_len, dtype=np.int32)
        self._fl = fl
        prm = cdlm_params(len(fl), fl.ctypes.data_as(C.POINTER(C.c_int32)), rho, sigma, lam, sherman_csc, sherman_cdl)
        g = int(fl.max())
        Lg = L4 + 2 * g
        d0 = None if D0 is None else _f32c(D0)
        x0 = None if X0 is None else _f32c(X0)
        if d0 is not None and d0.shape != (len(fl), Lg):
            raise ValueError("D0 must be (M, Lg) = %s, got %s" % ((len(fl), Lg), d0.shape))
        if x0 is not None and x0.shape != (N, len(fl), Lg):
            raise ValueError("X0 must be (N, M, Lg) = %s, got %s" % ((N, len(fl), Lg), x0.shape))
        h = C.c_void_p()
        d0p = None if d0 is None else d0.ctypes.data
        x0p = None if x0 is None else x0.ctypes.data
        if stream_budget is not None:
            if bases is not None:
                raise ValueError("stream_budget takes one-hot data, not bases")
            st = L.cdlm_problem_create_streamed(self.ctx.h, C.byref(prm), data.ctypes.data, L4, N, d0p, x0p,
                                                seed & 0xFFFFFFFFFFFFFFFF, int(stream_budget), C.byref(h))
        elif bases is not None:
            st = L.cdlm_problem_create_from_bases(self.ctx.h, C.byref(prm), bases.ctypes.data, L4 // 4, N, d0p, x0p,
                                                  seed & 0xFFFFFFFFFFFFFFFF, C.byref(h))
        else:
            st = L.cdlm_problem_create(self.ctx.h, C.byref(prm), data.ctypes.data, L4, N, d0p, x0p,
                                       seed & 0xFFFFFFFFFFFFFFFF, C.byref(h))
        check(st, self.ctx.h)
        self.h = h
        dims = [C.c_int32() for _ in range(5)]
        check(L.cdlm_problem_dims(h, *[C.byref(x) for x in dims]), self.ctx.h)
        self.Lg, self.gamma, self.L, self.N, self.M = [x.value for x in dims]
        self.lam = lam
        # the Sherman-Morrison constants are fixed when the problem is created (the device
        # descriptor carries them); calls that name another value are refused
        self.sherman_csc, self.sherman_cdl = float(np.float32(sherman_csc)), float(np.float32(sherman_cdl))

    # -- lifecycle
    def close(self):
        if getattr(self, "h", None):
            check(load().cdlm_problem_destroy(self.h), self.ctx.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self):
        check(load().cdlm_problem_reset(self.h), self.ctx.h)

    @property
    def device_bytes(self) -> int:
        return int(load().cdlm_problem_bytes(self.h))

    @property
    def chunks(self) -> int:
        """Chunks the sequences move through the device in (1: resident)."""
        c = C.c_int32()
        check(load().cdlm_problem_chunks(self.h, C.byref(c)), self.ctx.h)
        return c.value

    # -- reference methods
    def _check_sherman(self, name: str, value: Optional[float], fixed: float):
        if value is not None and float(np.float32(value)) != fixed:
            raise CdlmError(ERR_ARG, "%s=%r differs from the problem's constant %r, which is fixed when the problem is "
                                     "created (pass it to CDL(..., %s=...))" % (name, value, fixed, name.replace(
                                         "sherman_const_", "sherman_")))

    def code_update(self, num_iter: int, sherman_const_csc: Optional[float] = None):
        self._check_sherman("sherman_const_csc", sherman_const_csc, self.sherman_csc)
        check(load().cdlm_code_update(self.h, int(num_iter)), self.ctx.h)

    def dict_update(self, num_iter: int, sherman_const_cdl: Optional[float] = None):
        self._check_sherman("sherman_const_cdl", sherman_const_cdl, self.sherman_cdl)
        check(load().cdlm_dict_update(self.h, int(num_iter)), self.ctx.h)

    def update_best(self):
        obj, imp = C.c_double(), C.c_int32()
        check(load().cdlm_update_best(self.h, C.byref(obj), C.byref(imp)), self.ctx.h)
        return obj.value, bool(imp.value)

    def objective(self) -> float:
        obj = C.c_double()
        check(load().cdlm_objective(self.h, C.byref(obj)), self.ctx.h)
        return obj.value

    def obj_best(self) -> float:
        v = C.c_float()
        check(load().cdlm_obj_best(self.h, C.byref(v)), self.ctx.h)
        return v.value

    def optimize(self, iter_num: int, csc_iter: int, cdl_iter: int) -> np.ndarray:
        tr = np.zeros(max(iter_num, 1), np.float64)
        check(load().cdlm_optimize(self.h, iter_num, csc_iter, cdl_iter, tr.ctypes.data), self.ctx.h)
        return tr[:iter_num]

    def code_thresh(self, mult: float = 1.0):
        t, cnt = C.c_double(), C.c_int64()
        check(load().cdlm_code_thresh(self.h, mult, C.byref(t), C.byref(cnt)), self.ctx.h)
        return t.value, cnt.value

    def get(self, name: str) -> np.ndarray:
        fid = FIELD_IDS[name]
        Lg, M, N, L = self.Lg, self.M, self.N, self.L
        shape = {"X": (N, M, Lg), "d": (N, M, Lg), "y2": (N, M, Lg), "y3": (N, M, Lg), "g1": (N, M, Lg),
                 "g3": (N, M, Lg), "DX": (N, Lg), "G2": (N, Lg), "Y1": (N, Lg), "y1": (N, Lg), "g2": (N, Lg),
                 "Td": (N, Lg), "G1": (M, Lg), "G3": (M, Lg), "Dbest": (M, Lg), "Xbest_slice": (N, M, L)}[name]
        out = np.empty(shape, np.float32)
        check(load().cdlm_get(self.h, fid, out.ctypes.data, out.nbytes), self.ctx.h)
        return out

    def timing(self, enable: bool):
        """Start (True) or stop (False) per-launch HIP-event timing; starting clears totals."""
        check(load().cdlm_timing_enable(self.h, 1 if enable else 0), self.ctx.h)

    def timing_read(self, kernel: int):
        """(total device ms, launches) of kernel 0=CSC, 1=dict, 2=best, 3=dict mean reduction."""
        t, n = C.c_double(), C.c_int64()
        check(load().cdlm_timing_read(self.h, kernel, C.byref(t), C.byref(n)), self.ctx.h)
        return t.value, n.value

    def codes(self, mult: float = 1.0):
        """(codes, threshold): X_best[gamma+1:4:gamma+4L, :, :] and mult*std(X_best[X_best .> 0])
        (src/WRAPPER/wrap.jl:12; src/SEARCH/store.jl:15-17)."""
        th, _ = self.code_thresh(mult)
        return self.get("Xbest_slice"), th


class Batch:
    """Several CDL problems trained together: one launch per kernel over all of them (and a
    captured HIP graph of the whole schedule).  The reference trains one dataset per call of
    cdl_optimize (src/WRAPPER/wrap.jl:4-12); each problem of a batch gets the results it
    gets alone.  The batch keeps its problems alive; close() it before closing them."""

    def __init__(self, problems: Sequence[CDL], graph: bool = True):
        if not problems:
            raise ValueError("a batch needs at least one problem")
        self.ctx = problems[0].ctx
        self.problems = list(problems)
        arr = (C.c_void_p * len(self.problems))(*[p.h.value for p in self.problems])
        h = C.c_void_p()
        check(load().cdlm_batch_create(self.ctx.h, arr, len(self.problems), C.byref(h)), self.ctx.h)
        self.h = h
        if not graph:
            self.set_graph(False)

    def close(self):
        if getattr(self, "h", None):
            check(load().cdlm_batch_destroy(self.h), self.ctx.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.problems)

    @property
    def classes(self) -> int:
        """Shape classes (launch sets) the batch runs as."""
        n = C.c_int32()
        check(load().cdlm_batch_classes(self.h, C.byref(n)), self.ctx.h)
        return n.value

    def set_graph(self, enable: bool):
        check(load().cdlm_batch_set_graph(self.h, 1 if enable else 0), self.ctx.h)

    def reset(self):
        check(load().cdlm_batch_reset(self.h), self.ctx.h)

    def optimize(self, iter_num: int, csc_iter: int, cdl_iter: int) -> np.ndarray:
        """(count, iter_num) objective traces."""
        tr = np.zeros((len(self), max(iter_num, 1)), np.float64)
        check(load().cdlm_batch_optimize(self.h, iter_num, csc_iter, cdl_iter, tr.ctypes.data), self.ctx.h)
        return tr[:, :iter_num]

    def train(self, iter_num: int, csc_iter: int, cdl_iter: int, mult: float = 1.0):
        """One full training of every problem: (traces (count, iter_num), thresholds, counts)."""
        n = len(self)
        tr = np.zeros((n, max(iter_num, 1)), np.float64)
        th = np.zeros(n, np.float64)
        cnt = np.zeros(n, np.int64)
        check(load().cdlm_batch_train(self.h, iter_num, csc_iter, cdl_iter, mult, tr.ctypes.data, th.ctypes.data,
                                      cnt.ctypes.data), self.ctx.h)
        return tr[:, :iter_num], th, cnt

    def timing(self, enable: bool):
        check(load().cdlm_batch_timing_enable(self.h, 1 if enable else 0), self.ctx.h)

    def timing_read(self, kernel: int):
        t, n = C.c_double(), C.c_int64()
        check(load().cdlm_batch_timing_read(self.h, kernel, C.byref(t), C.byref(n)), self.ctx.h)
        return t.value, n.value


def autotune(L: int, M: int, max_filter_len: int, n_seq: int = 0, ctx: Optional[Context] = None) -> dict:
    """Measures the launch choices for problems of this shape and keeps the fastest for
    problems created afterwards (cdlm_autotune); returns the choice."""
    from ._lib import cdlm_tuning
    ctx = ctx or default_context()
    t = cdlm_tuning()
    check(load().cdlm_autotune(ctx.h, 4 * L, M, max_filter_len, n_seq, C.byref(t)), ctx.h)
    return {f: getattr(t, f) for f, _ in cdlm_tuning._fields_}


def tuning_set(L: int, M: int, max_filter_len: int, choice: Optional[dict], ctx: Optional[Context] = None):
    """Sets (dict with fft_order, fft4_order, dict_batch, csc_slots) or clears (None) the
    launch choices for a shape."""
    from ._lib import cdlm_tuning
    ctx = ctx or default_context()
    if choice is None:
        check(load().cdlm_tuning_set(ctx.h, 4 * L, M, max_filter_len, None), ctx.h)
        return
    t = cdlm_tuning(**{f: int(choice.get(f, 0)) for f, _ in cdlm_tuning._fields_})
    check(load().cdlm_tuning_set(ctx.h, 4 * L, M, max_filter_len, C.byref(t)), ctx.h)


def tuning_get(L: int, M: int, max_filter_len: int, ctx: Optional[Context] = None) -> dict:
    from ._lib import cdlm_tuning
    ctx = ctx or default_context()
    t = cdlm_tuning()
    check(load().cdlm_tuning_get(ctx.h, 4 * L, M, max_filter_len, C.byref(t)), ctx.h)
    return {f: getattr(t, f) for f, _ in cdlm_tuning._fields_}


def footprint_bytes(L: int, N: int, M: int, max_filter_len: int, keep_x0: bool = False) -> int:
    """Device bytes a problem of this shape holds (cdlm_footprint_bytes)."""
    return int(load().cdlm_footprint_bytes(4 * L, N, M, max_filter_len, 1 if keep_x0 else 0))


def pack_by_footprint(sizes: Sequence[int], budget: int):
    """Groups of indices whose summed device footprint stays within budget (first fit,
    largest first); an item larger than the budget gets a group of its own."""
    order = sorted(range(len(sizes)), key=lambda i: -sizes[i])
    groups, used = [], []
    for i in order:
        for g, u in enumerate(used):
            if u + sizes[i] <= budget:
                groups[g].append(i)
                used[g] += sizes[i]
                break
        else:
            groups.append([i])
            used.append(sizes[i])
    return [sorted(g) for g in groups]


# ---- reference-named free functions ---------------------------------------------------
@dataclass
class DNADataset:
    """The fields of the reference's datasets (FASTA_DNA / Sim_DNA, src/SETUP/LOAD_setup.jl:22-98)
    that the CDL path reads: data_matrix, the (4L, N) one-hot matrix (numpy (N, 4L) float32),
    N and L.  raw_data keeps the sequences when built from strings."""
    data_matrix: np.ndarray
    N: int = 0
    L: int = 0
    raw_data: Optional[Sequence[str]] = None

    def __post_init__(self):
        self.data_matrix = _f32c(self.data_matrix)
        self.N = int(self.data_matrix.shape[0])
        self.L = int(self.data_matrix.shape[1] // 4)

    @classmethod
    def from_strings(cls, dna_strings: Sequence[str]) -> "DNADataset":
        """data_2_dummy of the reads (src/LOAD/to_dummy.jl:29-38)."""
        from .io import data_2_dummy
        return cls(data_2_dummy(dna_strings), raw_data=list(dna_strings))

    @classmethod
    def from_fasta(cls, path: str, max_entries: Optional[int] = None) -> "DNADataset":
        """FASTA_DNA's reads: read_fasta(path; max_entries) then data_2_dummy (LOAD_setup.jl:72-98; the
        shuffled background it also builds comes from an external program and is not made here)."""
        from .io import MAX_NUM_READ_FASTA, read_fasta
        return cls.from_strings(read_fasta(path, MAX_NUM_READ_FASTA if max_entries is None else max_entries))


def init_cdl(setup: CDL_setup, dna_dataset, D0=None, X0=None, seed: int = 0, ctx: Optional[Context] = None) -> CDL:
    """init_cdl(cdl_setup, dna_dataset) (src/SETUP/CDL_setup.jl:273-283): filters of length
    4*filter_size, rho, sigma, lambda from the setup and the dataset's data_matrix.  dna_dataset
    is anything with a data_matrix attribute (DNADataset, or a reference-shaped object) or the
    (N, 4L) matrix itself."""
    data = getattr(dna_dataset, "data_matrix", dna_dataset)
    return CDL(data, setup.filter_lens(), setup.rho, setup.sigma, setup.lam, D0=D0, X0=X0, seed=seed, ctx=ctx)


def code_update(c: CDL, num_iter: int, type_complex=None, sherman_const_csc: float = 2.0):
    """code_update!(c, num_iter, ComplexF32; sherman_const_csc=2f0) (src/CDL/train.jl:1-44).  The
    Sherman-Morrison constant is a problem parameter fixed at creation: a call naming another
    value raises CdlmError instead of running with the problem's."""
    c.code_update(num_iter, sherman_const_csc=sherman_const_csc)


def dict_update(c: CDL, num_iter: int, type_complex=None, sherman_const_cdl: float = 2.0):
    """dict_update!(c, num_iter, ComplexF32; sherman_const_cdl=2f0) (src/CDL/train.jl:46-89); the
    constant is checked as in code_update."""
    c.dict_update(num_iter, sherman_const_cdl=sherman_const_cdl)


def update_best(c: CDL):
    """update_best!(c, dat) (src/CDL/train.jl:91-98) -> (objective, improved)."""
    return c.update_best()


def objective(c: CDL) -> float:
    """objective(c, signals) (src/SETUP/CDL_setup.jl:286)."""
    return c.objective()


def cdl_optimize(setup: CDL_setup, data: np.ndarray, D0=None, X0=None, seed: int = 0, ctx: Optional[Context] = None,
                 trainings: Optional[int] = None):
    """The CDL part of cdl_optimize (src/WRAPPER/wrap.jl:4-12): for each training, init,
    iter_num x (csc_iter_num code updates, update_best, cdl_iter_num dict updates), then the
    code threshold.  Returns a list of (codes (N, M, L), threshold, objective trace)."""
    out = []
    for t in range(trainings if trainings is not None else setup.num_train):
        c = init_cdl(setup, data, D0=D0, X0=X0, seed=seed + t, ctx=ctx)
        trace = c.optimize(setup.iter_num, setup.csc_iter_num, setup.cdl_iter_num)
        codes, th = c.codes(setup.std_threshold_multiplier)
        out.append((codes, th, trace))
        c.close()
    return out
