"""ctypes binding of libcdlm.so (the C ABI declared in include/cdlm.h).

The library is loaded from this directory; a missing library is an error (there is no
fallback path).
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcdlm.so")
HEADER = os.path.join(os.path.dirname(os.path.dirname(HERE)), "include", "cdlm.h")

OK, ERR_ARG, ERR_HIP, ERR_OOM, ERR_UNSUPPORTED, ERR_STATE = range(6)
STATUS_NAMES = {0: "CDLM_OK", 1: "CDLM_ERR_ARG", 2: "CDLM_ERR_HIP", 3: "CDLM_ERR_OOM", 4: "CDLM_ERR_UNSUPPORTED",
                5: "CDLM_ERR_STATE"}

_lib = None


class CdlmError(RuntimeError):
    """Raised for any non-OK status.  Deliberately not a ValueError/ArgumentError: the
    reference's retry loop treats ArgumentError as "no motif found" (wrap.jl:140-149)."""

    def __init__(self, status, msg):
        super().__init__("%s: %s" % (STATUS_NAMES.get(status, status), msg))
        self.status = status


class CdlmOOM(CdlmError):
    pass


class cdlm_params(C.Structure):
    _fields_ = [("M", C.c_int32), ("filter_len", C.POINTER(C.c_int32)), ("rho", C.c_float), ("sigma", C.c_float),
                ("lam", C.c_float), ("sherman_csc", C.c_float), ("sherman_cdl", C.c_float)]


class cdlm_tuning(C.Structure):
    _fields_ = [("fft_order", C.c_int32), ("fft4_order", C.c_int32), ("dict_batch", C.c_int32),
                ("csc_slots", C.c_int32), ("csc_waves", C.c_int32),
                ("csc_mfma", C.c_int32)]


class cdlm_server_opts(C.Structure):
    _fields_ = [("mem_budget", C.c_uint64), ("max_active", C.c_int32), ("csc_iter", C.c_int32), ("cdl_iter", C.c_int32),
                ("use_graph", C.c_int32), ("fill_seqs", C.c_int32), ("policy", C.c_int32),
                ("lockstep", C.c_int32)]


class cdlm_job_spec(C.Structure):
    _fields_ = [("params", cdlm_params), ("s", C.c_void_p), ("L4", C.c_int32), ("N", C.c_int32), ("D0", C.c_void_p),
                ("seed", C.c_uint64), ("iter_num", C.c_int32), ("thresh_mult", C.c_double), ("keep_codes", C.c_int32)]


class cdlm_job_result(C.Structure):
    _fields_ = [("latency_s", C.c_double), ("queue_s", C.c_double), ("service_s", C.c_double), ("thresh", C.c_double),
                ("count", C.c_int64),
                ("obj_best", C.c_float), ("iter_num", C.c_int32), ("L", C.c_int32), ("M", C.c_int32), ("N", C.c_int32)]


class cdlm_server_stats_t(C.Structure):
    _fields_ = [("completed", C.c_int64), ("active", C.c_int64), ("queued", C.c_int64), ("ticks", C.c_int64),
                ("p50_s", C.c_double), ("p99_s", C.c_double), ("bytes_in_use", C.c_uint64), ("budget_bytes", C.c_uint64)]


def header_functions():
    """Names of the functions declared in include/cdlm.h."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cdlm_[a-z0-9_]+)\s*\(", txt)))


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libcdlm.so not built at %s (run __graft_entry__.build())" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, f32, f64 = C.c_void_p, C.c_int32, C.c_int64, C.c_float, C.c_double
    P = C.POINTER
    sig = {
        "cdlm_create": ([P(vp), i32], C.c_int),
        "cdlm_destroy": ([vp], None),
        "cdlm_last_error": ([vp], C.c_char_p),
        "cdlm_version": ([], C.c_char_p),
        "cdlm_mem_info": ([vp, P(C.c_size_t), P(C.c_size_t)], C.c_int),
        "cdlm_problem_create": ([vp, P(cdlm_params), vp, i32, i32, vp, vp, C.c_uint64, P(vp)], C.c_int),
        "cdlm_problem_create_from_bases": ([vp, P(cdlm_params), vp, i32, i32, vp, vp, C.c_uint64, P(vp)], C.c_int),
        "cdlm_problem_create_streamed": ([vp, P(cdlm_params), vp, i32, i32, vp, vp, C.c_uint64, C.c_uint64, P(vp)],
                                         C.c_int),
        "cdlm_problem_chunks": ([vp, P(i32)], C.c_int),
        "cdlm_problem_destroy": ([vp], C.c_int),
        "cdlm_problem_reset": ([vp], C.c_int),
        "cdlm_problem_dims": ([vp, P(i32), P(i32), P(i32), P(i32), P(i32)], C.c_int),
        "cdlm_problem_bytes": ([vp], C.c_size_t),
        "cdlm_footprint_bytes": ([i32, i32, i32, i32, i32], C.c_size_t),
        "cdlm_code_update": ([vp, i32], C.c_int),
        "cdlm_dict_update": ([vp, i32], C.c_int),
        "cdlm_update_best": ([vp, P(f64), P(i32)], C.c_int),
        "cdlm_objective": ([vp, P(f64)], C.c_int),
        "cdlm_optimize": ([vp, i32, i32, i32, vp], C.c_int),
        "cdlm_code_thresh": ([vp, f64, P(f64), P(i64)], C.c_int),
        "cdlm_obj_best": ([vp, P(f32)], C.c_int),
        "cdlm_get": ([vp, i32, vp, C.c_size_t], C.c_int),
        "cdlm_timing_enable": ([vp, i32], C.c_int),
        "cdlm_timing_read": ([vp, i32, P(f64), P(i64)], C.c_int),
        "cdlm_copy_bench": ([vp, C.c_size_t, i32, P(f64)], C.c_int),
        "cdlm_autotune": ([vp, i32, i32, i32, i32, P(cdlm_tuning)], C.c_int),
        "cdlm_tuning_set": ([vp, i32, i32, i32, P(cdlm_tuning)], C.c_int),
        "cdlm_tuning_get": ([vp, i32, i32, i32, P(cdlm_tuning)], C.c_int),
        "cdlm_server_create": ([i32, P(cdlm_server_opts), P(vp)], C.c_int),
        "cdlm_server_destroy": ([vp], C.c_int),
        "cdlm_server_last_error": ([vp], C.c_char_p),
        "cdlm_server_submit": ([vp, P(cdlm_job_spec), P(C.c_uint64)], C.c_int),
        "cdlm_server_poll": ([vp, C.c_uint64, P(i32)], C.c_int),
        "cdlm_server_wait": ([vp, C.c_uint64, i32, P(i32)], C.c_int),
        "cdlm_server_result": ([vp, C.c_uint64, P(cdlm_job_result), vp, vp, C.c_size_t], C.c_int),
        "cdlm_server_release": ([vp, C.c_uint64], C.c_int),
        "cdlm_server_stats": ([vp, P(cdlm_server_stats_t)], C.c_int),
        "cdlm_record_create": ([vp, i32, i32, i32, i32, P(vp)], C.c_int),
        "cdlm_record_destroy": ([vp], C.c_int),
        "cdlm_record_add": ([vp, vp, f64, vp, vp, P(i32)], C.c_int),
        "cdlm_record_dims": ([vp, P(i32), P(i32), P(i32), P(i32)], C.c_int),
        "cdlm_record_filters": ([vp, vp, vp, vp], C.c_int),
        "cdlm_record_start_positions": ([vp, i32, i32, vp], C.c_int),
        "cdlm_record_topology_dims": ([vp, P(i32), P(i32), P(i32)], C.c_int),
        "cdlm_record_topology": ([vp, vp, vp, vp], C.c_int),
        "cdlm_topology": ([vp, vp, vp, vp, vp, i32, i32, i64, i32, i32, i32, vp, vp, vp], C.c_int),
        "cdlm_topology_last_ms": ([vp, P(f64)], C.c_int),
        "cdlm_record_pair_pfms": ([vp, vp, i32, vp, i32, vp, vp], C.c_int),
        "cdlm_selftest_fft": ([vp, i32, i32, i32, i32, i32, vp, vp], C.c_int),
        "cdlm_batch_create": ([vp, P(vp), i32, P(vp)], C.c_int),
        "cdlm_batch_destroy": ([vp], C.c_int),
        "cdlm_batch_set_graph": ([vp, i32], C.c_int),
        "cdlm_batch_reset": ([vp], C.c_int),
        "cdlm_batch_classes": ([vp, P(i32)], C.c_int),
        "cdlm_batch_optimize": ([vp, i32, i32, i32, vp], C.c_int),
        "cdlm_batch_train": ([vp, i32, i32, i32, f64, vp, vp, vp], C.c_int),
        "cdlm_batch_timing_enable": ([vp, i32], C.c_int),
        "cdlm_batch_timing_read": ([vp, i32, P(f64), P(i64)], C.c_int),
        "cdlm_selftest_mfma": ([vp, vp, vp, vp, vp], C.c_int),
        "cdlm_pwm_hits": ([vp, vp, i32, i64, vp, vp, vp, i32, i32, i32, vp, vp, vp, vp, i64, P(i64)], C.c_int),
        "cdlm_pwm_argmax": ([vp, vp, i32, i64, vp, vp, vp, i32, i32, vp], C.c_int),
        "cdlm_pwm_best": ([vp, vp, i32, i64, vp, vp, vp, i32, i32, vp, vp, vp], C.c_int),
        "cdlm_pwm_last_ms": ([vp, P(f64)], C.c_int),
        "cdlm_pwm_pvalue_thresh": ([vp, vp, vp, i32, i32, f64, f64, vp], C.c_int),
        "cdlm_pwm_pvalue_last_ms": ([vp, P(f64)], C.c_int),
        "cdlm_onehot_to_bases": ([vp, vp, i32, i64, vp], C.c_int),
        "cdlm_pwm_hits_device": ([vp, vp, i32, i64, vp, vp, vp, i32, i32, i32, vp, vp, vp, vp, i64, P(i64)], C.c_int),
        "cdlm_pwm_hits_device_noalloc": ([vp, vp, i32, i64, vp, vp, vp, i32, i32, i32, vp, vp, vp, vp, i64, P(i64)],
                                         C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def check(status, ctx=None):
    if status == OK:
        return
    msg = ""
    if ctx is not None:
        m = load().cdlm_last_error(ctx)
        msg = m.decode() if m else ""
    if status == ERR_OOM:
        raise CdlmOOM(status, msg)
    raise CdlmError(status, msg)
