"""Device context (one per GPU): wraps cdlm_create / cdlm_destroy."""
from __future__ import annotations

import ctypes as C

from ._lib import check, load

_default = None


class Context:
    def __init__(self, device: int = 0):
        L = load()
        h = C.c_void_p()
        st = L.cdlm_create(C.byref(h), device)
        if st != 0:
            check(st)
        self.h = h
        self.device = device

    def close(self):
        if getattr(self, "h", None):
            load().cdlm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def mem_info(self):
        f, t = C.c_size_t(), C.c_size_t()
        check(load().cdlm_mem_info(self.h, C.byref(f), C.byref(t)), self.h)
        return f.value, t.value


def default_context() -> Context:
    global _default
    if _default is None:
        _default = Context(0)
    return _default
