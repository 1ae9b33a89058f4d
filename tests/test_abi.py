"""The C-ABI library loads and exports every function include/cdlm.h declares (CPU only;
no compute calls)."""
import ctypes as C
import os
import subprocess

import pytest

from cdlm import _lib


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_lib.LIB_PATH), "libcdlm.so missing: run __graft_entry__.build()"
    L = C.CDLL(_lib.LIB_PATH)
    names = _lib.header_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert set(names) <= exported


def test_binding_declares_every_header_function():
    L = _lib.load()
    for n in _lib.header_functions():
        f = getattr(L, n)
        assert f.argtypes is not None or n == "cdlm_version", n


def test_footprint_is_host_arithmetic():
    L = _lib.load()
    b = L.cdlm_footprint_bytes(800, 20000, 64, 20, 0)
    E = 840 * 64 * 20000
    # Xt, y2, y3 (0.75), the g1 half spectra (~1), the stored Y3 half spectra (~0.25), two X_best
    # slices (2 L / Lg = 0.48); x0 not kept
    assert 4 * E * 4 < b < 5 * E * 4
    assert L.cdlm_version().decode().startswith("cdlm")


def test_errors_are_not_argument_errors():
    # the reference retries on ArgumentError (wrap.jl:140-149); library errors must differ
    assert not issubclass(_lib.CdlmError, ValueError)
    assert issubclass(_lib.CdlmOOM, _lib.CdlmError)
