"""The device team FFT against the DFT definition (numpy's pocketfft in float64), for every
padded length of the batched workload (Lg = 4L + 2 filter_len, L in {100,150,200,300,500},
filter_len in {16,20,24}; prime factors up to 127) and the quarter-length transforms the
dictionary step uses, both stage orders, forward and inverse.  This pins the transforms
that replace the reference's cuFFT plans (CDL_setup.jl:158-163,195-204).  FP32 tolerance:
relative 2-norm error <= 4e-6 (about 30 ulp of the norm)."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import flush

pytestmark = pytest.mark.gpu
cdlm = pytest.importorskip("cdlm")

LG = sorted({4 * L + 2 * f for L in (100, 150, 200, 300, 500) for f in (16, 20, 24)})
SMALL = [1, 2, 3, 5, 7, 11, 13, 17, 20, 24, 72, 127, 4 * 127, 4 * 79]


def _run(n, twstride, nb, inverse, order, x):
    from cdlm import _lib
    ctx = cdlm.default_context()
    xi = np.ascontiguousarray(np.stack([x.real, x.imag], axis=-1).astype(np.float32))
    out = np.empty_like(xi)
    _lib.check(_lib.load().cdlm_selftest_fft(ctx.h, n, twstride, nb, 1 if inverse else 0, order, xi.ctypes.data,
                                             out.ctypes.data), ctx.h)
    return out[..., 0].astype(np.float64) + 1j * out[..., 1].astype(np.float64)


def _check(n, twstride, nb, order, rng):
    x = (rng.standard_normal((nb, n)) + 1j * rng.standard_normal((nb, n))).astype(np.complex64).astype(np.complex128)
    for inverse in (False, True):
        got = _run(n, twstride, nb, inverse, order, x)
        ref = np.fft.ifft(x, axis=-1) * n if inverse else np.fft.fft(x, axis=-1)
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert err <= 4e-6, (n, twstride, nb, order, inverse, err)


@pytest.mark.parametrize("order", [0, 1])
def test_fft_padded_lengths(order):
    rng = np.random.default_rng(11 + order)
    for n in LG:
        flush("fft n=%d order=%d" % (n, order))
        _check(n, 1, 2, order, rng)
        _check(n // 4, 4, 3, order, rng)


@pytest.mark.parametrize("n", SMALL)
def test_fft_small_and_prime_lengths(n):
    rng = np.random.default_rng(n)
    for order in (0, 1):
        _check(n, 1, 1, order, rng)


def test_fft_matches_real_input_symmetry():
    """A real input has a Hermitian spectrum (the kernels rely on it to pack two real
    transforms into one complex one)."""
    rng = np.random.default_rng(5)
    n = 840
    x = rng.standard_normal((1, n)).astype(np.float32).astype(np.complex128)
    X = _run(n, 1, 1, False, 1, x)[0]
    k = np.arange(1, n)
    assert np.abs(X[k] - np.conj(X[n - k])).max() <= 4e-6 * np.abs(X).max()
