"""The allocation-free hit list (cdlm_pwm_hits_device_noalloc: strand-PWM tiles and hit counter
in the library's fixed per-GPU scratch array) against cdlm_pwm_hits_device, through the C ABI:
identical hit records (bit-exact; both sorted, the record order being unspecified) on two calls
in a row, the second with fewer sequences and PWMs, so the second call runs on the array the
first one left behind.  Inputs too large for the array are refused before any launch."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import flush

pytestmark = pytest.mark.gpu
cdlm = pytest.importorskip("cdlm")
torch = pytest.importorskip("torch")

from cdlm import synth                    # noqa: E402
from cdlm._lib import ERR_UNSUPPORTED, check, load   # noqa: E402


def _hits(name, ctx, d_bases, L, N, pwms, lens, thresh, both, cap):
    dev = torch.device("cuda:0")
    hk = torch.full((cap,), -1, dtype=torch.int32, device=dev)
    hn = torch.full((cap,), -1, dtype=torch.int64, device=dev)
    hp = torch.full((cap,), -1, dtype=torch.int32, device=dev)
    hs = torch.full((cap,), -1, dtype=torch.int8, device=dev)
    torch.cuda.synchronize()
    K, _, maxlen = pwms.shape
    n = C.c_int64(-1)
    st = getattr(load(), name)(ctx.h, d_bases.data_ptr(), L, N, pwms.ctypes.data, lens.ctypes.data, thresh.ctypes.data,
                               K, maxlen, 1 if both else 0, hk.data_ptr(), hn.data_ptr(), hp.data_ptr(), hs.data_ptr(),
                               cap, C.byref(n))
    check(st, ctx.h)
    c = n.value
    rows = np.stack([hk[:c].cpu().numpy().astype(np.int64), hn[:c].cpu().numpy(), hp[:c].cpu().numpy().astype(np.int64),
                     hs[:c].cpu().numpy().astype(np.int64)], 1)
    return rows[np.lexsort(rows.T[::-1])] if c else rows


def test_noalloc_hits_equal_allocating_entry_point_twice():
    ctx = cdlm.default_context()
    cases = [("first", 3000, 300, 96, True), ("second, smaller", 700, 300, 40, False)]
    for tag, N, L, K, both in cases:
        bases, pwms, lens, thresh = synth.pwm_workload(50 + K, N, L, K, len_range=(8, 24))
        d_bases = torch.from_numpy(np.ascontiguousarray(bases)).to("cuda:0")
        torch.cuda.synchronize()
        cap = 400000
        flush("no-alloc hit list %s: N=%d L=%d K=%d both=%s" % (tag, N, L, K, both))
        want = _hits("cdlm_pwm_hits_device", ctx, d_bases, L, N, pwms, lens, thresh, both, cap)
        got = _hits("cdlm_pwm_hits_device_noalloc", ctx, d_bases, L, N, pwms, lens, thresh, both, cap)
        flush("  %d hits (allocating entry point %d)" % (len(got), len(want)))
        assert len(want) > 0
        np.testing.assert_array_equal(got, want)


def test_noalloc_refuses_inputs_beyond_the_array():
    ctx = cdlm.default_context()
    K, maxlen = 4096, 64                     # both strands: 32 * K * maxlen = 8 MiB of tiles
    pwms = np.zeros((K, 4, maxlen), np.float32)
    lens = np.full(K, maxlen, np.int32)
    thresh = np.zeros(K, np.float32)
    d_bases = torch.zeros(1000, dtype=torch.uint8, device="cuda:0")
    torch.cuda.synchronize()
    n = C.c_int64(-1)
    st = load().cdlm_pwm_hits_device_noalloc(ctx.h, d_bases.data_ptr(), 100, 10, pwms.ctypes.data, lens.ctypes.data,
                                             thresh.ctypes.data, K, maxlen, 1, None, None, None, None, 0, C.byref(n))
    assert st == ERR_UNSUPPORTED
    assert "scratch" in load().cdlm_last_error(ctx.h).decode()
