"""What a training's codes give after thresholding, for fixtures and their checks.

cdl_optimize (src/WRAPPER/wrap.jl:12) takes code_thresh = std(X_best[X_best .> 0]) over all
of X_best; record_training_result! (src/SEARCH/store.jl:15-19) keeps the slice
X_best[gamma+1:4:gamma+4L, :, :], marks the codes above the threshold as active (a Float32
code compared with the Float64 threshold, i.e. in Float64) and takes uinds = the filters with
any active code.  summary() computes these from an oracle's X_best; check() compares a GPU
training's slice and threshold with a stored summary.

A code within the FP32 rounding envelope of the threshold can be active in one FP32
computation and not in another, so the active sets are compared exactly except at entries
the FP64 oracle puts within `band` (relative) of its threshold: those are stored with their
values, every flip must be one of them, and the flips are counted and bounded.
"""
from __future__ import annotations

import numpy as np


def xbest_slice(xb: np.ndarray, gamma: int, L4: int) -> np.ndarray:
    """(N, M, Lg) X_best -> (N, M, L) = the reference's X_best[gamma+1:4:gamma+L4, :, :]."""
    return xb[:, :, gamma:gamma + L4:4]


def summary(xb, gamma, L4, thresh, band=None, n_xs=0, n_act=0):
    xb = np.asarray(xb)
    pos = xb[xb > 0].astype(np.float64)
    xs = xbest_slice(xb, gamma, L4)
    act = xs.astype(np.float64) > float(thresh)
    out = {
        "count": np.int64(pos.size),
        "possum": np.float64(pos.sum()),
        "possq": np.float64((pos * pos).sum()),
        "actcnt": act.sum(axis=2).astype(np.uint16),
        "uinds": (np.nonzero(act.any(axis=(0, 2)))[0] + 1).astype(np.int32),
    }
    if band is not None:
        flat = xs.reshape(-1).astype(np.float64)
        idx = np.nonzero(np.abs(flat - float(thresh)) <= band * float(thresh))[0]
        out["band_idx"] = idx.astype(np.int64)
        out["band_val"] = flat[idx]
        out["band"] = np.float64(band)
    if n_xs:
        out["xs_head"] = np.ascontiguousarray(xs[:n_xs], np.float32)
    if n_act:
        out["act_head"] = np.packbits(act[:n_act].reshape(-1))
        out["act_head_shape"] = np.asarray(act[:n_act].shape, np.int64)
    return out


def active(xs: np.ndarray, thresh: float) -> np.ndarray:
    return np.asarray(xs).astype(np.float64) > float(thresh)


def check(tag, xs, thresh, count, z, sfx="f64", tol=1e-4, log=print, max_flip_frac=1e-4):
    """Compares a training's (N, M, L) X_best slice, threshold and positive count with the
    stored summary of the FP64 oracle.  Returns the number of flipped activities."""
    t64 = float(z["thresh_" + sfx])
    c64 = int(z["count_" + sfx])
    N, M, L = xs.shape
    log("%s code threshold %.9g (fp64 oracle %.9g, rel %.2e); positive codes %d (%d)" % (
        tag, thresh, t64, abs(thresh - t64) / t64, count, c64))
    assert abs(thresh - t64) <= tol * t64, (thresh, t64)
    # the positive count moves with every code near zero: within 1e-4, or 1.5 x the FP32 oracle's
    # own distance from FP64 where the fixture holds it
    c_tol = max(1, tol * c64)
    if "count_f32" in z:
        c_tol = max(c_tol, 1.5 * abs(int(z["count_f32"]) - c64))
    assert abs(count - c64) <= c_tol, (count, c64, c_tol)
    act = active(xs, thresh)
    # uinds: the filters with any active code
    uinds = (np.nonzero(act.any(axis=(0, 2)))[0] + 1).astype(np.int32)
    np.testing.assert_array_equal(uinds, z["uinds_" + sfx], err_msg="uinds")
    # per-(sequence, filter) counts: every difference is covered by near-threshold entries
    band_idx = z["band_idx_" + sfx]
    band = float(z["band_" + sfx])
    band_nm = np.bincount(band_idx // L, minlength=N * M).reshape(N, M)
    d = act.sum(axis=2).astype(np.int64) - z["actcnt_" + sfx].astype(np.int64)
    assert np.all(np.abs(d) <= band_nm), "activity differs away from the threshold band"
    # exact active set, apart from band entries, on the stored head; flips counted
    shape = tuple(int(x) for x in z["act_head_shape_" + sfx])
    nh = shape[0]
    a64 = np.unpackbits(z["act_head_" + sfx], count=int(np.prod(shape))).reshape(shape).astype(bool)
    flip = np.nonzero((act[:nh] != a64).reshape(-1))[0]
    in_band = np.isin(flip, band_idx)
    # flips on the band, over all sequences: decide each band entry with this run's value
    flat = xs.reshape(-1)
    band_act_gpu = flat[band_idx].astype(np.float64) > thresh
    band_act_64 = z["band_val_" + sfx] > t64
    n_flip_all = int((band_act_gpu != band_act_64).sum())
    n_band = band_idx.size
    log("%s active codes %d (fp64 oracle %d); uinds %d; flips in the first %d sequences %d (all on the %.0e band: %s); "
        "flips on the band over all %d sequences %d of %d band entries" % (
            tag, int(act.sum()), int(z["actcnt_" + sfx].sum()), len(uinds), nh, flip.size, band, bool(in_band.all()), N,
            n_flip_all, n_band))
    assert in_band.all(), "flipped activity away from the threshold band"
    assert int(np.abs(d).sum()) <= n_flip_all, (int(np.abs(d).sum()), n_flip_all)
    assert n_flip_all <= max(2, max_flip_frac * act.size), n_flip_all
    if "xs_head_" + sfx in z:
        h = z["xs_head_" + sfx]
        e = np.linalg.norm(xs[:h.shape[0]].astype(np.float64) - h) / np.linalg.norm(h)
        log("%s X_best slice of %d sequences: rel err vs fp64 %.2e" % (tag, h.shape[0], e))
        assert e <= tol, e
    return n_flip_all
