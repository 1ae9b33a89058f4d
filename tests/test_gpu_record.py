"""Activation record and topology on the device (record_training_result!, get_spatial_info!,
topology_at_nth_pt!; src/SEARCH/store.jl:1-144) against the oracle's index-for-index
restatement of the reference kernel: integer counts, compared exactly."""
import numpy as np
import pytest

import oracle as O
from gpu_util import flush

pytestmark = pytest.mark.gpu
cdlm = pytest.importorskip("cdlm")


def test_topology_hand_case_and_random():
    import sys, os
    sys.path.insert(0, os.path.dirname(__file__))
    from test_oracle_topology import _hand_case
    sp, train, lens, cs, maxfl, apart = _hand_case()
    got = cdlm.topology(sp, train, lens, cs, maxfl, apart)
    ref = O.topology(sp, train, lens, cs, maxfl, apart)
    for g, r in zip(got, ref):
        np.testing.assert_array_equal(g, r)
    for seed in range(4):
        rng = np.random.default_rng(100 + seed)
        N, K, C, T = 40, 9, 30, 3
        sp = (rng.random((N, K, C)) < (0.05 + 0.1 * seed)).astype(np.uint8)
        train = rng.integers(1, T + 1, K).astype(np.int32)
        lens = rng.integers(2, 7, K).astype(np.int32)
        cs = rng.integers(1, 5, (T, N)).astype(np.int32)
        maxfl, apart = int(lens.max()), C - int(lens.min()) - 1
        got = cdlm.topology(sp, train, lens, cs, maxfl, apart)
        ref = O.topology(sp, train, lens, cs, maxfl, apart)
        flush("topology seed %d: %d counts" % (seed, sum(int(r.sum()) for r in ref)))
        for g, r in zip(got, ref):
            np.testing.assert_array_equal(g, r)


def _train(seed, N, L, M, fl, with_data=False, data_seed=None):
    """One training (initialisation from seed) on the dataset of data_seed (default: seed);
    the reference's trainings of one search share the dataset (wrap.jl:4-19)."""
    from cdlm import synth
    bases, _ = synth.planted_bases(seed if data_seed is None else data_seed, N, L, [6, 8], 0.4)
    data = synth.onehot(bases)
    c = cdlm.CDL(data, [fl] * M, 1.0, 100.0, 0.05, seed=seed)
    c.optimize(4, 6, 1)
    return (c, data) if with_data else c


def test_record_two_trainings_matches_oracle():
    N, L, M, fl = 60, 40, 6, 12
    rec = cdlm.ActivationRecord(N, L, M, max_trainings=2)
    spec = [fl // 4] * M                      # filter_size in bases (wrap.jl:15)
    expect_sp, expect_u, expect_t, expect_f, expect_xs = [], [], [], [], []
    data = None
    for t, seed in enumerate((11, 12)):
        c, data = _train(seed, N, L, M, fl, with_data=True, data_seed=11)
        th, _ = c.code_thresh(1.0)
        xs = c.get("Xbest_slice")             # (N, M, L)
        act = xs.astype(np.float64) > th      # Float32 codes against a Float64 threshold
        uinds = [m for m in range(M) if act[:, m, :].any()]
        n_added = rec.add(c, th, spec)
        assert n_added == len(uinds)
        expect_u += [u + 1 for u in uinds]
        expect_t += [t + 1] * len(uinds)
        expect_f += [spec[u] for u in uinds]
        expect_sp.append(act[:, uinds, :])    # (N, k_t, L)
        expect_xs.append(xs)
        c.close()
    u, tn, fls = rec.filters()
    np.testing.assert_array_equal(u, expect_u)
    np.testing.assert_array_equal(tn, expect_t)
    np.testing.assert_array_equal(fls, expect_f)
    sp = np.concatenate(expect_sp, axis=1).astype(np.uint8)      # (N, K, C)
    np.testing.assert_array_equal(rec.start_positions(), sp)
    np.testing.assert_array_equal(rec.start_positions(7, 5), sp[7:12])
    K = sp.shape[1]
    maxfl, apart = int(max(fls)), L - int(min(fls)) - 1
    ref = O.topology(sp, tn, fls, np.ones((2, N), np.int32), maxfl, apart)
    got = rec.topology()
    flush("record: K=%d active entries %d, counts %s" % (K, int(sp.sum()), [int(r.sum()) for r in ref]))
    for g, r in zip(got, ref):
        np.testing.assert_array_equal(g, r)
    assert sum(int(r.sum()) for r in ref) > 0
    # pair PFMs (get_pair_pfms!, extractions.jl:18-100) of every overlapping pair; the
    # per-sequence rows follow the reference's order, the sum over sequences is chunked, so the
    # comparison is to 1e-5 relative
    pairs = cdlm.overlap_pairs(got[0], 0)
    assert len(pairs) > 0
    codes = np.stack([xs_t for xs_t in expect_xs])                 # (T, N, M, C)
    span, pfm = rec.pair_pfms(data, pairs, 2 * (fl // 4) - 1)
    rspan, rpfm = O.pair_pfms(sp, codes, u, tn, fls, data, pairs, 2 * (fl // 4) - 1)
    flush("pair PFMs: %d pairs" % len(pairs))
    np.testing.assert_array_equal(span, rspan)
    np.testing.assert_allclose(pfm, rpfm, rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(pfm[:, :, 0].sum(axis=1), 1.0, rtol=1e-5)
    with pytest.raises(cdlm.CdlmError):
        c = _train(13, N, L, M, fl)
        rec.add(c, 0.5, spec)                 # a third training in a two-training record
    rec.close()


def test_record_rejects_bad_threshold_and_shape():
    c = _train(21, 20, 30, 4, 8)
    rec = cdlm.ActivationRecord(20, 30, 4)
    with pytest.raises(cdlm.CdlmError):
        rec.add(c, 0.0, [2] * 4)              # thresh must be > 0 (store.jl:8)
    other = cdlm.ActivationRecord(21, 30, 4)
    with pytest.raises(cdlm.CdlmError):
        other.add(c, 0.1, [2] * 4)            # N differs
    rec.close()
    other.close()
