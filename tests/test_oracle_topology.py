"""The oracle's restatement of topology_at_nth_pt! (src/SEARCH/store.jl:64-107), pinned by a
hand-worked case and by an independent sparse restatement (pairs of active entries) on
random inputs.

Hand-worked case (1-based as the reference): C = 12, K = 3 filters: f1 from training 1 of
length 3, f2 and f3 from training 2 of lengths 2 and 3; apart = C - 2 - 1 = 9; max length 3.
Sequence 1 (cover starts 1): f1 at 2 and 5, f2 at 4, f3 at 2 and 12.
  f1@2 -> f2@4   d = 4 - (2+3-1) - 1 = -1            overlap[1,2,1]
  f2@4 -> f1@5   d = 5 - (4+2-1) - 1 = -1            overlap[2,1,1]
  f1@2 -> f3@12  d = 12 - 4 - 1 = 7                  non_overlap[1,3,7]
  f1@5 -> f3@12  d = 12 - 7 - 1 = 4                  non_overlap[1,3,4]
  f3@2 -> f1@5   d = 5 - 4 - 1 = 0                   next_to[3,1]
  f1@2 <-> f3@2  same start, same length: excluded; f2 <-> f3: same training: excluded
Sequence 2 (cover start of training 2 is 5): f1 at 1, f3 at 12 -> f3 starts at 16:
  f1@1 -> f3@16  d = 16 - 3 - 1 = 12 > apart         next_to[1,3]  (the else branch)
"""
import numpy as np
import pytest

import oracle as O


def _hand_case():
    C, K, N = 12, 3, 2
    sp = np.zeros((N, K, C), np.uint8)          # numpy (N, K, C) = reference (C, K, N)
    for k, p in [(0, 2), (0, 5), (1, 4), (2, 2), (2, 12)]:
        sp[0, k, p - 1] = 1
    for k, p in [(0, 1), (2, 12)]:
        sp[1, k, p - 1] = 1
    train = np.array([1, 2, 2], np.int32)
    lens = np.array([3, 2, 3], np.int32)
    cs = np.ones((2, N), np.int32)              # numpy (T, N) = reference (N, T)
    cs[1, 1] = 5
    return sp, train, lens, cs, 3, C - 2 - 1


def test_topology_hand_case():
    sp, train, lens, cs, maxfl, apart = _hand_case()
    over, non, nxt = O.topology(sp, train, lens, cs, maxfl, apart)
    e_over = np.zeros_like(over); e_non = np.zeros_like(non); e_nxt = np.zeros_like(nxt)
    e_over[0, 1, 0] = 1      # overlap[fr=1, fc=2, d=1]  (numpy axes [d, fc, fr])
    e_over[0, 0, 1] = 1      # overlap[2, 1, 1]
    e_non[6, 2, 0] = 1       # non_overlap[1, 3, 7]
    e_non[3, 2, 0] = 1       # non_overlap[1, 3, 4]
    e_nxt[0, 2] = 1          # next_to[3, 1]
    e_nxt[2, 0] = 1          # next_to[1, 3], distance beyond apart
    np.testing.assert_array_equal(over, e_over)
    np.testing.assert_array_equal(non, e_non)
    np.testing.assert_array_equal(nxt, e_nxt)


def sparse_topology(sp, train, lens, cs, maxfl, apart):
    """Independent restatement: for each sequence, all ordered pairs of active entries."""
    N, K, C = sp.shape
    over = np.zeros((maxfl, K, K), np.int64)
    non = np.zeros((max(apart, 1), K, K), np.int64)
    nxt = np.zeros((K, K), np.int64)
    for n in range(N):
        ks, ps = np.nonzero(sp[n])
        for fr, pr in zip(ks, ps):
            for fc, pc in zip(ks, ps):
                tr, tc = train[fr], train[fc]
                if fr == fc or tr == tc:
                    continue
                rs = pr + cs[tr - 1, n]             # 1-based start = (p0 + 1) + cs - 1
                cstart = pc + cs[tc - 1, n]
                if rs > cstart:
                    continue
                if lens[fr] == lens[fc] and rs == cstart:
                    continue
                d = cstart - (rs + lens[fr] - 1) - 1
                if d < 0:
                    over[-d - 1, fc, fr] += 1
                elif 0 < d <= apart:
                    non[d - 1, fc, fr] += 1
                else:
                    nxt[fc, fr] += 1
    return over, non[:apart], nxt


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_topology_matches_sparse_restatement(seed):
    rng = np.random.default_rng(seed)
    N, K, C, T = 7, 6, 15, 3
    sp = (rng.random((N, K, C)) < 0.15).astype(np.uint8)
    train = rng.integers(1, T + 1, K).astype(np.int32)
    lens = rng.integers(2, 6, K).astype(np.int32)
    cs = rng.integers(1, 4, (T, N)).astype(np.int32)
    maxfl, apart = int(lens.max()), C - int(lens.min()) - 1
    got = O.topology(sp, train, lens, cs, maxfl, apart)
    ref = sparse_topology(sp, train, lens, cs, maxfl, apart)
    for g, r in zip(got, ref):
        np.testing.assert_array_equal(g, r)
    assert sum(int(g.sum()) for g in got) > 0


def _pair_case():
    """C = 6, f1 (training 1, length 2), f2 (training 2, length 2), pair (1, 2, overlap 1):
    span = 2 - 1 + 2 = 3 (get_span_len), max span 3.
    Sequence 1 'ACGTAC': f1 at 2 (code 0.6), f2 at 3 (0.2): weight 0.4 over C,G,T;
      f1 at 5 (0.9) with f2 at 6 (0.9) is flagged but c = 5 > C - span + 1 = 4 (ignored).
    Sequence 2 'AAAAAA': f1 at 1 (1.0), f2 at 2 (0.6): weight 0.8 over A,A,A.
    Counts + 0.1: pos1 A .9 C .5 G .1 T .1 (total 1.6), pos2 A .9 C .1 G .5 T .1,
    pos3 A .9 C .1 G .1 T .5; divided by 1.6."""
    C, N = 6, 2
    sp = np.zeros((N, 2, C), np.uint8)
    codes = np.zeros((2, N, 1, C), np.float32)        # (T, N, M, C)
    for n, k, p, v in [(0, 0, 2, 0.6), (0, 1, 3, 0.2), (0, 0, 5, 0.9), (0, 1, 6, 0.9), (1, 0, 1, 1.0), (1, 1, 2, 0.6)]:
        sp[n, k, p - 1] = 1
        codes[k, n, 0, p - 1] = v
    seqs = ["ACGTAC", "AAAAAA"]
    data = np.zeros((N, 4 * C), np.float32)
    for n, sq in enumerate(seqs):
        for i, b in enumerate(sq):
            data[n, 4 * i + "ACGT".index(b)] = 1.0
    return sp, codes, np.array([1, 1]), np.array([1, 2]), np.array([2, 2]), data, np.array([[1, 2, 1]]), 3


def test_pair_pfms_hand_case():
    sp, codes, uinds, train, lens, data, pairs, S = _pair_case()
    span, pfm = O.pair_pfms(sp, codes, uinds, train, lens, data, pairs, S)
    assert span.tolist() == [3]
    e = np.array([[0.9, 0.9, 0.9], [0.5, 0.1, 0.1], [0.1, 0.5, 0.1], [0.1, 0.1, 0.5]], np.float32) / np.float32(1.6)
    np.testing.assert_allclose(pfm[0], e, rtol=1e-6)
