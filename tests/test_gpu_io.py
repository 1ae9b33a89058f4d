"""A problem built from 1-byte base codes (the one-hot expanded on the device,
cdlm_problem_create_from_bases) equals the problem built from the host one-hot matrix
(data_2_dummy, src/LOAD/to_dummy.jl:29-38): same objective at creation, bit-identical
optimisation trace and codes.  The input goes through the FASTA reader (read.jl:21-28)."""
import numpy as np
import pytest

from gpu_util import flush

pytestmark = pytest.mark.gpu
cdlm = pytest.importorskip("cdlm")


def test_problem_from_fasta_bases_equals_one_hot(tmp_path):
    from cdlm import io, synth
    bases, _ = synth.planted_bases(41, 37, 53, [6, 8], 0.4)
    seqs = ["".join("ACGT"[b] for b in row) for row in bases]
    seqs[5] = seqs[5][:-3]                      # another length: a zero one-hot row
    p = tmp_path / "in.fa"
    p.write_text("".join(">s%d\n%s\n%s\n" % (i, s[:20].lower(), s[20:]) for i, s in enumerate(seqs)))
    reads = io.read_fasta(str(p))
    assert reads == seqs
    data = io.data_2_dummy(reads)
    codes = io.bases_of(reads)
    assert (codes[5] == 4).all() and not data[5].any()
    fl = [12] * 5
    a = cdlm.CDL(data, fl, 1.0, 100.0, 0.05, seed=3)
    b = cdlm.CDL(None, fl, 1.0, 100.0, 0.05, seed=3, bases=codes)
    assert (a.N, a.L) == (b.N, b.L) == (37, 53)
    oa, ob = a.objective(), b.objective()
    flush("objective at creation %.9g / %.9g" % (oa, ob))
    assert oa == ob
    ta, tb = a.optimize(3, 4, 1), b.optimize(3, 4, 1)
    assert np.array_equal(ta, tb)
    assert np.array_equal(a.get("Xbest_slice"), b.get("Xbest_slice"))
    assert a.code_thresh(1.0) == b.code_thresh(1.0)
    a.close()
    b.close()
