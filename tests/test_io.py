"""Ingest and output formats on either side of the path (src/LOAD/read.jl:2-28,
src/LOAD/to_dummy.jl:2-38, src/SAVE/save_learned_results.jl:184-217), pinned by hand-written
FASTA text, hand-built one-hot matrices and hand-written TRANSFAC files.  The reference holds
no fixture for these functions; the expected values below follow its code read line by line."""
import os

import numpy as np
import pytest

from cdlm import io

FASTA = (">seq1 first\nACGT\nacgt\n"          # two lines joined, lowercase kept by reading
         ">seq2\nAANAA\n"                      # an 'N': dropped
         ">seq3\nccgg\nTTAA\n"
         ">seq4 soft-masked n\nACnT\n"         # an 'n': dropped
         ">seq5\nGGGG\n")


def test_reading_and_read_fasta(tmp_path):
    p = tmp_path / "x.fa"
    p.write_text(FASTA)
    assert io.reading(str(p)) == ["ACGTacgt", "ccggTTAA", "GGGG"]
    assert io.read_fasta(str(p)) == ["ACGTACGT", "CCGGTTAA", "GGGG"]
    assert io.read_fasta(str(p), max_entries=2) == ["ACGTACGT", "CCGGTTAA"]
    assert io.MAX_NUM_READ_FASTA == 10000


def test_reading_keeps_carriage_returns(tmp_path):
    """split on '\\n' only (read.jl:10): a CRLF file keeps its '\\r', which the one-hot
    encoder then refuses, as the reference's dictionary lookup does."""
    p = tmp_path / "crlf.fa"
    p.write_bytes(b">a\r\nAC\r\nGT\r\n")
    assert io.reading(str(p)) == ["AC\rGT\r"]
    with pytest.raises(KeyError):
        io.bases_of(io.read_fasta(str(p)))


def test_one_hot_encoding():
    v = io.dna2dummy("AcGt")
    np.testing.assert_array_equal(v, [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1])
    assert v.dtype == np.float32
    with pytest.raises(KeyError):
        io.dna2dummy("ACNT")
    S = io.data_2_dummy(["AC", "TG", "ACG", "GA"])
    e = np.array([[1, 0, 0, 0, 0, 1, 0, 0],
                  [0, 0, 0, 1, 0, 0, 1, 0],
                  [0, 0, 0, 0, 0, 0, 0, 0],            # length differs from the first: zero row
                  [0, 0, 1, 0, 1, 0, 0, 0]], np.float32)
    np.testing.assert_array_equal(S, e)
    b = io.bases_of(["AC", "TG", "ACG", "GA"])
    np.testing.assert_array_equal(b, [[0, 1], [3, 2], [4, 4], [2, 0]])
    with pytest.raises(ValueError):
        io.data_2_dummy([])


def test_julia_float32_strings():
    cases = {1000.0: "1000.0", 250.0: "250.0", 62.5: "62.5", 0.1: "0.1", 3.0: "3.0",
             1.0 / 3.0: "0.33333334", 1000.0 / 3.0: "333.33334", 0.0001: "0.0001", 9e-5: "9.0f-5",
             1e-6: "1.0f-6", 123456.0: "123456.0", 1e6: "1.0f6", 2.5e7: "2.5f7", -0.25: "-0.25", 0.0: "0.0"}
    for x, s in cases.items():
        assert io.julia_string_f32(x) == s, (x, io.julia_string_f32(x), s)


def test_save_pfms_as_transfac(tmp_path):
    p1 = np.array([[0.25, 1.0, 0.0], [0.25, 0.0, 0.5], [0.25, 0.0, 0.5], [0.25, 0.0, 0.0]], np.float32)
    p2 = np.full((4, 10), 0.25, np.float32)
    p2[:, 9] = [0.0625, 0.1875, 0.25, 0.5]
    io.save_pfms_as_transfac(str(tmp_path), [p1, p2], [2, 1])
    head = "ID\t\nXX\t\nBF\t\nXX\t\nP0\tA\tC\tG\tT\n"
    d2 = head + ("01\t250.0\t250.0\t250.0\t250.0\n"
                 "02\t1000.0\t0.0\t0.0\t0.0\n"
                 "03\t0.0\t500.0\t500.0\t0.0\n") + "XX\t\n"
    d2c = head + ("01\t0.0\t500.0\t500.0\t0.0\n"
                  "02\t0.0\t0.0\t0.0\t1000.0\n"
                  "03\t250.0\t250.0\t250.0\t250.0\n") + "XX\t\n"
    assert open(os.path.join(tmp_path, "d2.transfac")).read() == d2
    assert open(os.path.join(tmp_path, "d2_c.transfac")).read() == d2c
    d1 = open(os.path.join(tmp_path, "d1.transfac")).read().split("\n")
    assert d1[5] == "01\t250.0\t250.0\t250.0\t250.0"
    assert d1[13] == "09\t250.0\t250.0\t250.0\t250.0"
    assert d1[14] == "10\t62.5\t187.5\t250.0\t500.0"
    d1c = open(os.path.join(tmp_path, "d1_c.transfac")).read().split("\n")
    assert d1c[5] == "01\t500.0\t250.0\t187.5\t62.5"
    assert d1c[15] == "XX\t" and d1c[16] == ""
