"""Data formats on either side of the path: FASTA input, one-hot / base-code encoding,
TRANSFAC output (host code; the one-hot expansion can run on the device through
CDL(bases=...) / cdlm_problem_create_from_bases).

Reference (CDLmotif.jl):
  reading, read_fasta              src/LOAD/read.jl:2-28
  dna2dummy, data_2_dummy          src/LOAD/to_dummy.jl:2-38 (table `dummy`, TYPE_setup.jl:33-36)
  save_pfms_as_transfac            src/SAVE/save_learned_results.jl:184-217
  max_num_read_fasta = 10000       TYPE_setup.jl:41
"""
from __future__ import annotations

import os
from typing import List, Sequence

import numpy as np

MAX_NUM_READ_FASTA = 10000
_CODE = {"A": 0, "C": 1, "G": 2, "T": 3}


def reading(filepath: str, max_entries: int = MAX_NUM_READ_FASTA) -> List[str]:
    """reading(filepath, max_entries) (read.jl:2-19): records split on '>', the header line
    dropped and the remaining lines joined; records containing 'N' or 'n' dropped; the first
    max_entries kept."""
    with open(filepath, "r", newline="") as f:    # no newline translation: '\r' stays
        reads = f.read()
    out = []
    for rec in reads.split(">"):
        if rec:
            seq = "".join(rec.split("\n")[1:])
            if "N" not in seq and "n" not in seq:
                out.append(seq)
    return out[:max_entries] if len(out) > max_entries else out


def read_fasta(filepath: str, max_entries: int = MAX_NUM_READ_FASTA) -> List[str]:
    """read_fasta(filepath; max_entries) (read.jl:21-28): reading + uppercase."""
    return [s.upper() for s in reading(filepath, max_entries)]


def dna2dummy(dna_string: str) -> np.ndarray:
    """dna2dummy (to_dummy.jl:2-9): the (4L,) one-hot vector; a letter outside ACGT (after
    uppercasing) raises KeyError as the reference's dictionary lookup does."""
    v = np.zeros(4 * len(dna_string), np.float32)
    for i, ch in enumerate(dna_string):
        v[4 * i + _CODE[ch.upper()]] = 1.0
    return v


def bases_of(dna_strings: Sequence[str]) -> np.ndarray:
    """(N, L) uint8 base codes 0..3 of equal-length strings; a sequence whose length differs
    from the first gets code 4 everywhere (the reference leaves that column of the data matrix
    uninitialised, to_dummy.jl:33-35; here it is the all-zero one-hot column).  A letter outside
    ACGT raises KeyError."""
    if len(dna_strings) == 0:
        raise ValueError("There aren't DNA strings found in the input")
    L = len(dna_strings[0])
    lut = np.full(256, 255, np.uint8)
    for ch, c in _CODE.items():
        lut[ord(ch)] = c
        lut[ord(ch.lower())] = c
    out = np.full((len(dna_strings), L), 4, np.uint8)
    for n, s in enumerate(dna_strings):
        if len(s) != L:
            continue
        b = lut[np.frombuffer(s.encode("latin-1"), np.uint8)]
        if (b == 255).any():
            raise KeyError(s[int(np.argmax(b == 255))])
        out[n] = b
    return out


def data_2_dummy(dna_strings: Sequence[str]) -> np.ndarray:
    """data_2_dummy(dna_strings) (to_dummy.jl:29-38) as numpy (N, 4L) float32 (the reference's
    (4L, N)); a sequence of another length than the first gives a zero row."""
    b = bases_of(dna_strings)
    N, L = b.shape
    out = np.zeros((N, L, 4), np.float32)
    ok = b < 4
    n_idx, l_idx = np.nonzero(ok)
    out[n_idx, l_idx, b[ok]] = 1.0
    return out.reshape(N, 4 * L)


def julia_string_f32(x) -> str:
    """A Float32 as Julia 1.7 prints it in string interpolation (Base.Ryu.writeshortest): the
    shortest digits that round-trip; plain form when the decimal point position pt (value in
    [10^(pt-1), 10^pt)) satisfies -4 < pt <= 6, with '.0' after an integral value, else
    scientific d.ddd with the exponent marker 'f' (Float32).  The TRANSFAC writer reaches the
    scientific form only for counts below 1e-4, i.e. PFM entries below 1e-7, which the
    pseudocounted PFMs of at most 10000 sequences never hold."""
    x = np.float32(x)
    if np.isnan(x):
        return "NaN"
    if np.isinf(x):
        return "Inf" if x > 0 else "-Inf"
    if x == 0:
        return "-0.0" if np.signbit(x) else "0.0"
    sign = "-" if x < 0 else ""
    m, e = np.format_float_scientific(abs(x), unique=True, trim="-").split("e")
    digits = m.replace(".", "")
    olength, pt = len(digits), int(e) + 1
    xf = float(x)
    if -4 < pt <= 6 and not (pt >= olength and abs((xf + 0.05) % (10.0 ** (pt - olength)) - 0.05) > 0.05):
        if pt <= 0:
            body = "0." + "0" * (-pt) + digits
        elif pt >= olength:
            body = digits + "0" * (pt - olength) + ".0"
        else:
            body = digits[:pt] + "." + digits[pt:]
        return sign + body
    return "%s%s.%sf%d" % (sign, digits[0], digits[1:] or "0", pt - 1)


def _transfac_text(pfm: np.ndarray) -> str:
    q = np.asarray(pfm, np.float32) * np.float32(1000)      # "make it a count matrix"
    lines = ["ID\t", "XX\t", "BF\t", "XX\t", "P0\tA\tC\tG\tT"]
    for j in range(1, q.shape[1] + 1):
        row = ("0%d" % j) if j < 10 else str(j)
        lines.append(row + "\t" + "\t".join(julia_string_f32(q[a, j - 1]) for a in range(4)))
    lines.append("XX\t")
    return "\n".join(lines) + "\n"


def save_pfms_as_transfac(logo_folder: str, pfms: Sequence[np.ndarray], sort_perm: Sequence[int]):
    """save_pfms_as_transfac (save_learned_results.jl:184-217): for i, ind in enumerate(sort_perm)
    (1-based ind), d{i}.transfac with pfms[ind] x 1000 and d{i}_c.transfac with its reverse
    complement (reverse of all entries, pfm_complement).  The weblogo rendering the reference
    runs next is not part of this library."""
    os.makedirs(logo_folder, exist_ok=True)
    for i, ind in enumerate(sort_perm, start=1):
        pfm = np.asarray(pfms[ind - 1], np.float32)
        with open(os.path.join(logo_folder, "d%d.transfac" % i), "w") as f:
            f.write(_transfac_text(pfm))
        with open(os.path.join(logo_folder, "d%d_c.transfac" % i), "w") as f:
            f.write(_transfac_text(pfm[::-1, ::-1]))
