"""Tip encodings: the oracle's data-type functions pinned, and the engine's tip data against them.

The oracle's ``nucleotide_code`` / ``amino_acid_code`` / ``codon_code`` and their ``*_partial``
forms are pinned by the reference's own fixtures (``test/test_datatype.py:11-89``: the IUPAC
ambiguity sets of ``test_general``, the state indices and all-ones rows of the other tests).
The engine's tip codes (``alignment.encode_tips``, the uint8 codes plus row table that the
kernels read) are then compared with ``compress_alignment_states`` (``site_pattern.py:127-151``,
codes clamped to S, the all-ones column for S) and ``compress_alignment``
(``site_pattern.py:100-124``) of the oracle on random alignments that mix every kind of
character the data types distinguish.
"""
import itertools

import numpy as np
import pytest
import torch

from oracle import data as odata
from oracle.models import UNIVERSAL
from torchtree_amd import alignment as al

F64 = torch.float64

# test/test_datatype.py:13-26 (test_general): the IUPAC sets a nucleotide code stands for.
IUPAC_SETS = {"U": "T", "R": "AG", "Y": "CT", "M": "AC", "W": "AT", "S": "CG", "K": "GT",
              "B": "CGT", "D": "AGT", "H": "ACT", "V": "ACG"}


def test_oracle_nucleotide_pinned():
    assert odata.nucleotide_code("A") == 0
    assert odata.nucleotide_code("T") == 3 and odata.nucleotide_code("U") == 3
    assert odata.nucleotide_partial("A") == (1.0, 0.0, 0.0, 0.0)
    assert odata.nucleotide_partial("T") == (0.0, 0.0, 0.0, 1.0)
    assert odata.nucleotide_partial("-") == (1.0,) * 4
    for code, bases in IUPAC_SETS.items():
        want = tuple(1.0 if b in bases else 0.0 for b in "ACGT")
        assert odata.nucleotide_partial(code) == want
        assert odata.nucleotide_partial(code.lower()) == want
        # a single-base code encodes below 4, every other ambiguity at or above it
        c = odata.nucleotide_code(code)
        assert (c < 4) == (len(bases) == 1)
        if len(bases) == 1:
            assert c == "ACGT".index(bases)
        # without ambiguities only A, C, G, T, U keep their rows (datatype.py:102-110)
        assert odata.nucleotide_partial(code, use_ambiguities=False) == (
            want if code == "U" else (1.0,) * 4)
    # every printable character outside ACGTU (either case) is an unknown state after the clamp
    for ch in map(chr, range(33, 127)):
        clamped = min(odata.nucleotide_code(ch), 4)
        assert clamped == ("ACGT".index("T" if ch.upper() == "U" else ch.upper())
                           if ch.upper() in "ACGTU" else 4), ch


def test_oracle_amino_acid_pinned():
    assert odata.amino_acid_code("A") == 0 and odata.amino_acid_code("Y") == 19
    assert odata.amino_acid_partial("A") == (1.0,) + (0.0,) * 19
    assert odata.amino_acid_partial("Y") == (0.0,) * 19 + (1.0,)
    assert odata.amino_acid_partial("-") == (1.0,) * 20
    assert odata.amino_acid_partial("/") == (1.0,) * 20
    letters = "ACDEFGHIKLMNPQRSTVWY"
    for ch in map(chr, range(33, 127)):
        clamped = min(odata.amino_acid_code(ch), 20)
        assert clamped == (letters.index(ch.upper()) if ch.upper() in letters else 20), ch
    # the two ambiguity codes: B = D or N, Z = E or Q
    assert odata.amino_acid_partial("B") == tuple(1.0 if a in "DN" else 0.0 for a in letters)
    assert odata.amino_acid_partial("Z") == tuple(1.0 if a in "EQ" else 0.0 for a in letters)
    assert odata.amino_acid_partial("B", use_ambiguities=False) == (1.0,) * 20


def test_oracle_codon_pinned():
    assert odata.codon_code("AAA", UNIVERSAL) == 0 and odata.codon_code("TTT", UNIVERSAL) == 60
    assert odata.codon_code("TAC", UNIVERSAL) == 48 and odata.codon_code("TAT", UNIVERSAL) == 49
    assert odata.codon_code("TGC", UNIVERSAL) == 54 and odata.codon_code("TGG", UNIVERSAL) == 55
    assert odata.codon_partial("AAA", UNIVERSAL, 61) == (1.0,) + (0.0,) * 60
    assert odata.codon_partial("TTT", UNIVERSAL, 61) == (0.0,) * 60 + (1.0,)
    for amb in ("---", "///", "NNN", "ARA"):
        assert odata.codon_code(amb, UNIVERSAL) == 65
        assert odata.codon_partial(amb, UNIVERSAL, 61) == (1.0,) * 61


NUC_CHARS = list("ACGTUacgtu" + "RYMWSKBDHVN" + "?-X")
AA_CHARS = list("ACDEFGHIKLMNPQRSTVWY" + "acdy" + "BZXJ*?-/")


def _seqs(chars, n_taxa, length, seed):
    rng = np.random.default_rng(seed)
    # mostly clean characters so that patterns repeat, with every odd one present
    clean = [c for c in chars if c.isupper()][:4]
    seqs = []
    for _ in range(n_taxa):
        s = rng.choice(clean, length)
        odd = rng.random(length) < 0.2
        s[odd] = rng.choice(chars, int(odd.sum()))
        seqs.append("".join(s))
    return seqs


def _states_as_rows(codes, S):
    eye = torch.cat((torch.eye(S, dtype=F64), torch.ones(1, S, dtype=F64)))
    return eye[codes].t()  # [S, N]


@pytest.mark.parametrize("kind", ["nucleotide", "amino_acid", "codon"])
def test_engine_tip_states_match_oracle(kind):
    """encode_tips(tip_states=True) == compress_alignment_states, as one-hot rows."""
    if kind == "nucleotide":
        seqs = _seqs(NUC_CHARS, 7, 300, 1)
        dt, fn, S, size = al.NucleotideType(), odata.nucleotide_code, 4, 1
    elif kind == "amino_acid":
        seqs = _seqs(AA_CHARS, 7, 300, 2)
        dt, fn, S, size = al.AminoAcidType(), odata.amino_acid_code, 20, 1
    else:
        rng = np.random.default_rng(3)
        trip = ["".join(t) for t in itertools.product("ACGT", repeat=3)] + ["---", "NNN", "AR-"]
        seqs = ["".join(rng.choice(trip, 80)) for _ in range(6)]
        dt, S, size = al.CodonType(), 61, 3
        fn = lambda c: odata.codon_code(c, UNIVERSAL)  # noqa: E731
    ref_codes, ref_w = odata.compress_alignment_states(seqs, fn, S, size=size)
    cols, w = al.compress_bytes(seqs, size)
    codes, tbl = al.encode_tips(cols, dt, tip_states=True)
    assert torch.equal(w, ref_w)
    for t in range(len(seqs)):
        assert int(ref_codes[t].max()) <= S
        assert torch.equal(tbl[codes[t].long()].t(), _states_as_rows(ref_codes[t], S))


@pytest.mark.parametrize("use_ambiguities", [False, True])
def test_engine_nucleotide_partials_match_oracle(use_ambiguities):
    seqs = _seqs(NUC_CHARS, 6, 250, 4)
    ref, ref_w = odata.compress_alignment(
        seqs, lambda c: odata.nucleotide_partial(c, use_ambiguities))
    cols, w = al.compress_bytes(seqs)
    codes, tbl = al.encode_tips(cols, al.NucleotideType(), use_ambiguities=use_ambiguities)
    assert torch.equal(w, ref_w)
    for t in range(len(seqs)):
        assert torch.equal(tbl[codes[t].long()].t(), ref[t])
