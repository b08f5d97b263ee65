"""Alignments, data types and pattern compression for the engine's tip data.

Follows torchtree 1.0.2: ``NucleotideDataType``, ``AminoAcidDataType``,
``CodonDataType``, ``GeneralDataType`` (``evolution/datatype.py:58-335``),
``read_fasta_sequences`` (``alignment.py:82-92``) and ``compress`` /
``compress_alignment`` / ``compress_alignment_states``
(``site_pattern.py:69-151``). Patterns are the unique alignment columns in
lexicographic order of their characters, with integer counts; each tip is
turned into uint8 codes into a table of its distinct partial rows, which is
the device representation of ``HX_TIPS_CODES``.

Compression runs on numpy byte arrays (``np.unique`` over columns), which
keeps registration of 10^5-pattern alignments off the Python per-column path.
"""
from __future__ import annotations

import itertools

import numpy as np
import torch

_IUPAC = {"A": "A", "C": "C", "G": "G", "T": "T", "U": "T", "R": "AG", "Y": "CT", "M": "AC",
          "W": "AT", "S": "CG", "K": "GT", "B": "CGT", "D": "AGT", "H": "ACT", "V": "ACG",
          "N": "ACGT"}
AMINO_ACIDS = "ACDEFGHIKLMNPQRSTVWY"
UNIVERSAL_CODE = "KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSS*CWCLFLF"
GENETIC_CODES = {
    "universal": UNIVERSAL_CODE,
    "vertebrate mitochondrial": "KNKNTTTT*S*SMIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSSWCWCLFLF",
    "no stops": "KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVVYYQYSSSSWCWCLFLF",
}


class NucleotideType:
    size = 1
    state_count = 4

    def partial(self, ch, use_ambiguities=True):
        u = ch.upper()
        if not use_ambiguities and u not in "ACGTU":
            return (1.0,) * 4
        s = _IUPAC.get(u)
        if s is None:
            return (1.0,) * 4
        return tuple(1.0 if b in s else 0.0 for b in "ACGT")

    def state(self, ch):
        u = ch.upper()
        return "ACGT".index(_IUPAC[u]) if u in _IUPAC and len(_IUPAC[u]) == 1 else 4


class AminoAcidType:
    size = 1
    state_count = 20

    def partial(self, ch, use_ambiguities=True):
        u = ch.upper()
        if u in AMINO_ACIDS:
            return tuple(1.0 if a == u else 0.0 for a in AMINO_ACIDS)
        if use_ambiguities and u == "B":
            return tuple(1.0 if a in "DN" else 0.0 for a in AMINO_ACIDS)
        if use_ambiguities and u == "Z":
            return tuple(1.0 if a in "EQ" else 0.0 for a in AMINO_ACIDS)
        return (1.0,) * 20

    def state(self, ch):
        u = ch.upper()
        return AMINO_ACIDS.index(u) if u in AMINO_ACIDS else 20


class CodonType:
    size = 3

    def __init__(self, genetic_code="universal"):
        self.table = GENETIC_CODES[genetic_code.lower()]
        self.sense = [i for i in range(64) if self.table[i] != "*"]
        self.state_count = len(self.sense)
        self._trip = ["".join(t) for t in itertools.product("ACGT", repeat=3)]
        self._stops_upto = np.cumsum([c == "*" for c in self.table])

    def state(self, codon):
        idx = []
        for ch in codon:
            u = ch.upper()
            if u not in "ACGTU":
                return self.state_count
            idx.append("ACGT".index("T" if u == "U" else u))
        full = idx[0] * 16 + idx[1] * 4 + idx[2]
        return int(full - self._stops_upto[full])

    def partial(self, codon, use_ambiguities=True):
        c = self.state(codon)
        if c >= self.state_count:
            return (1.0,) * self.state_count
        v = [0.0] * self.state_count
        v[c] = 1.0
        return tuple(v)


class GeneralType:
    size = 1

    def __init__(self, codes, ambiguities=None):
        self.codes = list(codes)
        self.state_count = len(self.codes)
        self.amb = dict(ambiguities or {})

    def partial(self, ch, use_ambiguities=True):
        if ch in self.codes:
            return tuple(1.0 if c == ch else 0.0 for c in self.codes)
        if ch in self.amb:
            al = self.amb[ch] if isinstance(self.amb[ch], list) else [self.amb[ch]]
            return tuple(1.0 if c in al else 0.0 for c in self.codes)
        return (1.0,) * self.state_count

    def state(self, ch):
        if ch in self.codes:
            return self.codes.index(ch)
        a = self.amb.get(ch)
        if isinstance(a, str) and a in self.codes:
            return self.codes.index(a)
        return self.state_count


def read_fasta(path):
    seqs, name = {}, None
    with open(path) as fp:
        for line in fp:
            line = line.strip()
            if line.startswith(">"):
                name = line[1:]
                seqs[name] = []
            elif name is not None:
                seqs[name].append(line)
    return [(k, "".join(v)) for k, v in seqs.items()]


def compress_bytes(sequences, size=1):
    """Unique columns as bytes [T][size][N], lexicographic by characters, and counts int64 [N]."""
    arr = np.array([np.frombuffer(s.encode("latin-1"), dtype=np.uint8) for s in sequences])
    T, L = arr.shape
    if L % size:
        raise ValueError("sequence length is not a multiple of the data type size")
    n = L // size
    # rows (taxon, position-in-unit) so lexicographic column order matches tuple-of-strings order
    grid = arr.reshape(T, n, size).transpose(0, 2, 1).reshape(T * size, n)
    uniq, counts = np.unique(grid, axis=1, return_counts=True)
    return uniq.reshape(T, size, -1), torch.tensor(counts, dtype=torch.int64)


def compress(sequences, size=1):
    """Unique columns (lexicographic by characters) and counts (site_pattern.py:69-97).

    Returns (patterns [T][N] as strings of length ``size``, weights int64 [N]).
    """
    cols, counts = compress_bytes(sequences, size)
    T, _, N = cols.shape
    per_taxon = [["".join(chr(x) for x in cols[t, :, j]) for j in range(N)] for t in range(T)]
    return per_taxon, counts


def encode_tips(cols, data_type, use_ambiguities=False, tip_states=False):
    """Tip codes [T][N] (uint8) and the table of distinct partial rows they index.

    ``cols`` is the byte array of ``compress_bytes``. ``tip_states`` follows
    ``compress_alignment_states`` (one-hot states, anything else the all-ones
    row); otherwise each character's ``partial`` row (``compress_alignment``).
    """
    S = data_type.state_count
    T, size, N = cols.shape
    keys = np.zeros((T, N), dtype=np.int64)
    for q in range(size):
        keys = keys * 256 + cols[:, q, :].astype(np.int64)
    uniq, inv = np.unique(keys, return_inverse=True)
    rows, index, lut = [], {}, np.zeros(len(uniq), dtype=np.int64)
    for u_i, key in enumerate(uniq.tolist()):
        chars = "".join(chr((key >> (8 * (size - 1 - q))) & 255) for q in range(size))
        if tip_states:
            st = data_type.state(chars)
            row = tuple(1.0 if k == st else 0.0 for k in range(S)) if st < S else (1.0,) * S
        else:
            row = data_type.partial(chars, use_ambiguities)
        code = index.get(row)
        if code is None:
            code = len(rows)
            if code > 255:
                raise ValueError("more than 256 distinct tip partial rows")
            index[row] = code
            rows.append(row)
        lut[u_i] = code
    codes = lut[inv.reshape(T, N)].astype(np.uint8)
    return torch.from_numpy(codes), torch.tensor(rows, dtype=torch.float64)
