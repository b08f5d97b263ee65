"""Tree, alignment and pattern handling, CPU restatement (test infrastructure only).

Follows torchtree 1.0.2:

* ``Tree.index_nodes``        -> ``setup_indexes`` (``evolution/tree_model.py:37-53``):
  tips 0..T-1 in taxa order, internal nodes T..2T-2 numbered in post-order
  (root last).
* ``Tree.postorder``          -> ``AbstractTreeModel.update_traversals`` (``tree_model.py:187-195``)
* ``Tree.unrooted_branch_lengths`` -> ``UnRootedTreeModel.from_json`` with
  ``keep_branch_lengths`` (``tree_model.py:328-344``)
* ``Tree.heights_from_branch_lengths`` -> ``heights_from_branch_lengths`` (``tree_model.py:110-122``)
* ``Tree.time_branch_lengths`` -> ``TimeTreeModel.branch_lengths`` (``tree_model.py:407-424``)
* ``leaf_heights``            -> ``TimeTreeModel.update_leaf_heights`` (``tree_model.py:365-379``)
* ``nucleotide_partial`` / ``nucleotide_code`` -> ``NucleotideDataType`` (``datatype.py:58-114``)
* ``amino_acid_partial`` / ``amino_acid_code`` -> ``AminoAcidDataType`` (``datatype.py:118-177``)
* ``codon_code`` / ``codon_partial`` -> ``CodonDataType`` (``datatype.py:181-297``)
* ``compress`` / ``compress_alignment`` / ``compress_alignment_states`` ->
  ``site_pattern.py:69-151``
* ``read_fasta``              -> ``read_fasta_sequences`` (``alignment.py:82-92``)

The reference parses Newick with DendroPy (absent here); this module has its
own parser. Polytomies are resolved with zero-length edges, as DendroPy's
``resolve_polytomies`` does; the likelihood does not depend on where the
zero-length edge goes. Pinned through the tiny and fluA log-likelihood
goldens and the pattern/encoding tests of the reference
(``test/test_site_pattern.py``, ``test/test_datatype.py``).
"""
from __future__ import annotations

import collections

import torch


class Node:
    __slots__ = ("label", "length", "children", "parent", "index", "date")

    def __init__(self, label=None, length=None):
        self.label = label
        self.length = length
        self.children = []
        self.parent = None
        self.index = -1
        self.date = 0.0

    def is_leaf(self):
        return not self.children


def parse_newick(text: str) -> Node:
    """Recursive-descent Newick parser (labels, lengths, [comments] skipped)."""
    s = text.strip()
    pos = 0

    def skip():
        nonlocal pos
        while pos < len(s):
            if s[pos].isspace():
                pos += 1
            elif s[pos] == "[":
                pos = s.index("]", pos) + 1
            else:
                break

    def label():
        nonlocal pos
        skip()
        if pos < len(s) and s[pos] in "'\"":
            q = s[pos]
            end = s.index(q, pos + 1)
            lab = s[pos + 1:end]
            pos = end + 1
            return lab
        start = pos
        while pos < len(s) and s[pos] not in ",():;[":
            pos += 1
        lab = s[start:pos].strip()
        return lab or None

    def length():
        nonlocal pos
        skip()
        if pos < len(s) and s[pos] == ":":
            pos += 1
            skip()
            start = pos
            while pos < len(s) and s[pos] not in ",();[":
                pos += 1
            return float(s[start:pos])
        return None

    def subtree():
        nonlocal pos
        skip()
        node = Node()
        if s[pos] == "(":
            pos += 1
            while True:
                child = subtree()
                child.parent = node
                node.children.append(child)
                skip()
                if s[pos] == ",":
                    pos += 1
                    continue
                if s[pos] == ")":
                    pos += 1
                    break
                raise ValueError(f"bad newick at {pos}")
        node.label = label()
        node.length = length()
        return node

    root = subtree()
    return root


def resolve_polytomies(node: Node):
    """Make every node binary with zero-length edges (post-order)."""
    for ch in list(node.children):
        resolve_polytomies(ch)
    while len(node.children) > 2:
        a = node.children.pop()
        b = node.children.pop()
        joint = Node(None, 0.0)
        for x in (b, a):
            x.parent = joint
            joint.children.append(x)
        joint.parent = node
        node.children.append(joint)


def postorder_nodes(node: Node):
    out = []
    stack = [(node, False)]
    while stack:
        n, done = stack.pop()
        if done or n.is_leaf():
            out.append(n)
        else:
            stack.append((n, True))
            for ch in reversed(n.children):
                stack.append((ch, False))
    return out


class Tree:
    """A rooted binary tree indexed the torchtree way."""

    def __init__(self, newick: str, taxa: list[str] | None = None):
        self.root = parse_newick(newick)
        resolve_polytomies(self.root)
        nodes = postorder_nodes(self.root)
        leaves = [n for n in nodes if n.is_leaf()]
        if taxa is None:
            # taxon namespace in order of appearance in the Newick string
            taxa = []
            for n in self._preorder_leaves(self.root):
                taxa.append(n.label)
        self.taxa = list(taxa)
        pos = {t: i for i, t in enumerate(self.taxa)}
        if len(leaves) != len(self.taxa):
            raise ValueError("taxa and tree tips differ")
        T = len(self.taxa)
        nxt = T
        for n in nodes:
            if n.is_leaf():
                n.index = pos[n.label]
            else:
                n.index = nxt
                nxt += 1
        self.nodes = sorted(nodes, key=lambda n: n.index)
        self.T = T

    @staticmethod
    def _preorder_leaves(node):
        out = []
        stack = [node]
        while stack:
            n = stack.pop()
            if n.is_leaf():
                out.append(n)
            else:
                stack.extend(reversed(n.children))
        return out

    @property
    def postorder(self):
        """(node, left, right) for every internal node in post-order."""
        return [
            (n.index, n.children[0].index, n.children[1].index)
            for n in postorder_nodes(self.root)
            if not n.is_leaf()
        ]

    def edge_lengths(self):
        """Newick edge lengths of every non-root node, by node index [2T-2]."""
        return [float(n.length or 0.0) for n in self.nodes if n.parent is not None]

    def unrooted_branch_lengths(self):
        """2T-3 lengths with the two root edges merged (tree_model.py:328-344)."""
        bl = self.edge_lengths()
        c1, c2 = self.root.children
        l1, l2 = float(c1.length or 0.0), float(c2.length or 0.0)
        bl[c1.index] += l2
        bl[c2.index] += l1
        return bl[:-1]

    def set_dates(self, dates: list[float]):
        """Leaf heights from dates (TimeTreeModel.update_leaf_heights, tree_model.py:365-379)."""
        self.leaf_heights = leaf_heights(dates)
        for n in self.nodes:
            if n.is_leaf():
                n.date = float(self.leaf_heights[n.index])

    def heights_from_branch_lengths(self, eps=1.0e-6):
        """Internal heights: max over children of child height + max(eps, edge) (tree_model.py:110-122)."""
        h = [0.0] * (2 * self.T - 1)
        for n in postorder_nodes(self.root):
            if n.is_leaf():
                h[n.index] = n.date
            else:
                h[n.index] = max(h[c.index] + max(eps, float(c.length or 0.0)) for c in n.children)
        return torch.tensor(h[self.T:], dtype=torch.float64)

    def time_branch_lengths(self, internal_heights: torch.Tensor):
        """parent height - node height for every non-root node (tree_model.py:407-424)."""
        lh = torch.as_tensor(self.leaf_heights, dtype=internal_heights.dtype)
        heights = torch.cat((lh.expand(internal_heights.shape[:-1] + lh.shape), internal_heights), -1)
        child = [n.index for n in self.nodes if n.parent is not None]
        par = [n.parent.index for n in self.nodes if n.parent is not None]
        return heights[..., par] - heights[..., child]


def leaf_heights(dates):
    dates = list(dates)
    mx = max(dates)
    if min(dates) == 0.0:
        return torch.tensor(dates, dtype=torch.float64)
    return torch.tensor([mx - d for d in dates], dtype=torch.float64)


def read_fasta(path):
    """Ordered (taxon, sequence) pairs (alignment.py:82-92)."""
    seqs = collections.OrderedDict()
    name = None
    with open(path) as fp:
        for line in fp:
            line = line.strip()
            if line.startswith(">"):
                name = line[1:]
                seqs[name] = ""
            elif name is not None:
                seqs[name] += line
    return list(seqs.items())


# --- data types -------------------------------------------------------------

_NUC_ONE = {"A": 0, "C": 1, "G": 2, "T": 3, "U": 3}
_NUC_AMBIG = {
    "R": "AG", "Y": "CT", "M": "AC", "W": "AT", "S": "CG", "K": "GT",
    "B": "CGT", "D": "AGT", "H": "ACT", "V": "ACG", "N": "ACGT",
}
_NUC_CODE = {"A": 0, "C": 1, "G": 2, "T": 3, "U": 3, "R": 5, "Y": 6, "M": 7, "W": 8,
             "S": 9, "K": 10, "B": 11, "D": 12, "H": 13, "V": 14, "N": 15}


def nucleotide_code(ch: str) -> int:
    """NucleotideDataType.encoding: 0-3 for ACGT/U, 4-15 IUPAC, 16 '?', 17 '-'/other."""
    u = ch.upper()
    if u in _NUC_CODE:
        return _NUC_CODE[u]
    if ch == "?":
        return 16
    if ch.isalpha():
        return 16
    return 17


def nucleotide_partial(ch: str, use_ambiguities=True):
    """NucleotideDataType.partial (datatype.py:102-110)."""
    u = ch.upper()
    if not use_ambiguities and u not in "ACGTU":
        return (1.0,) * 4
    if u in _NUC_ONE:
        v = [0.0] * 4
        v[_NUC_ONE[u]] = 1.0
        return tuple(v)
    if u in _NUC_AMBIG:
        return tuple(1.0 if b in _NUC_AMBIG[u] else 0.0 for b in "ACGT")
    return (1.0,) * 4


AMINO_ACIDS = "ACDEFGHIKLMNPQRSTVWY"


def amino_acid_code(ch: str) -> int:
    u = ch.upper()
    if u in AMINO_ACIDS:
        return AMINO_ACIDS.index(u)
    special = {"B": 20, "Z": 21, "X": 22, "*": 23, "?": 24, "-": 25}
    if ch in special or u in special:
        return special.get(ch, special.get(u))
    return 24 if ch.isalpha() else 25


def amino_acid_partial(ch: str, use_ambiguities=True):
    """AminoAcidDataType.partial (datatype.py:162-173)."""
    u = ch.upper()
    if not use_ambiguities and u not in AMINO_ACIDS:
        return (1.0,) * 20
    if u in AMINO_ACIDS:
        v = [0.0] * 20
        v[AMINO_ACIDS.index(u)] = 1.0
        return tuple(v)
    if u == "B":
        return tuple(1.0 if a in "DN" else 0.0 for a in AMINO_ACIDS)
    if u == "Z":
        return tuple(1.0 if a in "EQ" else 0.0 for a in AMINO_ACIDS)
    return (1.0,) * 20


def codon_code(triplet: str, table: str) -> int:
    """CodonDataType.encoding: sense-codon index, 65 for anything ambiguous (datatype.py:274-284)."""
    idx = []
    for ch in triplet:
        u = ch.upper()
        if u not in _NUC_ONE:
            return 65
        idx.append(_NUC_ONE[u])
    full = idx[0] * 16 + idx[1] * 4 + idx[2]
    stops_before = sum(1 for i in range(full + 1) if table[i] == "*")
    return full - stops_before


def codon_partial(triplet: str, table: str, state_count: int):
    c = codon_code(triplet, table)
    if c == 65:
        return (1.0,) * state_count
    v = [0.0] * state_count
    v[c] = 1.0
    return tuple(v)


# --- pattern compression ----------------------------------------------------

def compress(sequences: list[str], size: int = 1, indices=None):
    """Unique columns sorted lexicographically with integer counts (site_pattern.py:69-97).

    ``sequences`` are in taxa order; returns (patterns per taxon, weights). ``indices`` (ints
    and slices, one-character data only) first replaces each sequence by the concatenation of
    ``sequence[index]`` over the list, as the reference does.
    """
    if indices is not None:
        sequences = ["".join(s[i] for i in indices) for s in sequences]
    if size > 1:
        cols = list(zip(*[[s[i:i + size] for i in range(0, len(s), size)] for s in sequences]))
    else:
        cols = list(zip(*sequences))
    counts = collections.Counter(cols)
    order = sorted(counts.keys())
    weights = torch.tensor([counts[p] for p in order])
    per_taxon = list(zip(*order))
    return per_taxon, weights


def compress_alignment(sequences, partial_fn, size=1, dtype=torch.float64):
    """Tip partials [S, N] per taxon plus weights (site_pattern.py:100-124)."""
    per_taxon, weights = compress(sequences, size)
    tips = [torch.tensor([partial_fn(c) for c in pats], dtype=dtype).t().contiguous()
            for pats in per_taxon]
    return tips, weights


def compress_alignment_states(sequences, code_fn, state_count, size=1):
    """Tip codes [N] clamped to ``state_count`` plus weights (site_pattern.py:127-151)."""
    per_taxon, weights = compress(sequences, size)
    tips = [torch.clamp(torch.tensor([code_fn(c) for c in pats], dtype=torch.long), max=state_count)
            for pats in per_taxon]
    return tips, weights
