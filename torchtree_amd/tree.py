"""Topology handling for the engine: Newick parsing and torchtree node indexing.

Follows torchtree 1.0.2 semantics (DendroPy is not used):

* node indices: tips 0..T-1 in taxa order, internal nodes T..2T-2 numbered in
  post-order with the root last (``setup_indexes``, ``evolution/tree_model.py:37-53``);
* post-order triples (node, left, right) (``update_traversals``, ``tree_model.py:187-195``);
* unrooted branch lengths with the two root edges merged (``UnRootedTreeModel.from_json``
  with ``keep_branch_lengths``, ``tree_model.py:328-344``);
* heights from branch lengths (``heights_from_branch_lengths``, ``tree_model.py:110-122``)
  and the ratio/root-height transform of ``GeneralNodeHeightTransform``
  (``tree_height_transform.py:7-86``).

Polytomies are resolved with zero-length edges (as DendroPy's
``resolve_polytomies`` does); the likelihood does not depend on the placement.
"""
from __future__ import annotations

import re

import numpy as np
import torch

_TOKEN = re.compile(r"\s*(\(|\)|,|;|:|\[[^\]]*\]|'[^']*'|\"[^\"]*\"|[^\s(),:;\[\]]+)")


class _N:
    __slots__ = ("label", "length", "kids", "parent", "index")

    def __init__(self):
        self.label = None
        self.length = None
        self.kids = []
        self.parent = None
        self.index = -1


def _parse(text: str) -> _N:
    toks = [t for t in _TOKEN.findall(text) if not t.startswith("[")]
    pos = 0

    def node():
        nonlocal pos
        n = _N()
        if toks[pos] == "(":
            pos += 1
            while True:
                ch = node()
                ch.parent = n
                n.kids.append(ch)
                if toks[pos] == ",":
                    pos += 1
                    continue
                if toks[pos] == ")":
                    pos += 1
                    break
                raise ValueError(f"malformed Newick near token {pos}: {toks[pos]!r}")
        if pos < len(toks) and toks[pos] not in "(),:;":
            lab = toks[pos]
            n.label = lab[1:-1] if lab[0] in "'\"" else lab
            pos += 1
        if pos < len(toks) and toks[pos] == ":":
            n.length = float(toks[pos + 1])
            pos += 2
        return n

    return node()


def _binarize(n: _N):
    for ch in list(n.kids):
        _binarize(ch)
    while len(n.kids) > 2:
        a, b = n.kids.pop(), n.kids.pop()
        j = _N()
        j.length = 0.0
        j.kids = [b, a]
        a.parent = b.parent = j
        j.parent = n
        n.kids.append(j)


def _postorder(root: _N):
    out, stack = [], [(root, False)]
    while stack:
        n, done = stack.pop()
        if done or not n.kids:
            out.append(n)
        else:
            stack.append((n, True))
            stack.extend((c, False) for c in reversed(n.kids))
    return out


class Topology:
    """Rooted binary tree with torchtree indices.

    ``taxa``: taxon names defining tip indices; when None the order of
    appearance in the Newick string is used (a fresh DendroPy namespace).
    """

    def __init__(self, newick: str, taxa=None):
        root = _parse(newick.strip())
        _binarize(root)
        order = _postorder(root)
        leaves = [n for n in order if not n.kids]
        if taxa is None:
            taxa = [n.label for n in self._leaves_in_text_order(root)]
        taxa = list(taxa)
        pos = {t: i for i, t in enumerate(taxa)}
        if len(pos) != len(leaves) or any(n.label not in pos for n in leaves):
            raise ValueError("tree tips and taxa differ")
        T = len(taxa)
        nxt = T
        for n in order:
            if n.kids:
                n.index = nxt
                nxt += 1
            else:
                n.index = pos[n.label]
        self.T = T
        self.taxa = taxa
        self._root = root
        self._nodes = sorted(order, key=lambda n: n.index)
        self.postorder = np.array([(n.index, n.kids[0].index, n.kids[1].index)
                                   for n in order if n.kids], dtype=np.int32)
        self.parent = np.full(2 * T - 1, -1, dtype=np.int64)
        for v, a, b in self.postorder:
            self.parent[a] = v
            self.parent[b] = v

    @staticmethod
    def _leaves_in_text_order(root):
        out, stack = [], [root]
        while stack:
            n = stack.pop()
            if n.kids:
                stack.extend(reversed(n.kids))
            else:
                out.append(n)
        return out

    @property
    def B(self):
        return 2 * self.T - 2

    def edge_lengths(self):
        """Newick lengths of every non-root node by index [2T-2]."""
        return np.array([n.length or 0.0 for n in self._nodes if n.parent is not None])

    def unrooted_lengths(self):
        """2T-3 free lengths with the two root edges merged."""
        bl = self.edge_lengths()
        c1, c2 = self._root.kids
        l1, l2 = c1.length or 0.0, c2.length or 0.0
        bl[c1.index] += l2
        bl[c2.index] += l1
        return bl[:-1]

    def preorder_pairs(self):
        """(parent, child) for every non-root node, parents first."""
        out = []
        for v, a, b in self.postorder[::-1]:
            out.append((int(v), int(a)))
            out.append((int(v), int(b)))
        return out

    def heights_from_lengths(self, leaf_heights, eps=1.0e-6):
        h = np.zeros(2 * self.T - 1)
        h[: self.T] = leaf_heights
        for v, a, b in self.postorder:
            la = self._nodes[a].length or 0.0
            lb = self._nodes[b].length or 0.0
            h[v] = max(h[a] + max(eps, la), h[b] + max(eps, lb))
        return h[self.T:]


def leaf_heights_from_dates(dates):
    """Sampling times: dates themselves when time starts at 0, else max - date (tree_model.py:365-379)."""
    dates = np.asarray(dates, dtype=np.float64)
    if dates.min() == 0.0:
        return dates
    return dates.max() - dates


def time_branch_lengths(topo: Topology, leaf_heights: torch.Tensor, internal: torch.Tensor):
    """parent height - child height for every non-root node (TimeTreeModel.branch_lengths)."""
    lh = leaf_heights.to(internal.dtype)
    h = torch.cat((lh.expand(internal.shape[:-1] + lh.shape), internal), -1)
    child = torch.arange(topo.B)
    par = torch.as_tensor(topo.parent[: topo.B])
    return h[..., par] - h[..., child]


def ratio_bounds(topo: Topology, leaf_heights):
    """Lower bound of every node: max sampling time below it (GeneralNodeHeightTransform.update_bounds)."""
    b = np.zeros(2 * topo.T - 1)
    b[: topo.T] = leaf_heights
    for v, a, c in topo.postorder:
        b[v] = max(b[a], b[c])
    return b


def ratios_to_heights(topo: Topology, bounds, x: torch.Tensor) -> torch.Tensor:
    """Ratios of the T-2 non-root internal nodes then root height -> internal heights."""
    T = topo.T
    bnd = torch.as_tensor(bounds[T:], dtype=x.dtype)
    cols = [None] * (T - 1)
    cols[T - 2] = x[..., T - 2]
    for par, ch in topo.preorder_pairs():
        if ch >= T:
            i, p = ch - T, par - T
            cols[i] = bnd[i] + x[..., i] * (cols[p] - bnd[i])
    return torch.stack(cols, -1)


def heights_to_ratios(topo: Topology, bounds, h: torch.Tensor) -> torch.Tensor:
    T = topo.T
    bnd = torch.as_tensor(bounds[T:], dtype=h.dtype)
    out = h.clone()
    for ch in range(T, 2 * T - 2):
        i, p = ch - T, int(topo.parent[ch]) - T
        out[..., i] = (h[..., i] - bnd[i]) / (h[..., p] - bnd[i])
    return out
