"""JSON ``TreeLikelihoodModel`` backed by the HIP engine.

Accepts the object graph of torchtree 1.0.2 for this path and returns the
same ``[*sample_shape, 1]`` log-likelihood, differentiable with respect to
every parameter tensor (``requires_grad``):

* ``TreeLikelihoodModel`` (``evolution/tree_likelihood.py:281-450``): keys ``id``,
  ``tree_model``, ``site_model``, ``substitution_model``, ``site_pattern``, optional
  ``branch_model``, ``use_ambiguities``, ``use_tip_states``; an extra optional
  ``precision`` key ("float32" | "float64" | "bf16x3": FP32 partials with the dense-state
  products as split BF16 on the matrix cores) picks the partials precision
  (default: torch's default dtype, as the reference follows it).
* ``Parameter`` (``core/parameter.py:189``): ``tensor``, ``full``, ``zeros``, ``ones``
  (+ ``value``, ``dtype``, ``requires_grad``).
* ``Taxa``/``Taxon``, ``Alignment`` (nucleotide / amino acid / codon / general data
  types; ``sequences`` or ``file``), ``SitePattern``, ``AttributePattern``
  (``attribute_pattern.py:13-74``, discrete traits).
* ``UnRootedTreeModel`` (``tree_model.py:225-344``), ``TimeTreeModel`` (349-463),
  ``ReparameterizedTimeTreeModel`` with ``ratios``/``root_height`` (527-707).
* ``JC69``, ``HKY``, ``GTR`` (``nucleotide.py``), ``GeneralJC69``,
  ``GeneralSymmetricSubstitutionModel``, ``EmpiricalSubstitutionModel``, ``LG``, ``WAG``
  (``amino_acid.py:12-105``),
  ``GeneralNonSymmetricSubstitutionModel`` (``general.py``; P = expm(Q t) through
  torch, as ``abstract.py:88-94``), ``MG94`` (``codon.py``).
* ``ConstantSiteModel``, ``InvariantSiteModel``, ``WeibullSiteModel``
  (``site_model.py``) and ``GammaSiteModel`` (new; torchtree has none).
* ``StrictClockModel``, ``SimpleClockModel`` (``branch_model.py:38-77``).

Types are matched on the last component of the (possibly dotted) type name,
as ``core/utils.py:116-125`` resolves registered classes. Objects given as a
string are looked up by id in ``dic``.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import alignment as al
from . import models as hm
from .empirical import empirical
from .engine import TreeJob
from .likelihood import TipData, _PruneFn
from .tree import (Topology, heights_to_ratios, leaf_heights_from_dates, ratio_bounds,
                   ratios_to_heights, time_branch_lengths)

_REGISTRY = {}


def register(name):
    def deco(fn):
        _REGISTRY[name] = fn
        return fn

    return deco


class JSONParseError(ValueError):
    pass


def process_object(data, dic):
    if isinstance(data, str):
        if data not in dic:
            raise JSONParseError(f"Object with ID `{data}' not found")
        return dic[data]
    if not isinstance(data, dict) or "type" not in data:
        raise JSONParseError(f"cannot build an object from {data!r}")
    kind = data["type"].split(".")[-1]
    fn = _REGISTRY.get(kind)
    if fn is None:
        raise JSONParseError(f"unknown type `{data['type']}'")
    obj = fn(data, dic)
    if data.get("id") is not None:
        if data["id"] in dic and dic[data["id"]] is not obj:
            raise JSONParseError(f"Object with ID `{data['id']}' already exists")
        dic[data["id"]] = obj
    return obj


def _opt(data, key, dic):
    return process_object(data[key], dic) if key in data else None


# ------------------------------------------------------------- parameters


class Parameter:
    """A named tensor whose value models read at call time (like torchtree.Parameter)."""

    def __init__(self, id_, tensor):
        self.id = id_
        self.tensor = tensor

    @property
    def shape(self):
        return self.tensor.shape


_DTYPES = {"torch.float64": torch.float64, "float64": torch.float64, "torch.float32": torch.float32,
           "float32": torch.float32, "torch.int64": torch.int64, "torch.long": torch.int64}


@register("Parameter")
def _parameter(data, dic):
    dtype = _DTYPES.get(data.get("dtype")) if "dtype" in data else None
    if "tensor" in data:
        t = torch.tensor(data["tensor"], dtype=dtype)
        if dtype is None and t.is_floating_point():
            t = t.to(torch.get_default_dtype())
    elif "full" in data:
        size = data["full"] if isinstance(data["full"], list) else [data["full"]]
        t = torch.full(size, data["value"], dtype=dtype or torch.get_default_dtype())
    elif "zeros" in data:
        size = data["zeros"] if isinstance(data["zeros"], list) else [data["zeros"]]
        t = torch.zeros(size, dtype=dtype or torch.get_default_dtype())
    elif "ones" in data:
        size = data["ones"] if isinstance(data["ones"], list) else [data["ones"]]
        t = torch.ones(size, dtype=dtype or torch.get_default_dtype())
    else:
        raise JSONParseError(f"Parameter `{data.get('id')}' has no value")
    if data.get("requires_grad", False):
        t.requires_grad_(True)
    return Parameter(data.get("id"), t)


# ------------------------------------------------------------ taxa / data


class Taxon:
    def __init__(self, id_, attributes):
        self.id = id_
        self.attributes = dict(attributes or {})

    def __getitem__(self, k):
        return self.attributes[k]


@register("Taxon")
def _taxon(data, dic):
    return Taxon(data["id"], data.get("attributes", {}))


@register("Taxa")
def _taxa(data, dic):
    return [process_object(t, dic) for t in data["taxa"]]


def _data_type(spec, dic):
    if isinstance(spec, str):
        if spec == "nucleotide":
            return al.NucleotideType()
        if spec in dic:
            return dic[spec]
        raise JSONParseError(f"unknown data type `{spec}'")
    kind = spec["type"].split(".")[-1]
    if kind == "NucleotideDataType":
        dt = al.NucleotideType()
    elif kind == "AminoAcidDataType":
        dt = al.AminoAcidType()
    elif kind == "CodonDataType":
        dt = al.CodonType(spec.get("genetic_code", "universal"))
    elif kind == "GeneralDataType":
        dt = al.GeneralType(spec["codes"], spec.get("ambiguities", {}))
    else:
        raise JSONParseError(f"unknown data type `{spec['type']}'")
    if spec.get("id") is not None:
        dic[spec["id"]] = dt
    return dt


for _k in ("NucleotideDataType", "AminoAcidDataType", "CodonDataType", "GeneralDataType"):
    register(_k)(lambda data, dic: _data_type(data, dic))


class Alignment:
    def __init__(self, sequences, taxa, data_type):
        order = {t.id: i for i, t in enumerate(taxa)}
        self.sequences = sorted(sequences, key=lambda s: order[s[0]])
        self.taxa = taxa
        self.data_type = data_type


@register("Alignment")
def _alignment(data, dic):
    taxa = process_object(data["taxa"], dic)
    dt = _data_type(data["datatype"], dic)
    if "sequences" in data:
        seqs = [(s["taxon"], s["sequence"]) for s in data["sequences"]]
    elif "file" in data:
        seqs = al.read_fasta(data["file"])
    else:
        raise JSONParseError("sequences or file should be specified in Alignment")
    return Alignment(seqs, taxa, dt)


def parse_indices(text):
    """``"0:100,250,300::2"`` -> [slice(0, 100), 250, slice(300, None, 2)], as the reference's
    ``string_to_list_index`` reads each comma-separated item (core/utils.py:404-419)."""
    out = []
    for item in text.split(","):
        parts = item.split(":")
        if len(parts) == 1:
            out.append(int(parts[0]))
            continue
        start = int(parts[0]) if parts[0] != "" else None
        stop = int(parts[1]) if parts[1] != "" else None
        step = int(parts[2]) if len(parts) == 3 and parts[2] != "" else None
        out.append(slice(start, stop, step))
    return out


class SitePattern:
    """Alignment columns, optionally a selection of them (``indices``, site_pattern.py:69-97:
    each index or slice of every sequence, concatenated in the order given)."""

    def __init__(self, alignment, indices=None):
        self.alignment = alignment
        self.indices = indices

    def sequences(self):
        seqs = [s for _, s in self.alignment.sequences]
        if self.indices is None:
            return seqs
        return ["".join(s[i] for i in self.indices) for s in seqs]

    def compressed(self):
        return al.compress_bytes(self.sequences(), self.alignment.data_type.size)


class AttributePattern:
    """Discrete-trait pattern (attribute_pattern.py:13-74): one pattern of weight 1 whose
    state at each taxon is that taxon's ``attribute`` value under ``data_type``."""

    def __init__(self, taxa, data_type, attribute):
        self.taxa, self.data_type, self.attribute = taxa, data_type, attribute

    def tip_codes(self, use_ambiguities=False, use_tip_states=False):
        """uint8 codes [T][1], the table of their partial rows, weights [1].

        Tip states (compute_tips_states) are one-hot, with states past the last one clamped to
        the all-ones row; otherwise each value's ``partial`` row (compute_tips_partials)."""
        S = self.data_type.state_count
        rows = []
        for t in self.taxa:
            v = t[self.attribute]
            if use_tip_states:
                st = min(self.data_type.state(v), S)
                rows.append(tuple(1.0 if i == st else 0.0 for i in range(S)) if st < S else (1.0,) * S)
            else:
                rows.append(tuple(self.data_type.partial(v, use_ambiguities)))
        uniq = sorted(set(rows))
        codes = torch.tensor([[uniq.index(r)] for r in rows], dtype=torch.uint8)
        return codes, torch.tensor(uniq, dtype=torch.float64), torch.ones(1, dtype=torch.float64)


@register("AttributePattern")
def _attribute_pattern(data, dic):
    taxa = process_object(data["taxa"], dic)
    return AttributePattern(taxa, _data_type(data["data_type"], dic), data["attribute"])


@register("SitePattern")
def _site_pattern(data, dic):
    alignment = process_object(data["alignment"], dic)
    indices = data.get("indices")
    if indices is None:
        return SitePattern(alignment)
    if alignment.data_type.size > 1:
        # the reference selects characters of each sequence, which it cannot do for codons
        # (compress zips their triplets first, site_pattern.py:82-91)
        raise JSONParseError("SitePattern `indices' applies to one-character data types only")
    return SitePattern(alignment, parse_indices(indices))


# ------------------------------------------------------------ tree models


class UnRootedTreeModel:
    def __init__(self, topo, taxa, branch_lengths: Parameter):
        self.topo, self.taxa, self._bl = topo, taxa, branch_lengths

    def branch_lengths(self):
        return self._bl.tensor

    @property
    def sample_shape(self):
        return tuple(self._bl.tensor.shape[:-1])


def _newick(data):
    if "newick" in data:
        return data["newick"]
    if "file" in data:
        with open(data["file"]) as fp:
            return fp.read().strip()
    raise JSONParseError("Tree model requires a file or newick element to be specified")


@register("UnRootedTreeModel")
def _unrooted(data, dic):
    taxa = process_object(data["taxa"], dic)
    topo = Topology(_newick(data), [t.id for t in taxa])
    bl = process_object(data["branch_lengths"], dic)
    if data.get("keep_branch_lengths"):
        bl.tensor = torch.tensor(topo.unrooted_lengths(), dtype=bl.tensor.dtype)
    return UnRootedTreeModel(topo, taxa, bl)


class TimeTreeModel:
    """Branch lengths from node heights (tree_model.py:407-424)."""

    def __init__(self, topo, taxa, internal: Parameter, ratios=False):
        self.topo, self.taxa, self._p = topo, taxa, internal
        self.leaf_heights = torch.tensor(leaf_heights_from_dates([t["date"] for t in taxa]))
        self.ratios = ratios
        self.bounds = ratio_bounds(topo, self.leaf_heights.numpy())

    def internal_heights(self):
        x = self._p.tensor
        return ratios_to_heights(self.topo, self.bounds, x) if self.ratios else x

    def branch_lengths(self):
        return time_branch_lengths(self.topo, self.leaf_heights, self.internal_heights())

    @property
    def sample_shape(self):
        return tuple(self._p.tensor.shape[:-1])


@register("TimeTreeModel")
def _time_tree(data, dic):
    taxa = process_object(data["taxa"], dic)
    topo = Topology(_newick(data), [t.id for t in taxa])
    p = process_object(data["internal_heights"], dic)
    model = TimeTreeModel(topo, taxa, p)
    if data.get("keep_branch_lengths"):
        h = topo.heights_from_lengths(model.leaf_heights.numpy())
        p.tensor = torch.tensor(h, dtype=p.tensor.dtype)
    return model


@register("ReparameterizedTimeTreeModel")
def _reparam_tree(data, dic):
    taxa = process_object(data["taxa"], dic)
    topo = Topology(_newick(data), [t.id for t in taxa])
    if "shifts" in data:
        raise JSONParseError("ReparameterizedTimeTreeModel with shifts is not supported")
    ratios = process_object(data["ratios"], dic)
    root = process_object(data["root_height"], dic)

    class _Cat(Parameter):
        @property
        def tensor(self):
            return torch.cat((ratios.tensor, root.tensor), -1)

        @tensor.setter
        def tensor(self, value):
            ratios.tensor = value[..., :-1]
            root.tensor = value[..., -1:]

    cat = _Cat.__new__(_Cat)
    cat.id = None
    model = TimeTreeModel(topo, taxa, cat, ratios=True)
    if data.get("keep_branch_lengths"):
        h = torch.tensor(topo.heights_from_lengths(model.leaf_heights.numpy()),
                         dtype=ratios.tensor.dtype)
        cat.tensor = heights_to_ratios(topo, model.bounds, h)
    return model


# ---------------------------------------------------- substitution models


class SubstModel:
    """Normalised Q [..., S, S] and frequencies [..., S] read at call time."""

    nonsymmetric = False  # P through expm (non-reversible Q) instead of the eigensystem

    def __init__(self, S, qfn, ffn, shape_fn):
        self.state_count = S
        self._q = qfn
        self._f = ffn
        self._shape = shape_fn

    def frequencies(self):
        return self._f()

    def q_norm(self):
        f = self.frequencies()
        return hm.normalize(self._q(), f)

    @property
    def sample_shape(self):
        return self._shape()


def _shape_of(*params):
    shapes = [tuple(p.tensor.shape[:-1]) for p in params if p is not None]
    return max(shapes, key=len) if shapes else ()


@register("JC69")
def _jc69(data, dic):
    pi = torch.full((4,), 0.25, dtype=torch.float64)
    return SubstModel(4, lambda: hm.gtr_q(torch.ones(6, dtype=torch.float64), pi), lambda: pi,
                      lambda: ())


@register("GeneralJC69")
def _gjc69(data, dic):
    S = int(data["state_count"])
    pi = torch.full((S,), 1.0 / S, dtype=torch.float64)
    ones = torch.ones(S * (S - 1) // 2, dtype=torch.float64)
    return SubstModel(S, lambda: hm.q_from_exchangeabilities(ones, pi), lambda: pi, lambda: ())


@register("HKY")
def _hky(data, dic):
    kappa = process_object(data["kappa"], dic)
    freqs = process_object(data["frequencies"], dic)
    return SubstModel(4, lambda: hm.hky_q(kappa.tensor, freqs.tensor), lambda: freqs.tensor,
                      lambda: _shape_of(kappa, freqs))


@register("GTR")
def _gtr(data, dic):
    rates = process_object(data["rates"], dic)
    freqs = process_object(data["frequencies"], dic)
    return SubstModel(4, lambda: hm.gtr_q(rates.tensor, freqs.tensor), lambda: freqs.tensor,
                      lambda: _shape_of(rates, freqs))


@register("GeneralSymmetricSubstitutionModel")
def _general(data, dic):
    dt = _data_type(data["data_type"], dic)
    rates = process_object(data["rates"], dic)
    freqs = process_object(data["frequencies"], dic)
    S = dt.state_count
    if "mapping" not in data:
        mapping = torch.arange(S * (S - 1) // 2)
    elif isinstance(data["mapping"], list):
        mapping = torch.tensor(data["mapping"])
    else:
        mapping = process_object(data["mapping"], dic).tensor.long()
    return SubstModel(S, lambda: hm.general_symmetric_q(rates.tensor, mapping, freqs.tensor),
                      lambda: freqs.tensor, lambda: _shape_of(rates, freqs))


@register("GeneralNonSymmetricSubstitutionModel")
def _general_nonsymmetric(data, dic):
    """Non-reversible model (general.py:202-291); P = expm(Q t) as in abstract.py:88-94."""
    dt = _data_type(data["data_type"], dic)
    rates = process_object(data["rates"], dic)
    freqs = process_object(data["frequencies"], dic)
    S = dt.state_count
    if "mapping" not in data:
        mapping = torch.arange(S * (S - 1))
    elif isinstance(data["mapping"], list):
        mapping = torch.tensor(data["mapping"])
    else:
        mapping = process_object(data["mapping"], dic).tensor.long()
    model = SubstModel(S, lambda: hm.general_nonsymmetric_q(rates.tensor, mapping, freqs.tensor),
                       lambda: freqs.tensor, lambda: _shape_of(rates, freqs))
    model.nonsymmetric = True
    return model


@register("EmpiricalSubstitutionModel")
def _empirical(data, dic):
    rates = torch.tensor(data["rates"], dtype=torch.float64)
    freqs = torch.tensor(data["frequencies"], dtype=torch.float64)
    S = freqs.shape[-1]
    return SubstModel(S, lambda: hm.q_from_exchangeabilities(rates, freqs), lambda: freqs,
                      lambda: ())


def _fixed_empirical(name):
    """LG / WAG (amino_acid.py:12-105): an EmpiricalSubstitutionModel with a fixed table."""

    def build(data, dic):
        rates, freqs = empirical(name)
        q = hm.q_from_exchangeabilities(rates, freqs)
        return SubstModel(20, lambda: q, lambda: freqs, lambda: ())

    return build


register("LG")(_fixed_empirical("LG"))
register("WAG")(_fixed_empirical("WAG"))


@register("MG94")
def _mg94(data, dic):
    dt = _data_type(data["data_type"], dic)
    alpha = process_object(data["alpha"], dic)
    beta = process_object(data["beta"], dic)
    kappa = process_object(data["kappa"], dic)
    freqs = process_object(data["frequencies"], dic)
    flags = hm.mg94_flags(dt.table)

    def q():
        a, b, k = torch.broadcast_tensors(alpha.tensor, beta.tensor, kappa.tensor)
        return hm.mg94_q(a, b, k, freqs.tensor, flags)

    return SubstModel(dt.state_count, q, lambda: freqs.tensor,
                      lambda: _shape_of(alpha, beta, kappa, freqs))


# ------------------------------------------------------------ site models


class SiteModel:
    def __init__(self, fn, shape_fn):
        self._fn = fn
        self._shape = shape_fn

    def rates_probs(self):
        return self._fn()

    @property
    def sample_shape(self):
        return self._shape()


@register("ConstantSiteModel")
def _constant(data, dic):
    mu = _opt(data, "mu", dic)

    def fn():
        r = mu.tensor if mu is not None else torch.ones(1, dtype=torch.float64)
        return r, torch.ones_like(r)

    return SiteModel(fn, lambda: _shape_of(mu))


@register("InvariantSiteModel")
def _invariant(data, dic):
    inv = process_object(data["invariant"], dic)
    mu = _opt(data, "mu", dic)
    return SiteModel(lambda: hm.invariant_rates(inv.tensor, mu.tensor if mu else None),
                     lambda: _shape_of(inv, mu))


@register("WeibullSiteModel")
def _weibull(data, dic):
    shape = process_object(data["shape"], dic)
    cats = int(data["categories"])
    inv = _opt(data, "invariant", dic)
    mu = _opt(data, "mu", dic)
    return SiteModel(lambda: hm.weibull_rates(shape.tensor, cats, inv.tensor if inv else None,
                                              mu.tensor if mu else None),
                     lambda: _shape_of(shape, inv, mu))


@register("GammaSiteModel")
def _gamma(data, dic):
    shape = process_object(data["shape"], dic)
    cats = int(data["categories"])

    def fn():
        """Rates [*batch, C] (one row per sample of a batched shape [*batch, 1]), probs [C]."""
        x = shape.tensor
        if x.requires_grad:
            raise JSONParseError("GammaSiteModel shape gradients are not provided")
        batch = tuple(x.shape[:-1])
        flat = x.detach().reshape(-1)
        if x.shape[-1:] != (1,) and x.dim() > 0:
            raise JSONParseError("GammaSiteModel shape must have a trailing dimension of 1")
        rows = [hm.gamma_rates(float(v), cats) for v in flat.tolist()]
        rates = torch.stack([r for r, _ in rows]).reshape(batch + (cats,))
        return rates, rows[0][1]

    return SiteModel(fn, lambda: _shape_of(shape))


# ---------------------------------------------------------- branch models


class ClockModel:
    def __init__(self, rate: Parameter, B, strict):
        self._r, self.B, self.strict = rate, B, strict

    def rates(self):
        r = self._r.tensor
        return r.expand(r.shape[:-1] + (self.B,)) if self.strict else r

    @property
    def sample_shape(self):
        return _shape_of(self._r)


@register("StrictClockModel")
def _strict(data, dic):
    tree = process_object(data["tree_model"], dic)
    return ClockModel(process_object(data["rate"], dic), tree.topo.B, True)


@register("SimpleClockModel")
def _simple(data, dic):
    tree = process_object(data["tree_model"], dic)
    return ClockModel(process_object(data["rate"], dic), tree.topo.B, False)


# ------------------------------------------------------- the likelihood


# ------------------------------------------- objects built by torchtree itself
#
# A JSON object may name sub-objects that torchtree's own loader built earlier (a string id
# resolved through ``dic``). They are used through the methods torchtree's TreeLikelihoodModel
# calls on them (tree_likelihood.py:305-356): ``branch_lengths()`` / ``postorder`` / ``taxa``
# (TreeModel), ``q()`` / ``frequencies`` (SubstitutionModel, abstract.py:15-97), ``rates()`` /
# ``probabilities()`` (SiteModel, site_model.py:13-35), ``rates`` (BranchModel), and
# ``compute_tips_partials`` / ``compute_tips_states`` (SitePattern, site_pattern.py:27-31).


def _shape_attr(obj):
    return tuple(getattr(obj, "sample_shape", ()) or ())


class _ForeignTree:
    def __init__(self, obj):
        self.obj = obj
        post = np.asarray([list(map(int, r)) for r in obj.postorder], dtype=np.int32).reshape(-1, 3)
        self.topo = type("Topo", (), {"T": post.shape[0] + 1, "B": 2 * post.shape[0],
                                      "postorder": post})()

    def branch_lengths(self):
        return self.obj.branch_lengths()

    @property
    def sample_shape(self):
        return _shape_attr(self.obj)


class _ForeignSubst:
    def __init__(self, obj):
        self.obj = obj
        self.state_count = obj.frequencies.shape[-1]
        self.nonsymmetric = any("NonSymmetric" in k.__name__ for k in type(obj).__mro__)

    def frequencies(self):
        return self.obj.frequencies

    def q_norm(self):
        return hm.normalize(self.obj.q(), self.obj.frequencies)

    @property
    def sample_shape(self):
        return _shape_attr(self.obj)


class _ForeignSite:
    def __init__(self, obj):
        self.obj = obj

    def rates_probs(self):
        return self.obj.rates(), self.obj.probabilities()

    @property
    def sample_shape(self):
        return _shape_attr(self.obj)


class _ForeignClock:
    def __init__(self, obj):
        self.obj = obj

    def rates(self):
        r = self.obj.rates
        return r() if callable(r) else r

    @property
    def sample_shape(self):
        return _shape_attr(self.obj)


def _adapt(obj, ours, foreign):
    return obj if obj is None or isinstance(obj, ours) else foreign(obj)


class TreeLikelihoodModel:
    """HIP-backed ``TreeLikelihoodModel``: ``model()`` -> log-likelihood [*sample_shape, 1].

    Sub-models are this module's objects (built from inline JSON) or torchtree's own objects
    (resolved by id from ``dic``), used through the methods torchtree's model calls on them.
    The result lives on the device of the branch lengths, as in the reference; nothing is
    copied to the host unless the parameters are host tensors.
    """

    def __init__(self, id_, site_pattern, tree_model, subst_model, site_model, clock_model=None,
                 use_ambiguities=False, use_tip_states=False, precision=None, device=None):
        self.id = id_
        tree_model = _adapt(tree_model, (UnRootedTreeModel, TimeTreeModel), _ForeignTree)
        subst_model = _adapt(subst_model, SubstModel, _ForeignSubst)
        site_model = _adapt(site_model, SiteModel, _ForeignSite)
        clock_model = _adapt(clock_model, ClockModel, _ForeignClock)
        self.site_pattern, self.tree_model = site_pattern, tree_model
        self.subst_model, self.site_model, self.clock_model = subst_model, site_model, clock_model
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        self.device = torch.device(device)
        if precision is None:
            precision = torch.get_default_dtype()
        # "bf16x3": FP32 partials with the dense-state products as split BF16 (hi + lo)
        self.products = "bf16x3" if precision == "bf16x3" else "fp32"
        if precision == "bf16x3":
            precision = torch.float32
        self.dtype = _DTYPES[precision] if isinstance(precision, str) else precision
        topo = tree_model.topo
        if isinstance(site_pattern, AttributePattern):
            dt = site_pattern.data_type
            codes, table, weights = site_pattern.tip_codes(use_ambiguities, use_tip_states)
        elif isinstance(site_pattern, SitePattern):
            cols, weights = site_pattern.compressed()
            dt = site_pattern.alignment.data_type
            codes, table = al.encode_tips(cols, dt, use_ambiguities, use_tip_states)
        else:  # torchtree's SitePattern: its own tip tensors, in taxa order
            S = subst_model.state_count
            if use_tip_states:
                states, weights = site_pattern.compute_tips_states()
                tips = TipData.from_states(list(states)[:topo.T], S, self.device)
            else:
                partials, weights = site_pattern.compute_tips_partials(use_ambiguities)
                tips = TipData.from_partials(list(partials)[:topo.T], self.device)
            self._finish(topo, tips, torch.as_tensor(weights), S)
            return
        S = dt.state_count
        tips = TipData(0, codes.contiguous().to(self.device), table.to(self.device), S,
                       codes.shape[1], topo.T)
        self._finish(topo, tips, weights, S)

    def _finish(self, topo, tips, weights, S):
        self.weights = weights
        self.job = TreeJob([tuple(int(x) for x in r) for r in topo.postorder], tips, weights, S,
                           dtype=self.dtype, device=self.device, products=self.products)

    @property
    def sample_shape(self):
        shapes = [self.tree_model.sample_shape, self.subst_model.sample_shape,
                  self.site_model.sample_shape]
        if self.clock_model is not None:
            shapes.append(self.clock_model.sample_shape)
        return max(shapes, key=len)

    def __call__(self):
        shape = tuple(self.sample_shape)
        if len(shape) > 1:
            raise ValueError("sample shapes with more than one dimension are not supported")
        K = shape[0] if shape else 1
        bl = self.tree_model.branch_lengths()
        if self.clock_model is None:
            bl = bl.expand((K,) + bl.shape[-1:])
            bls = torch.cat((bl, torch.zeros(K, 1, dtype=bl.dtype, device=bl.device)), -1)
        else:
            bls = (self.clock_model.rates() * bl).expand((K,) + bl.shape[-1:])
        rates, probs = self.site_model.rates_probs()
        rates = rates.reshape(-1, rates.shape[-1]).expand(K, -1) if rates.dim() > 1 else rates.unsqueeze(0)
        probs = probs.reshape(-1, probs.shape[-1]).expand(K, -1) if probs.dim() > 1 else probs.unsqueeze(0)
        q = self.subst_model.q_norm()
        f = self.subst_model.frequencies()
        out_dev = bl.device
        if self.subst_model.nonsymmetric:
            ll = self._loglik_expm(bls, rates, probs, q, f)
            return ll.to(out_dev, torch.get_default_dtype()).reshape(shape + (1,))
        if q.dim() == 3 and q.shape[0] == 1:
            q, f = q[0], f.reshape(-1, f.shape[-1])[0]
        ll = self.job.loglik(bls.to(torch.float64), rates.to(torch.float64),
                             probs.to(torch.float64), q.to(torch.float64), f.to(torch.float64))
        out = ll.to(out_dev, torch.get_default_dtype())
        return out.reshape(shape + (1,))

    def _loglik_expm(self, bls, rates, probs, q, f):
        """Non-reversible Q: P = exp(Q bl r) by the HIP matrix exponential in FP64
        (hx_pmatrix_expm; abstract.py:88-94 uses torch.matrix_exp), then the pruning kernels
        and their d/dP through the drop-in autograd function; the gradient reaches Q, the
        branch lengths and the rates through the Frechet derivative (hx_pmatrix_expm_grad)."""
        dev = self.device
        q = q.to(dev, torch.float64)
        if q.dim() == 2:
            q = q.unsqueeze(0)
        mats = hm.pmatrix_expm(q, bls.to(dev, torch.float64), rates.to(dev, torch.float64),
                               self.dtype)  # [K, B, C, S, S]
        f2 = f.to(dev, self.dtype).reshape(-1, f.shape[-1]).contiguous()
        p2 = probs.to(dev, self.dtype).contiguous()
        return _PruneFn.apply(mats, f2, p2, self.job.pruning)

    @classmethod
    def from_json_safe(cls, data, dic):
        """torchtree's loader entry (core/serializable.py:33-61): a missing key is a
        JSONParseError naming the key and the object."""
        try:
            return cls.from_json(data, dic)
        except KeyError as e:
            raise JSONParseError(f"Missing key `{e.args[0]}' for object of type "
                                 f"`TreeLikelihoodModel' with ID `{data.get('id')}'") from None

    @classmethod
    def from_json(cls, data, dic):
        tree = process_object(data["tree_model"], dic)
        site = process_object(data["site_model"], dic)
        subst = process_object(data["substitution_model"], dic)
        pattern = process_object(data["site_pattern"], dic)
        clock = _opt(data, "branch_model", dic)
        return cls(data["id"], pattern, tree, subst, site, clock, data.get("use_ambiguities", False),
                   data.get("use_tip_states", False), data.get("precision"))


register("TreeLikelihoodModel")(lambda data, dic: TreeLikelihoodModel.from_json(data, dic))
