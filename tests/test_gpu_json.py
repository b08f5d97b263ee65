"""The JSON TreeLikelihoodModel on the GPU against the reference goldens.

Object graphs follow test/test_tree_likelihood.py:73-342: the tiny JC69
unrooted model (-83.329016), the fluA JC69 + Weibull(0.1)x4 + strict clock
model on a reparameterised time tree (-4618.2062529058, single and a batch
of 3), and batch == single (exact equality, as the reference asserts).
"""
import pytest
import torch

from oracle import data as odata

pytestmark = pytest.mark.gpu


def _taxa(path, dated=False):
    out = []
    for t, _ in odata.read_fasta(path):
        d = {"id": t, "type": "Taxon"}
        if dated:
            d["attributes"] = {"date": float(t.split("_")[-1])}
        out.append(d)
    return {"id": "taxa", "type": "Taxa", "taxa": out}


@pytest.mark.parametrize("precision,tol", [("float64", 1e-9), ("float32", 1e-6)])
def test_tiny_json(gpu, tiny_newick, tiny_fasta, precision, tol):
    print(f"case: tiny JSON {precision}", flush=True)
    from torchtree_amd.treelikelihood_model import TreeLikelihoodModel

    data = {
        "id": "like", "type": "TreeLikelihoodModel", "precision": precision,
        "tree_model": {"id": "tree", "type": "UnRootedTreeModel", "newick": tiny_newick,
                       "branch_lengths": {"id": "branches", "type": "torchtree.Parameter",
                                          "tensor": [0.0], "dtype": "torch.float64"},
                       "keep_branch_lengths": True, "taxa": _taxa(tiny_fasta)},
        "site_model": {"id": "sm", "type": "ConstantSiteModel"},
        "substitution_model": {"id": "m", "type": "JC69"},
        "site_pattern": {"id": "sp", "type": "SitePattern",
                         "alignment": {"id": "alignment", "type": "Alignment",
                                       "datatype": "nucleotide", "file": tiny_fasta, "taxa": "taxa"}},
    }
    like = TreeLikelihoodModel.from_json(data, {})
    out = like()
    assert out.shape == (1,)
    assert abs(float(out) - (-83.329016)) < 83.33 * max(tol, 1e-7) + 1e-6


def _flu(flu_newick, flu_fasta, batch):
    ratios = {"id": "ratios", "type": "Parameter", "tensor": [0.5] * 67, "dtype": "torch.float64"}
    root = {"id": "root_height", "type": "Parameter", "tensor": [20.0], "dtype": "torch.float64"}
    return {
        "id": "like", "type": "torchtree.tree_likelihood.TreeLikelihoodModel", "precision": "float64",
        "tree_model": {"id": "tree_model", "type": "ReparameterizedTimeTreeModel", "newick": flu_newick,
                       "ratios": ratios, "root_height": root, "keep_branch_lengths": True,
                       "taxa": _taxa(flu_fasta, dated=True)},
        "site_model": {"id": "site_model", "type": "WeibullSiteModel", "categories": 4,
                       "shape": {"id": "shape", "type": "Parameter", "tensor": [[0.1]] * batch,
                                 "dtype": "torch.float64"}},
        "substitution_model": {"id": "jc", "type": "JC69"},
        "site_pattern": {"id": "sp", "type": "torchtree.evolution.site_pattern.SitePattern",
                         "alignment": {"id": "alignment", "type": "torchtree.evolution.alignment.Alignment",
                                       "datatype": "nucleotide", "file": flu_fasta, "taxa": "taxa"}},
        "branch_model": {"id": "clock", "type": "StrictClockModel", "tree_model": "tree_model",
                         "rate": {"id": "rate", "type": "Parameter", "tensor": [[0.001]] * batch,
                                  "dtype": "torch.float64"}},
    }


def test_flu_json_golden_and_batch(gpu, flu_newick, flu_fasta):
    print("case: fluA JSON reparameterised time tree", flush=True)
    from torchtree_amd.treelikelihood_model import TreeLikelihoodModel

    like = TreeLikelihoodModel.from_json(_flu(flu_newick, flu_fasta, 1), {})
    single = like()
    assert single.shape == (1, 1)
    assert abs(float(single) - (-4618.2062529058)) < 4618.2 * 1e-5
    like3 = TreeLikelihoodModel.from_json(_flu(flu_newick, flu_fasta, 3), {})
    b3 = like3()
    assert b3.shape == (3, 1)
    assert torch.all(b3 == single)


def test_batch_equals_single_exactly(gpu):
    """test_tree_likelihood.py:172-265: time tree heights, Weibull, strict clock."""
    print("case: batch == single", flush=True)
    from torchtree_amd.treelikelihood_model import TreeLikelihoodModel

    def graph(heights, shape, rate):
        return {
            "id": "like", "type": "TreeLikelihoodModel", "precision": "float64",
            "tree_model": {"id": "tree", "type": "TimeTreeModel", "newick": "(((A,B),C),D);",
                           "internal_heights": {"id": "heights", "type": "Parameter", "tensor": heights,
                                                "dtype": "torch.float64"},
                           "taxa": {"id": "taxa", "type": "Taxa", "taxa": [
                               {"id": n, "type": "Taxon", "attributes": {"date": d}}
                               for n, d in (("A", 0.0), ("B", 1.0), ("C", 4.0), ("D", 5.0))]}},
            "site_model": {"id": "sm", "type": "WeibullSiteModel", "categories": 4,
                           "shape": {"id": "s", "type": "Parameter", "tensor": shape, "dtype": "torch.float64"}},
            "substitution_model": {"id": "jc", "type": "JC69"},
            "site_pattern": {"id": "sites", "type": "SitePattern", "alignment": {
                "id": "alignment", "type": "Alignment", "datatype": "nucleotide", "taxa": "taxa",
                "sequences": [{"taxon": "A", "sequence": "AAG"}, {"taxon": "B", "sequence": "AAC"},
                              {"taxon": "C", "sequence": "AAC"}, {"taxon": "D", "sequence": "AAT"}]}},
            "branch_model": {"id": "clock", "type": "StrictClockModel", "tree_model": "tree",
                             "rate": {"id": "r", "type": "Parameter", "tensor": rate, "dtype": "torch.float64"}},
        }

    batch = TreeLikelihoodModel.from_json(graph([[10.0, 20.0, 30.0], [100.0, 200.0, 300.0]],
                                                [[1.0], [1.0]], [[0.01], [0.001]]), {})
    single = TreeLikelihoodModel.from_json(graph([100.0, 200.0, 300.0], [1.0], [0.001]), {})
    assert single() == batch()[1]


def test_json_gradients_flow(gpu, tiny_newick, tiny_fasta):
    """Gradients reach JSON parameters (branch lengths, GTR rates, frequencies)."""
    print("case: JSON gradients", flush=True)
    from torchtree_amd.treelikelihood_model import TreeLikelihoodModel

    dic = {}
    data = {
        "id": "like", "type": "TreeLikelihoodModel", "precision": "float64",
        "tree_model": {"id": "tree", "type": "UnRootedTreeModel", "newick": tiny_newick,
                       "branch_lengths": {"id": "branches", "type": "Parameter", "tensor": [0.0],
                                          "dtype": "torch.float64"},
                       "keep_branch_lengths": True, "taxa": _taxa(tiny_fasta)},
        "site_model": {"id": "sm", "type": "ConstantSiteModel"},
        "substitution_model": {"id": "gtr", "type": "GTR",
                               "rates": {"id": "rates", "type": "Parameter", "tensor": [1.0, 2.0, 1.0, 1.0, 2.0, 1.0],
                                         "dtype": "torch.float64"},
                               "frequencies": {"id": "freqs", "type": "Parameter", "tensor": [0.3, 0.2, 0.2, 0.3],
                                               "dtype": "torch.float64"}},
        "site_pattern": {"id": "sp", "type": "SitePattern",
                         "alignment": {"id": "alignment", "type": "Alignment",
                                       "datatype": "nucleotide", "file": tiny_fasta, "taxa": "taxa"}},
    }
    prev_dtype = torch.get_default_dtype()
    torch.set_default_dtype(torch.float64)  # the model returns the default dtype; FD needs FP64
    try:
        _json_gradients(TreeLikelihoodModel.from_json(data, dic), dic)
    finally:
        torch.set_default_dtype(prev_dtype)


def _json_gradients(like, dic):
    for k in ("branches", "rates", "freqs"):
        dic[k].tensor.requires_grad_(True)
    out = like()
    out.sum().backward()
    for k in ("branches", "rates", "freqs"):
        g = dic[k].tensor.grad
        assert g is not None and torch.isfinite(g).all(), k
    # central differences on one branch length
    eps = 1e-6
    bl = dic["branches"].tensor.detach().clone()
    def f(x):
        dic["branches"].tensor = x
        return float(like())
    e = torch.zeros_like(bl)
    e[2] = eps
    fd = (f(bl + e) - f(bl - e)) / (2 * eps)
    g = float(like_grad(dic, bl, like))
    assert abs(fd - g) < 1e-6 * max(1.0, abs(fd)), (fd, g)


def like_grad(dic, bl, like):
    dic["branches"].tensor = bl.clone().requires_grad_(True)
    like().sum().backward()
    return dic["branches"].tensor.grad[2]


def _nonsym_graph(tiny_newick, tiny_fasta, kind, rates, mapping):
    subst = {"id": "q", "type": kind, "data_type": "nucleotide",
             "rates": {"id": "rates", "type": "Parameter", "tensor": rates, "dtype": "torch.float64"},
             "frequencies": {"id": "freqs", "type": "Parameter", "tensor": [0.3, 0.2, 0.15, 0.35],
                             "dtype": "torch.float64"},
             "mapping": mapping}
    return {
        "id": "like", "type": "TreeLikelihoodModel", "precision": "float64",
        "tree_model": {"id": "tree", "type": "UnRootedTreeModel", "newick": tiny_newick,
                       "branch_lengths": {"id": "branches", "type": "Parameter", "tensor": [0.0],
                                          "dtype": "torch.float64"},
                       "keep_branch_lengths": True, "taxa": _taxa(tiny_fasta)},
        "site_model": {"id": "sm", "type": "ConstantSiteModel"},
        "substitution_model": subst,
        "site_pattern": {"id": "sp", "type": "SitePattern",
                         "alignment": {"id": "alignment", "type": "Alignment",
                                       "datatype": "nucleotide", "file": tiny_fasta, "taxa": "taxa"}},
    }


def test_json_nonsymmetric_model(gpu, tiny_newick, tiny_fasta):
    """GeneralNonSymmetricSubstitutionModel: P = expm(Q t), pruning on the HIP kernels. With the
    same rates on both triangles it equals GeneralSymmetricSubstitutionModel (P through the
    eigensystem); with different ones, gradients reach rates and frequencies and a rate's
    gradient matches central differences."""
    print("case: JSON non-symmetric model", flush=True)
    from torchtree_amd.treelikelihood_model import TreeLikelihoodModel

    prev_dtype = torch.get_default_dtype()
    torch.set_default_dtype(torch.float64)
    try:
        six = [0.8, 2.1, 0.6, 1.3, 2.4, 0.9]
        sym = TreeLikelihoodModel.from_json(
            _nonsym_graph(tiny_newick, tiny_fasta, "GeneralSymmetricSubstitutionModel", six,
                          list(range(6))), {})()
        non = TreeLikelihoodModel.from_json(
            _nonsym_graph(tiny_newick, tiny_fasta, "GeneralNonSymmetricSubstitutionModel", six,
                          list(range(6)) + list(range(6))), {})()
        assert abs(float(sym) - float(non)) < 1e-9 * abs(float(sym)), (float(sym), float(non))

        dic = {}
        twelve = [0.8, 2.1, 0.6, 1.3, 2.4, 0.9, 1.7, 0.4, 1.1, 0.5, 3.0, 1.2]
        like = TreeLikelihoodModel.from_json(
            _nonsym_graph(tiny_newick, tiny_fasta, "GeneralNonSymmetricSubstitutionModel", twelve,
                          list(range(12))), dic)
        assert abs(float(like()) - float(non)) > 1e-3  # the lower triangle matters
        for k in ("rates", "freqs"):
            dic[k].tensor.requires_grad_(True)
        like().sum().backward()
        g = dic["rates"].tensor.grad.clone()
        assert torch.isfinite(g).all() and torch.isfinite(dic["freqs"].tensor.grad).all()
        base = dic["rates"].tensor.detach().clone()
        eps = 1e-6
        for i in (1, 8):
            e = torch.zeros_like(base)
            e[i] = eps
            dic["rates"].tensor = base + e
            up = float(like())
            dic["rates"].tensor = base - e
            dn = float(like())
            fd = (up - dn) / (2 * eps)
            assert abs(fd - float(g[i])) < 1e-6 * max(1.0, abs(fd)), (i, fd, float(g[i]))
    finally:
        torch.set_default_dtype(prev_dtype)


@pytest.mark.parametrize("kind", ["GeneralSymmetricSubstitutionModel", "GeneralNonSymmetricSubstitutionModel"])
def test_json_attribute_pattern(gpu, tiny_newick, tiny_fasta, kind):
    """Discrete trait (AttributePattern, attribute_pattern.py:13-74): one pattern, each taxon's
    attribute as its state, under a 3-state general model; against the oracle on the same tree,
    branch lengths and P."""
    print(f"case: JSON attribute pattern {kind}", flush=True)
    from oracle import models as omodels
    from oracle import pruning as opruning
    from torchtree_amd.treelikelihood_model import TreeLikelihoodModel

    taxa = _taxa(tiny_fasta)
    regions = ["east", "west", "north"]
    for i, t in enumerate(taxa["taxa"]):
        t["attributes"] = {"region": regions[(i * 2 + 1) % 3]}
    sym = kind.startswith("GeneralSymmetric")
    rates = [0.7, 1.9, 0.4] if sym else [0.7, 1.9, 0.4, 1.2, 0.3, 2.6]
    dtype_spec = {"id": "regions", "type": "GeneralDataType", "codes": regions}
    data = {
        "id": "like", "type": "TreeLikelihoodModel", "precision": "float64",
        "tree_model": {"id": "tree", "type": "UnRootedTreeModel", "newick": tiny_newick,
                       "branch_lengths": {"id": "branches", "type": "Parameter", "tensor": [0.0],
                                          "dtype": "torch.float64"},
                       "keep_branch_lengths": True, "taxa": taxa},
        "site_model": {"id": "sm", "type": "ConstantSiteModel"},
        "substitution_model": {"id": "q", "type": kind, "data_type": dtype_spec,
                               "rates": {"id": "rates", "type": "Parameter", "tensor": rates,
                                         "dtype": "torch.float64"},
                               "frequencies": {"id": "freqs", "type": "Parameter",
                                               "tensor": [0.5, 0.3, 0.2], "dtype": "torch.float64"}},
        "site_pattern": {"id": "sp", "type": "AttributePattern", "taxa": "taxa",
                         "data_type": "regions", "attribute": "region"},
    }
    prev_dtype = torch.get_default_dtype()
    torch.set_default_dtype(torch.float64)
    try:
        dic = {}
        like = TreeLikelihoodModel.from_json(data, dic)
        dic["rates"].tensor.requires_grad_(True)
        out = like()
        out.sum().backward()
        got, grad = float(out), dic["rates"].tensor.grad.clone()
        topo = like.tree_model.topo
        bl = like.tree_model.branch_lengths().detach().reshape(-1)
        bls = torch.cat((bl, torch.zeros(1, dtype=torch.float64)))
        pi = torch.tensor([0.5, 0.3, 0.2], dtype=torch.float64)
        r = torch.tensor(rates, dtype=torch.float64, requires_grad=True)
        if sym:
            q = omodels.general_symmetric_q(r, torch.arange(3), pi)
            mats = omodels.p_t_eigh(q, pi, bls[None, :, None])
        else:
            q = omodels.general_nonsymmetric_q(r, torch.arange(6), pi)
            mats = omodels.p_t_expm(q, pi, bls[None, :, None])
        tips = []
        for t in taxa["taxa"]:
            s = regions.index(t["attributes"]["region"])
            tips.append(torch.nn.functional.one_hot(torch.tensor([s]), 3).to(torch.float64).t())
        post = [tuple(int(x) for x in row) for row in topo.postorder]
        ref = opruning.treelik_discrete_rescaled(tips + [None] * (len(tips) - 1),
                                                 torch.ones(1, dtype=torch.float64), post, mats,
                                                 pi.reshape(1, -1),
                                                 torch.ones(1, 1, 1, dtype=torch.float64))
        assert abs(got - float(ref)) < 1e-10 * abs(float(ref)), (got, float(ref))
        # one pattern (N = 1): the gradient pass reads past the only row of every slab
        ref.sum().backward()
        err = float((grad - r.grad).norm() / r.grad.norm())
        assert err < 1e-9, (err, grad, r.grad)
    finally:
        torch.set_default_dtype(prev_dtype)


class _TorchtreeLikeTree:
    """The interface torchtree's TreeLikelihoodModel uses of a TreeModel (tree_likelihood.py:305-356)."""

    def __init__(self, tree, bl):
        self.postorder = [list(r) for r in tree.postorder]
        self.taxa = list(tree.taxa)
        self._bl = bl
        self.sample_shape = torch.Size([])

    def branch_lengths(self):
        return self._bl


class _TorchtreeLikeJC69:
    sample_shape = torch.Size([])

    def __init__(self):
        self.frequencies = torch.full((4,), 0.25, dtype=torch.float64)

    def q(self):
        q = torch.full((4, 4), 1.0 / 3.0, dtype=torch.float64)
        q.fill_diagonal_(-1.0)
        return q


class _TorchtreeLikeSiteModel:
    sample_shape = torch.Size([])

    def rates(self):
        return torch.ones(1, dtype=torch.float64)

    def probabilities(self):
        return torch.ones(1, dtype=torch.float64)


class _TorchtreeLikeSitePattern:
    def __init__(self, seqs):
        self.seqs = seqs

    def compute_tips_partials(self, use_ambiguities=False):
        return odata.compress_alignment(self.seqs,
                                        lambda c: odata.nucleotide_partial(c, use_ambiguities))

    def compute_tips_states(self):
        return odata.compress_alignment_states(self.seqs, odata.nucleotide_code, 4)


@pytest.mark.parametrize("use_tip_states", [False, True])
def test_json_with_objects_built_by_torchtree(gpu, default_f64, tiny_newick, tiny_fasta,
                                              use_tip_states):
    """Sub-objects resolved by id from ``dic`` that torchtree's own loader would have built
    (stand-ins with the same methods): the tiny golden -83.329016, via from_json_safe."""
    print(f"case: JSON with torchtree-built objects, tip states {use_tip_states}", flush=True)
    from torchtree_amd.treelikelihood_model import JSONParseError, TreeLikelihoodModel

    recs = odata.read_fasta(tiny_fasta)
    taxa = [t for t, _ in recs]
    tree = odata.Tree(tiny_newick, taxa)
    bl = torch.tensor(tree.unrooted_branch_lengths(), dtype=torch.float64, requires_grad=True)
    dic = {"tree": _TorchtreeLikeTree(tree, bl), "jc": _TorchtreeLikeJC69(),
           "sm": _TorchtreeLikeSiteModel(), "sp": _TorchtreeLikeSitePattern([s for _, s in recs])}
    data = {"id": "like", "type": "torchtree_amd.treelikelihood_model.TreeLikelihoodModel",
            "tree_model": "tree", "site_model": "sm", "substitution_model": "jc",
            "site_pattern": "sp", "precision": "float64", "use_tip_states": use_tip_states}
    like = TreeLikelihoodModel.from_json_safe(data, dic)
    out = like()
    assert out.shape == (1,) and out.device == bl.device
    assert abs(float(out) - (-83.329016)) < 1e-5
    out.sum().backward()
    assert bl.grad is not None and torch.isfinite(bl.grad).all()
    with pytest.raises(JSONParseError):
        TreeLikelihoodModel.from_json_safe({"id": "x", "tree_model": "tree"}, dic)


def _sequences(S, T, L, rng):
    letters = "ACGT" if S == 4 else "ACDEFGHIKLMNPQRSTVWY"
    pool = list(letters) * 6 + list("-?N" if S == 4 else "-?XB")
    return ["".join(rng.choice(pool, L)) for _ in range(T)]


@pytest.fixture
def default_f64():
    """torchtree's runner default dtype (torchtree.py:43-48): the model returns it."""
    prev = torch.get_default_dtype()
    torch.set_default_dtype(torch.float64)
    yield
    torch.set_default_dtype(prev)


@pytest.mark.parametrize("model", ["LG", "WAG", "HKY", "EmpiricalSubstitutionModel"])
def test_json_substitution_models(gpu, default_f64, model):
    """LG, WAG (amino_acid.py:12-105), HKY (nucleotide.py:238-264) and a 4-state
    EmpiricalSubstitutionModel (general.py:294-361) through JSON, with a Weibull site model:
    log-likelihood and gradients (branch lengths, Weibull shape, and HKY's kappa and
    frequencies) against the oracle, FP64 relative 1e-12 / 1e-9."""
    print(f"case: JSON substitution model {model}", flush=True)
    import json
    from pathlib import Path

    import numpy as np

    from oracle import models as omodels
    from oracle import pruning as opruning
    from torchtree_amd import synthetic
    from torchtree_amd.treelikelihood_model import TreeLikelihoodModel

    S = 20 if model in ("LG", "WAG") else 4
    rng = np.random.default_rng(len(model))
    T, L = 12, 250
    topo = synthetic.random_topology(T, rng)
    names = [f"taxon{i}" for i in range(T)]
    newick = topo.newick(synthetic.unrooted_lengths(topo, rng, 0.15), names)
    seqs = _sequences(S, T, L, rng)
    kappa = torch.tensor([2.7], dtype=torch.float64)
    pi4 = torch.tensor([0.31, 0.19, 0.22, 0.28], dtype=torch.float64)
    ex4 = torch.tensor([0.8, 2.1, 0.5, 1.1, 3.0, 1.0], dtype=torch.float64)
    if model in ("LG", "WAG"):
        subst = {"id": "m", "type": model}
        fix = json.loads((Path(__file__).parent / "golden" / "empirical_aa.json").read_text())[model]
        rates = torch.tensor(fix["rates"], dtype=torch.float64)
        pi = torch.tensor(fix["frequencies"], dtype=torch.float64)
        oq = lambda: (omodels.q_from_exchangeabilities(rates, pi), pi)  # noqa: E731
        dtype_spec = {"id": "aa", "type": "AminoAcidDataType"}
    elif model == "HKY":
        subst = {"id": "m", "type": "HKY",
                 "kappa": {"id": "kappa", "type": "Parameter", "tensor": kappa.tolist(),
                           "dtype": "torch.float64", "requires_grad": True},
                 "frequencies": {"id": "freqs", "type": "Parameter", "tensor": pi4.tolist(),
                                 "dtype": "torch.float64", "requires_grad": True}}
        k_o, f_o = kappa.clone().requires_grad_(True), pi4.clone().requires_grad_(True)
        oq = lambda: (omodels.hky_q(k_o, f_o), f_o)  # noqa: E731
        dtype_spec = "nucleotide"
    else:
        subst = {"id": "m", "type": "EmpiricalSubstitutionModel", "rates": ex4.tolist(),
                 "frequencies": pi4.tolist()}
        oq = lambda: (omodels.gtr_q(ex4, pi4), pi4)  # noqa: E731
        dtype_spec = "nucleotide"
    taxa = {"id": "taxa", "type": "Taxa", "taxa": [{"id": n, "type": "Taxon"} for n in names]}
    data = {
        "id": "like", "type": "TreeLikelihoodModel", "precision": "float64",
        "tree_model": {"id": "tree", "type": "UnRootedTreeModel", "newick": newick,
                       "branch_lengths": {"id": "branches", "type": "Parameter", "tensor": [0.0],
                                          "dtype": "torch.float64"},
                       "keep_branch_lengths": True, "taxa": taxa},
        "site_model": {"id": "sm", "type": "WeibullSiteModel", "categories": 4,
                       "shape": {"id": "shape", "type": "Parameter", "tensor": [0.7],
                                 "dtype": "torch.float64", "requires_grad": True}},
        "substitution_model": subst,
        "site_pattern": {"id": "sp", "type": "SitePattern",
                         "alignment": {"id": "aln", "type": "Alignment", "datatype": dtype_spec,
                                       "taxa": "taxa",
                                       "sequences": [{"taxon": n, "sequence": s}
                                                     for n, s in zip(names, seqs)]}},
    }
    dic = {}
    like = TreeLikelihoodModel.from_json(data, dic)
    dic["branches"].tensor.requires_grad_(True)
    out = like()
    out.sum().backward()

    tree = odata.Tree(newick, names)
    part = odata.nucleotide_partial if S == 4 else odata.amino_acid_partial
    tips, weights = odata.compress_alignment(seqs, lambda c: part(c, False))
    bl = torch.tensor(tree.unrooted_branch_lengths(), dtype=torch.float64, requires_grad=True)
    shape = torch.tensor([0.7], dtype=torch.float64, requires_grad=True)
    r, p = omodels.weibull_rates_probs(shape, 4)
    q, f = oq()
    bls = torch.cat((bl, torch.zeros(1, dtype=torch.float64)))
    mats = omodels.p_t_eigh(q, f, bls.reshape(-1, 1) * r.reshape(1, -1))
    ref = opruning.treelik_discrete_rescaled(tips + [None] * (T - 1), weights, tree.postorder,
                                             mats, f.reshape(1, -1), p.reshape(-1, 1, 1))
    ref.sum().backward()
    assert abs(float(out) - float(ref)) < 1e-12 * abs(float(ref))
    g = dic["branches"].tensor.grad
    assert float((g - bl.grad).norm() / bl.grad.norm()) < 1e-9
    gs = dic["shape"].tensor.grad
    assert abs(float(gs) - float(shape.grad)) < 1e-9 * abs(float(shape.grad))
    if model == "HKY":
        assert abs(float(dic["kappa"].tensor.grad) - float(k_o.grad)) < 1e-9 * abs(float(k_o.grad))
        assert float((dic["freqs"].tensor.grad - f_o.grad).norm() / f_o.grad.norm()) < 1e-9
