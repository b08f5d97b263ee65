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
