"""HIP kernels against the FP64 oracle on identical seeded inputs.

Tolerances: FP64 kernels relative 1e-12 (log-likelihood) and 1e-10
(gradients); FP32-with-rescaling kernels relative 1e-6 (log-likelihood) and
1e-5 (gradient, norm-wise), the targets of the engine.
"""
import math

import numpy as np
import pytest
import torch

from oracle import data as odata
from oracle import models as omodels
from oracle import pruning as opruning

pytestmark = pytest.mark.gpu

F64 = torch.float64


def _flush(msg):
    print(msg, flush=True)


def oracle_model_loglik(wl, bls, exch, freqs, rates, props, model="gtr"):
    """Oracle logL [K] of a synthetic workload with autograd (reference method)."""
    if model == "gtr":
        q = omodels.gtr_q(exch, freqs)
    else:
        q = omodels.q_from_exchangeabilities(exch, freqs)
    mats = omodels.p_t_eigh(q, freqs, bls.unsqueeze(-1) * rates.reshape(1, 1, -1))
    tips = wl.tip_partials()
    post = [tuple(int(x) for x in r) for r in wl.topo.postorder]
    lp = opruning.treelik_discrete_rescaled(tips + [None] * (wl.T - 1), wl.weights, post, mats,
                                            freqs.reshape(1, -1), props.reshape(-1, 1, 1))
    return lp.reshape(-1)


def test_tiny_dropin_fp64(gpu, tiny_newick, tiny_fasta):
    _flush("case: tiny drop-in calculate_treelikelihood fp64")
    from torchtree_amd import likelihood as hl

    tree = odata.Tree(tiny_newick)
    seqs = dict(odata.read_fasta(tiny_fasta))
    tips, weights = odata.compress_alignment([seqs[t] for t in tree.taxa], odata.nucleotide_partial)
    bl = torch.tensor(tree.edge_lengths(), dtype=F64)
    mats = omodels.jc69_p_t(bl)
    freqs = torch.full((1, 4), 0.25, dtype=F64)
    ref = opruning.treelik(tips + [None] * (tree.T - 1), weights, tree.postorder, mats, freqs)
    got = hl.calculate_treelikelihood(tips + [None] * (tree.T - 1), weights, tree.postorder,
                                      mats.to(gpu), freqs.to(gpu))
    assert got.shape == ref.shape
    assert abs(float(got) - (-83.329016)) < 1e-5
    assert math.isclose(float(got), float(ref), rel_tol=1e-12)


def test_flu_dropin_fp64_and_fp32(gpu, flu_newick, flu_fasta):
    _flush("case: fluA drop-in discrete_rescaled fp64/fp32")
    from torchtree_amd import likelihood as hl
    from tests.test_oracle_pruning import flu_setup

    tree, tips, weights, bl = flu_setup(flu_newick, flu_fasta)
    rates, props = omodels.weibull_rates_probs(torch.tensor([0.1], dtype=F64), 4)
    mats = omodels.jc69_p_t(bl.reshape(-1, 1) * rates.reshape(1, -1))
    freqs = torch.full((1, 4), 0.25, dtype=F64)
    pr = props.reshape(-1, 1, 1)
    partials = tips + [None] * (tree.T - 1)
    ref = opruning.treelik_discrete_rescaled(partials, weights, tree.postorder, mats, freqs, pr)
    got = hl.calculate_treelikelihood_discrete_rescaled(partials, weights, tree.postorder,
                                                        mats.to(gpu), freqs.to(gpu), pr.to(gpu))
    assert math.isclose(float(got), float(ref), rel_tol=1e-12)
    assert abs(float(got) - (-4618.2062529058)) < 4618.2 * 1e-5
    got32 = hl.calculate_treelikelihood_discrete(partials, weights, tree.postorder,
                                                 mats.float().to(gpu), freqs.float().to(gpu),
                                                 pr.float().to(gpu))
    assert math.isclose(float(got32), float(ref), rel_tol=1e-6)


def test_config1_fp64(gpu):
    _flush("case: config1 JC69 64 taxa fp64 via engine")
    from torchtree_amd import synthetic
    from torchtree_amd.engine import TreeJob

    wl = synthetic.config1()
    post = [tuple(int(x) for x in r) for r in wl.topo.postorder]
    job = TreeJob(post, list(wl.codes), wl.weights, 4, dtype=F64, device=gpu, tip_kind="states")
    got = job.loglik(wl.bls.to(gpu), wl.rates.reshape(1, -1), wl.props.reshape(1, -1),
                     wl.q_norm, wl.freqs)
    ref = oracle_model_loglik(wl, wl.bls, wl.exch, wl.freqs, wl.rates, wl.props)
    assert math.isclose(float(got[0]), float(ref[0]), rel_tol=1e-12)
    job32 = TreeJob(post, list(wl.codes), wl.weights, 4, dtype=torch.float32, device=gpu,
                    tip_kind="states")
    got32 = job32.loglik(wl.bls.to(gpu), wl.rates.reshape(1, -1), wl.props.reshape(1, -1),
                         wl.q_norm, wl.freqs)
    assert math.isclose(float(got32[0]), float(ref[0]), rel_tol=1e-6)


@pytest.mark.parametrize("dtype,rtol_l,rtol_g", [(F64, 1e-12, 1e-10), (torch.float32, 1e-6, 1e-5)])
def test_dropin_grad_mats(gpu, dtype, rtol_l, rtol_g):
    _flush(f"case: drop-in d/dmats {dtype}")
    from torchtree_amd import likelihood as hl
    from torchtree_amd import synthetic

    wl = synthetic.make_workload(24, 300, 3, model="gtr", seed=11, bl_mean=0.1, gap_fraction=0.05)
    post = [tuple(int(x) for x in r) for r in wl.topo.postorder]
    tips = wl.tip_partials()
    q = omodels.gtr_q(wl.exch, wl.freqs)
    tau = wl.bls.unsqueeze(-1) * wl.rates.reshape(1, 1, -1)
    mats = omodels.p_t_eigh(q, wl.freqs, tau).detach().requires_grad_(True)
    freqs = wl.freqs.reshape(1, -1).clone().requires_grad_(True)
    props = wl.props.reshape(-1, 1, 1).clone().requires_grad_(True)
    ref = opruning.treelik_discrete_rescaled(tips + [None] * (wl.T - 1), wl.weights, post, mats,
                                             freqs, props)
    ref.sum().backward()
    m2 = mats.detach().to(gpu, dtype).requires_grad_(True)
    f2 = freqs.detach().to(gpu, dtype).requires_grad_(True)
    p2 = props.detach().to(gpu, dtype).requires_grad_(True)
    got = hl.calculate_treelikelihood_discrete(tips + [None] * (wl.T - 1), wl.weights, post, m2, f2, p2)
    assert got.shape == ref.shape
    rel = (got.double().cpu() - ref.detach()).abs() / ref.detach().abs()
    assert float(rel.max()) < rtol_l
    got.sum().backward()
    for a, b in ((m2.grad, mats.grad), (f2.grad, freqs.grad), (p2.grad, props.grad)):
        a = a.double().cpu()
        err = float((a - b).norm() / b.norm())
        assert err < rtol_g, err


@pytest.mark.parametrize("dtype,rtol_l,rtol_g", [(F64, 1e-12, 1e-9), (torch.float32, 1e-6, 1e-5)])
def test_engine_grad_model(gpu, dtype, rtol_l, rtol_g):
    _flush(f"case: engine GTR+W4 full gradient {dtype}")
    from torchtree_amd import models as hm
    from torchtree_amd import synthetic
    from torchtree_amd.engine import TreeJob

    wl = synthetic.make_workload(32, 500, 4, model="gtr", seed=12, bl_mean=0.05)
    post = [tuple(int(x) for x in r) for r in wl.topo.postorder]
    # oracle (reference method, autograd through eigh)
    bls = wl.bls.clone().requires_grad_(True)
    exch = wl.exch.clone().requires_grad_(True)
    freqs = wl.freqs.clone().requires_grad_(True)
    ref = oracle_model_loglik(wl, bls, exch, freqs, wl.rates, wl.props)
    ref.sum().backward()
    # engine
    job = TreeJob(post, list(wl.codes), wl.weights, 4, dtype=dtype, device=gpu, tip_kind="states")
    bls2 = wl.bls.clone().to(gpu).requires_grad_(True)
    exch2 = wl.exch.clone().requires_grad_(True)
    freqs2 = wl.freqs.clone().requires_grad_(True)
    qn = hm.normalize(hm.gtr_q(exch2, freqs2), freqs2)
    got = job.loglik(bls2, wl.rates.reshape(1, -1), wl.props.reshape(1, -1), qn, freqs2)
    rel = (got.cpu() - ref.detach()).abs() / ref.detach().abs()
    assert float(rel.max()) < rtol_l
    got.sum().backward()
    for a, b, name in ((bls2.grad.cpu(), bls.grad, "bls"), (exch2.grad, exch.grad, "exch"),
                       (freqs2.grad, freqs.grad, "freqs")):
        err = float((a - b).norm() / b.norm())
        assert err < rtol_g, (name, err)


# The FP32 pre-order runs as the walk at C = 1, 3 and 4 (dense tips included), by level at
# C = 8 and 16; at (24, 33000) it has ceil(516 / 4) * 2 = 258 workgroups
@pytest.mark.parametrize("T,N,C,dense", [(70, 700, 8, False), (40, 260, 16, False), (30, 333, 4, True),
                                         (24, 33000, 4, False), (24, 33000, 1, False),
                                         (24, 33000, 3, False), (24, 33000, 4, True)])
@pytest.mark.parametrize("dtype,rtol_l,rtol_g", [(F64, 1e-12, 1e-10), (torch.float32, 1e-6, 1e-5)])
def test_dropin_dna_shapes(gpu, T, N, C, dense, dtype, rtol_l, rtol_g):
    """Four-state kernels beyond the default shape: 8 and 16 categories (one-level
    pre-order; at 16 the walk stack no longer fits in LDS and spills to HBM) and
    dense tip partials (more than 256 distinct tip columns)."""
    _flush(f"case: drop-in DNA T={T} N={N} C={C} dense={dense} {dtype}")
    from torchtree_amd import likelihood as hl
    from torchtree_amd import synthetic

    wl = synthetic.make_workload(T, N, 2, categories=C, model="gtr", seed=T + C, bl_mean=0.1)
    post = [tuple(int(x) for x in r) for r in wl.topo.postorder]
    tips = wl.tip_partials()
    if dense:
        rng = np.random.default_rng(5)
        tips = [t * torch.tensor(rng.uniform(0.2, 1.0, t.shape), dtype=t.dtype) + 0.01 for t in tips]
        assert hl.TipData.from_partials(tips, "cpu").kind == hl.HX_TIPS_DENSE
    q = omodels.gtr_q(wl.exch, wl.freqs)
    rates = torch.linspace(0.2, 2.5, C, dtype=F64)
    props = torch.full((C,), 1.0 / C, dtype=F64)
    tau = wl.bls.unsqueeze(-1) * rates.reshape(1, 1, -1)
    mats = omodels.p_t_eigh(q, wl.freqs, tau).detach().requires_grad_(True)
    freqs = wl.freqs.reshape(1, -1).clone().requires_grad_(True)
    pr = props.reshape(-1, 1, 1).clone().requires_grad_(True)
    ref = opruning.treelik_discrete_rescaled(tips + [None] * (wl.T - 1), wl.weights, post, mats,
                                             freqs, pr)
    ref.sum().backward()
    m2 = mats.detach().to(gpu, dtype).requires_grad_(True)
    f2 = freqs.detach().to(gpu, dtype).requires_grad_(True)
    p2 = pr.detach().to(gpu, dtype).requires_grad_(True)
    got = hl.calculate_treelikelihood_discrete(tips + [None] * (wl.T - 1), wl.weights, post, m2, f2, p2)
    rel = (got.double().cpu() - ref.detach()).abs() / ref.detach().abs()
    assert float(rel.max()) < rtol_l
    got.sum().backward()
    for a, b in ((m2.grad, mats.grad), (f2.grad, freqs.grad), (p2.grad, pr.grad)):
        a = a.double().cpu()
        err = float((a - b).norm() / b.norm())
        assert err < rtol_g, err


def _shape_post(shape, T):
    """Post-order triples of a caterpillar or a balanced rooted tree."""
    post, nxt = [], T
    if shape == "caterpillar":
        prev = 0
        for t in range(1, T):
            post.append((nxt, prev, t))
            prev = nxt
            nxt += 1
        return post
    level = list(range(T))
    while len(level) > 1:
        new = []
        for i in range(0, len(level) - 1, 2):
            post.append((nxt, level[i], level[i + 1]))
            new.append(nxt)
            nxt += 1
        if len(level) % 2:
            new.append(level[-1])
        level = new
    return post


@pytest.mark.parametrize("shape", ["caterpillar", "balanced"])
def test_pre_walk_tree_shapes(gpu, shape):
    """The FP32 pre-order walk on trees with no pushes (caterpillar) and with
    the deepest stack (balanced, 32 taxa: depth 4, every push through memory)
    at a walk-sized pattern count, against the FP64 oracle."""
    _flush(f"case: pre-order walk {shape}")
    from torchtree_amd import likelihood as hl

    T, N, K, C = 32, 33000, 2, 4
    rng = np.random.default_rng(9)
    post = _shape_post(shape, T)
    codes = rng.integers(0, 4, size=(T, N))
    tbl = torch.eye(4, dtype=F64)
    tips = [tbl[torch.as_tensor(c)].t().contiguous() for c in codes]
    weights = torch.as_tensor(rng.integers(1, 4, size=N), dtype=F64)
    freqs0 = torch.as_tensor(rng.dirichlet(np.ones(4) * 5), dtype=F64)
    q = omodels.gtr_q(torch.as_tensor(rng.gamma(2.0, 0.5, size=6), dtype=F64), freqs0)
    rates = torch.linspace(0.2, 2.5, C, dtype=F64)
    bls = torch.as_tensor(rng.exponential(0.1, size=(K, 2 * T - 2)), dtype=F64)
    tau = bls.unsqueeze(-1) * rates.reshape(1, 1, -1)
    mats = omodels.p_t_eigh(q, freqs0, tau).detach().requires_grad_(True)
    freqs = freqs0.reshape(1, -1).clone().requires_grad_(True)
    pr = torch.full((C, 1, 1), 1.0 / C, dtype=F64).requires_grad_(True)
    ref = opruning.treelik_discrete_rescaled(tips + [None] * (T - 1), weights, post, mats, freqs, pr)
    ref.sum().backward()
    m2 = mats.detach().to(gpu, torch.float32).requires_grad_(True)
    f2 = freqs.detach().to(gpu, torch.float32).requires_grad_(True)
    p2 = pr.detach().to(gpu, torch.float32).requires_grad_(True)
    got = hl.calculate_treelikelihood_discrete(tips + [None] * (T - 1), weights, post, m2, f2, p2)
    rel = (got.double().cpu() - ref.detach()).abs() / ref.detach().abs()
    assert float(rel.max()) < 1e-6
    got.sum().backward()
    for a, b in ((m2.grad, mats.grad), (f2.grad, freqs.grad), (p2.grad, pr.grad)):
        err = float((a.double().cpu() - b).norm() / b.norm())
        assert err < 1e-5, err


@pytest.mark.parametrize("C", [1, 4])
def test_fp32_gradient_batch_invariant(gpu, C):
    """An FP32 four-state sample's log-likelihood and d logL/dP are bitwise the same evaluated
    alone (K = 1: 8 workgroups) and in a batch of 32 (256 workgroups): every batch size runs
    the same pre-order walk and chunking (the reference asserts batch == single,
    test/test_tree_likelihood.py:263-265)."""
    _flush(f"case: FP32 batch == single C={C}")
    from torchtree_amd import likelihood as hl
    from torchtree_amd import synthetic

    K = 32
    wl = synthetic.make_workload(24, 2048, K, categories=C, model="gtr", seed=11, bl_mean=0.1)
    post = [tuple(int(x) for x in r) for r in wl.topo.postorder]
    tips = wl.tip_partials() + [None] * (wl.T - 1)
    q = omodels.gtr_q(wl.exch, wl.freqs)
    rates = torch.linspace(0.3, 2.0, C, dtype=F64)
    tau = wl.bls.unsqueeze(-1) * rates.reshape(1, 1, -1)
    mats = omodels.p_t_eigh(q, wl.freqs, tau).to(gpu, torch.float32)
    freqs = wl.freqs.reshape(1, -1).to(gpu, torch.float32)
    props = torch.full((C, 1, 1), 1.0 / C, dtype=torch.float32, device=gpu)
    mb = mats.clone().requires_grad_(True)
    lb = hl.calculate_treelikelihood_discrete(tips, wl.weights, post, mb, freqs, props)
    lb.sum().backward()
    for j in (0, 17, K - 1):
        mj = mats[j:j + 1].clone().requires_grad_(True)
        lj = hl.calculate_treelikelihood_discrete(tips, wl.weights, post, mj, freqs, props)
        lj.sum().backward()
        assert torch.equal(lj[0], lb[j]), j
        assert torch.equal(mj.grad[0], mb.grad[j]), j


@pytest.mark.parametrize("T,N,K,C", [(2, 1, 1, 1), (2, 65, 3, 2), (3, 1, 2, 3), (3, 200, 1, 4),
                                     (5, 64, 3, 4)])
@pytest.mark.parametrize("dtype,rtol_l,rtol_g", [(F64, 1e-12, 1e-10), (torch.float32, 1e-6, 1e-5)])
def test_smallest_trees(gpu, T, N, K, C, dtype, rtol_l, rtol_g):
    """Two and three taxa, one pattern, ragged tiles, odd category counts: the whole-tree walks
    with a single entry, a workgroup with one real tile of four, and the padding lanes."""
    _flush(f"case: smallest trees T={T} N={N} K={K} C={C} {dtype}")
    from torchtree_amd import likelihood as hl

    rng = np.random.default_rng(100 * T + N)
    post = [(2, 0, 1)] if T == 2 else ([(3, 0, 1), (4, 3, 2)] if T == 3 else
                                       [(5, 0, 1), (6, 2, 3), (7, 5, 6), (8, 7, 4)])
    B = 2 * T - 2
    codes = rng.integers(0, 5, size=(T, N))  # 4 = gap (all ones)
    table = np.vstack([np.eye(4), np.ones((1, 4))])
    tips = [torch.tensor(table[codes[t]].T.copy(), dtype=F64) for t in range(T)]
    weights = torch.tensor(rng.integers(1, 4, size=N), dtype=F64)
    exch = torch.tensor(rng.gamma(2.0, 2.0, size=6), dtype=F64)
    freqs = torch.tensor(rng.dirichlet(np.full(4, 10.0)), dtype=F64)
    bls = torch.tensor(rng.exponential(0.2, size=(K, B)), dtype=F64)
    rates = torch.linspace(0.4, 1.8, C, dtype=F64)
    q = omodels.gtr_q(exch, freqs)
    mats = omodels.p_t_eigh(q, freqs, bls.unsqueeze(-1) * rates.reshape(1, 1, -1)).detach()
    mats.requires_grad_(True)
    fr = freqs.reshape(1, -1).clone().requires_grad_(True)
    pr = torch.full((C, 1, 1), 1.0 / C, dtype=F64, requires_grad=True)
    ref = opruning.treelik_discrete_rescaled(tips + [None] * (T - 1), weights, post, mats, fr, pr)
    ref.sum().backward()
    m2 = mats.detach().to(gpu, dtype).requires_grad_(True)
    f2 = fr.detach().to(gpu, dtype).requires_grad_(True)
    p2 = pr.detach().to(gpu, dtype).requires_grad_(True)
    got = hl.calculate_treelikelihood_discrete(tips + [None] * (T - 1), weights, post, m2, f2, p2)
    rel = (got.double().cpu() - ref.detach()).abs() / ref.detach().abs()
    assert float(rel.max()) < rtol_l
    got.sum().backward()
    for a, b in ((m2.grad, mats.grad), (f2.grad, fr.grad), (p2.grad, pr.grad)):
        err = float((a.double().cpu() - b).norm() / b.norm())
        assert err < rtol_g, err


@pytest.mark.parametrize("dtype,rtol_l,rtol_g", [(F64, 1e-12, 1e-10), (torch.float32, 1e-6, 1e-5)])
def test_walks_with_a_large_code_table(gpu, dtype, rtol_l, rtol_g):
    """About 200 distinct tip columns (random ambiguity vectors) stay on the tip-code path: the
    walks stage a 200-row code table in LDS ahead of their stacks and transposes."""
    _flush(f"case: large code table {dtype}")
    from torchtree_amd import likelihood as hl
    from torchtree_amd import synthetic

    rng = np.random.default_rng(21)
    T, N, K, C = 40, 3000, 2, 4
    wl = synthetic.make_workload(T, N, K, categories=C, model="gtr", seed=21, bl_mean=0.1)
    post = [tuple(int(x) for x in r) for r in wl.topo.postorder]
    cols = np.vstack([np.eye(4), rng.uniform(0.05, 1.0, size=(196, 4))])  # 200 distinct columns
    codes = rng.integers(0, cols.shape[0], size=(T, wl.N))
    tips = [torch.tensor(cols[codes[t]].T.copy(), dtype=F64) for t in range(T)]
    td = hl.TipData.from_partials(tips, "cpu")
    assert td.kind == hl.HX_TIPS_CODES and 150 <= td.ncodes <= 256
    rates = torch.linspace(0.3, 2.2, C, dtype=F64)
    q = omodels.gtr_q(wl.exch, wl.freqs)
    mats = omodels.p_t_eigh(q, wl.freqs, wl.bls.unsqueeze(-1) * rates.reshape(1, 1, -1)).detach()
    mats.requires_grad_(True)
    fr = wl.freqs.reshape(1, -1).clone().requires_grad_(True)
    pr = torch.full((C, 1, 1), 1.0 / C, dtype=F64, requires_grad=True)
    ref = opruning.treelik_discrete_rescaled(tips + [None] * (T - 1), wl.weights, post, mats, fr, pr)
    ref.sum().backward()
    m2 = mats.detach().to(gpu, dtype).requires_grad_(True)
    f2 = fr.detach().to(gpu, dtype).requires_grad_(True)
    p2 = pr.detach().to(gpu, dtype).requires_grad_(True)
    got = hl.calculate_treelikelihood_discrete(tips + [None] * (T - 1), wl.weights, post, m2, f2, p2)
    rel = (got.double().cpu() - ref.detach()).abs() / ref.detach().abs()
    assert float(rel.max()) < rtol_l
    got.sum().backward()
    for a, b in ((m2.grad, mats.grad), (f2.grad, fr.grad), (p2.grad, pr.grad)):
        err = float((a.double().cpu() - b).norm() / b.norm())
        assert err < rtol_g, err
