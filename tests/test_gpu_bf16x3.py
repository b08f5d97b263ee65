"""Dense-state products in split BF16 precision (HX_F32_BF16X3) against the FP64 oracle.

Each FP32 operand is carried as hi = bf16(x) plus lo = bf16(x - hi) and each product formed as
three BF16 matrix-core products with FP32 accumulation. The partials stay FP32 in HBM with the
same power-of-two rescaling, so the errors stated for FP32 apply: log-likelihood relative
1e-6, gradients relative 1e-5 norm-wise (2e-6 / 2e-5 over many tiles, as for FP32). Covered:
the drop-in with tip codes and dense tips at S = 17, 20, 33, 61 and C = 1..4, many tiles with a
ragged last tile, and the amino-acid and codon workloads at their full sizes through the
engine (branch lengths, model parameters, frequencies).
"""
import math

import pytest
import torch

from oracle import models as omodels
from oracle import pruning as opruning
from tests.test_gpu_wide import _random_case

pytestmark = pytest.mark.gpu
F64 = torch.float64


@pytest.fixture
def split_products():
    from torchtree_amd import likelihood as hl

    prev = hl.set_dense_products("bf16x3")
    yield
    hl.set_dense_products(prev)


def _rel(a, b):
    a, b = a.detach().double().cpu(), b.detach()
    return float((a - b).norm() / b.norm())


@pytest.mark.parametrize("S,C,N,T,dense", [(20, 4, 37, 10, False), (61, 4, 37, 10, False),
                                           (61, 1, 200, 14, False), (17, 1, 515, 12, False),
                                           (20, 3, 1000, 12, True), (33, 4, 70, 12, False),
                                           (61, 2, 333, 12, True)])
def test_split_dropin(gpu, split_products, S, C, N, T, dense):
    print(f"case: bf16x3 drop-in S={S} C={C} N={N} dense={dense}", flush=True)
    from torchtree_amd import likelihood as hl

    K = 2
    post, tips, codes, w, pi, ex, mats, rates, props, bl = _random_case(S, T, N, K, C, 3 * S + N)
    if dense:  # non-code tips: a partial with three states set
        tips = [t.clone() for t in tips]
        for t in tips:
            t[: min(3, S), ::5] = 1.0
    m = mats.detach().clone().requires_grad_(True)
    f = pi.reshape(1, -1).clone().requires_grad_(True)
    p = props.reshape(-1, 1, 1).clone().requires_grad_(True)
    ref = opruning.treelik_discrete_rescaled(tips + [None] * (T - 1), w, post, m, f, p)
    ref.sum().backward()
    m2 = mats.detach().to(gpu, torch.float32).requires_grad_(True)
    f2 = pi.reshape(1, -1).to(gpu, torch.float32).requires_grad_(True)
    p2 = props.reshape(-1, 1, 1).to(gpu, torch.float32).requires_grad_(True)
    got = hl.calculate_treelikelihood_discrete(tips + [None] * (T - 1), w, post, m2, f2, p2)
    rel = float(((got.detach().double().cpu() - ref.detach()).abs() / ref.detach().abs()).max())
    many = N > 64
    assert rel < (2e-6 if many else 1e-6), rel
    got.sum().backward()
    for a, b, name in ((m2.grad, m.grad, "mats"), (f2.grad, f.grad, "freqs"),
                       (p2.grad, p.grad, "props")):
        err = _rel(a, b)
        assert err < (2e-5 if many else 1e-5), (name, err)


def _workload(kind):
    from torchtree_amd import synthetic

    if kind == "aa":
        return synthetic.make_workload(256, 20_000, 1, model="aa", site="gamma", shape=0.5,
                                       bl_mean=0.05, seed=41)
    return synthetic.make_workload(128, 10_000, 1, model="mg94", site="gamma", shape=0.5,
                                   bl_mean=0.05, seed=42)


@pytest.mark.parametrize("kind", ["aa", "codon"])
def test_split_engine_full_size(gpu, kind):
    """The amino-acid (256 taxa x 20,000 patterns) and codon (128 x 10,000) workloads with
    split products, against the FP64 oracle: logL 1e-6, gradients 1e-5."""
    print(f"case: bf16x3 engine full size {kind}", flush=True)
    from torchtree_amd import models as hm
    from torchtree_amd.engine import TreeJob
    from torchtree_amd.likelihood import TipData

    wl = _workload(kind)
    post = [tuple(int(x) for x in r) for r in wl.topo.postorder]
    job = TreeJob(post, TipData.from_states(list(wl.codes), wl.S, gpu), wl.weights, wl.S,
                  dtype=torch.float32, device=gpu, products="bf16x3")
    if kind == "aa":
        par0 = wl.exch.clone()
        qf = lambda x, fr: hm.q_from_exchangeabilities(x, fr)  # noqa: E731
        oqf = lambda x, fr: omodels.q_from_exchangeabilities(x, fr)  # noqa: E731
    else:
        par0 = torch.tensor([1.0, 0.3, 2.0], dtype=F64)
        qf = lambda x, fr: hm.mg94_q(x[0:1], x[1:2], x[2:3], fr)  # noqa: E731
        oqf = lambda x, fr: omodels.mg94_q(x[0:1], x[1:2], x[2:3], fr)  # noqa: E731
    x = par0.clone().requires_grad_(True)
    fr = wl.freqs.clone().requires_grad_(True)
    bd = wl.bls.to(gpu).requires_grad_(True)
    ll = job.loglik(bd, wl.rates.reshape(1, -1), wl.props.reshape(1, -1),
                    hm.normalize(qf(x, fr), fr), fr)
    ll.sum().backward()
    xo = par0.clone().requires_grad_(True)
    fo = wl.freqs.clone().requires_grad_(True)
    bo = wl.bls.clone().requires_grad_(True)
    mats = omodels.p_t_eigh(oqf(xo, fo), fo, bo.unsqueeze(-1) * wl.rates.reshape(1, 1, -1))
    ref = opruning.treelik_discrete_rescaled(wl.tip_partials() + [None] * (wl.T - 1), wl.weights,
                                             post, mats, fo.reshape(1, -1),
                                             wl.props.reshape(-1, 1, 1)).reshape(-1)
    ref.sum().backward()
    rel = abs(float(ll[0].detach()) - float(ref[0].detach())) / abs(float(ref[0].detach()))
    print(f"  logL rel err {rel:.2e}, bls {_rel(bd.grad, bo.grad):.2e}, "
          f"model {_rel(x.grad, xo.grad):.2e}, freqs {_rel(fr.grad, fo.grad):.2e}", flush=True)
    assert rel < 1e-6, rel
    assert _rel(bd.grad, bo.grad) < 1e-5
    assert _rel(x.grad, xo.grad) < 1e-5
    assert _rel(fr.grad, fo.grad) < 1e-5
    assert math.isfinite(float(ll[0].detach()))
