"""Dense state spaces (S != 4) on the GPU against the FP64 oracle.

20-state amino-acid and 61-state codon (MG94) models, plus S=3 to exercise
the generic path. Drop-in d/dmats and the engine's full gradient (branch
lengths, exchangeabilities / kappa / omega, frequencies) are compared with
torch autograd through the oracle (the reference method).
"""
import math

import numpy as np
import pytest
import torch

from oracle import models as omodels
from oracle import pruning as opruning

pytestmark = pytest.mark.gpu
F64 = torch.float64


def _random_case(S, T, N, K, C, seed):
    rng = np.random.default_rng(seed)
    from torchtree_amd import synthetic

    topo = synthetic.random_topology(T, rng)
    post = [tuple(int(x) for x in r) for r in topo.postorder]
    codes = rng.integers(0, S + 1, size=(T, N))  # S = unknown
    tbl = torch.cat((torch.eye(S, dtype=F64), torch.ones(1, S, dtype=F64)))
    tips = [tbl[torch.tensor(c)].t().contiguous() for c in codes]
    w = torch.tensor(rng.integers(1, 5, N), dtype=F64)
    pi = torch.tensor(rng.dirichlet(np.ones(S) * 2), dtype=F64)
    ex = torch.tensor(rng.gamma(1.0, 1.0, S * (S - 1) // 2), dtype=F64)
    q = omodels.q_from_exchangeabilities(ex, pi)
    bl = torch.tensor(rng.exponential(0.1, (K, 2 * T - 2)), dtype=F64)
    rates = torch.tensor(rng.gamma(2.0, 0.5, C), dtype=F64)
    props = torch.tensor(rng.dirichlet(np.ones(C)), dtype=F64)
    mats = omodels.p_t_eigh(q, pi, bl.unsqueeze(-1) * rates.reshape(1, 1, -1))
    return post, tips, codes, w, pi, ex, mats, rates, props, bl


@pytest.mark.parametrize("S,C", [(3, 2), (20, 4), (61, 4), (61, 1)])
@pytest.mark.parametrize("dtype,rl,rg", [(F64, 1e-12, 1e-10), (torch.float32, 1e-6, 1e-5)])
def test_dropin_wide(gpu, S, C, dtype, rl, rg):
    print(f"case: wide drop-in S={S} C={C} {dtype}", flush=True)
    from torchtree_amd import likelihood as hl

    T, N, K = 10, 37, 2
    post, tips, codes, w, pi, ex, mats, rates, props, bl = _random_case(S, T, N, K, C, S * 7 + C)
    m = mats.detach().clone().requires_grad_(True)
    f = pi.reshape(1, -1).clone().requires_grad_(True)
    p = props.reshape(-1, 1, 1).clone().requires_grad_(True)
    ref = opruning.treelik_discrete_rescaled(tips + [None] * (T - 1), w, post, m, f, p)
    ref.sum().backward()
    m2 = mats.detach().to(gpu, dtype).requires_grad_(True)
    f2 = pi.reshape(1, -1).to(gpu, dtype).requires_grad_(True)
    p2 = props.reshape(-1, 1, 1).to(gpu, dtype).requires_grad_(True)
    got = hl.calculate_treelikelihood_discrete(tips + [None] * (T - 1), w, post, m2, f2, p2)
    rel = float(((got.detach().double().cpu() - ref.detach()).abs() / ref.detach().abs()).max())
    assert rel < rl, rel
    got.sum().backward()
    for a, b, name in ((m2.grad, m.grad, "mats"), (f2.grad, f.grad, "freqs"), (p2.grad, p.grad, "props")):
        err = float((a.double().cpu() - b).norm() / b.norm())
        assert err < rg, (name, err)


@pytest.mark.parametrize("S,C,T", [(20, 4, 10), (61, 4, 10), (20, 3, 25), (61, 2, 25)])
@pytest.mark.parametrize("dtype,rl,rg", [(F64, 1e-12, 1e-10), (torch.float32, 1e-6, 1e-5)])
def test_dropin_wide_stale_workspace(gpu, S, C, T, dtype, rl, rg, monkeypatch):
    """Every workspace byte starts as 0xFF (NaN in both precisions), so memory the kernels did
    not write cannot reach a result unnoticed. The dense pre-order reads past the rows of the
    last 16-pattern tile (N = 37); those reads land on other partials, on the padding after
    them or on the head of the root's first slab, which the post-order never stores and
    zeroes at its end. T = 10 and 25 put the root at different slots."""
    print(f"case: wide drop-in, NaN-filled workspace S={S} C={C} T={T} {dtype}", flush=True)
    from torchtree_amd import likelihood as hl

    real_empty = torch.empty

    def empty_nan(*args, **kwargs):
        t = real_empty(*args, **kwargs)
        if t.dtype == torch.uint8 and t.is_cuda:
            t.fill_(0xFF)
        return t

    N, K = 37, 2
    post, tips, codes, w, pi, ex, mats, rates, props, bl = _random_case(S, T, N, K, C, S + 3 * T + C)
    m = mats.detach().clone().requires_grad_(True)
    f = pi.reshape(1, -1).clone().requires_grad_(True)
    p = props.reshape(-1, 1, 1).clone().requires_grad_(True)
    ref = opruning.treelik_discrete_rescaled(tips + [None] * (T - 1), w, post, m, f, p)
    ref.sum().backward()
    m2 = mats.detach().to(gpu, dtype).requires_grad_(True)
    f2 = pi.reshape(1, -1).to(gpu, dtype).requires_grad_(True)
    p2 = props.reshape(-1, 1, 1).to(gpu, dtype).requires_grad_(True)
    monkeypatch.setattr(torch, "empty", empty_nan)
    got = hl.calculate_treelikelihood_discrete(tips + [None] * (T - 1), w, post, m2, f2, p2)
    got.sum().backward()
    monkeypatch.setattr(torch, "empty", real_empty)
    rel = float(((got.detach().double().cpu() - ref.detach()).abs() / ref.detach().abs()).max())
    assert rel < rl, rel
    for a, b, name in ((m2.grad, m.grad, "mats"), (f2.grad, f.grad, "freqs"), (p2.grad, p.grad, "props")):
        err = float((a.double().cpu() - b).norm() / b.norm())
        assert err < rg, (name, err)


@pytest.mark.parametrize("S,C", [(4, 4), (20, 2)])
@pytest.mark.parametrize("dtype,rl,rg", [(F64, 1e-12, 1e-10), (torch.float32, 1e-6, 1e-5)])
def test_dropin_nonreversible(gpu, S, C, dtype, rl, rg):
    """Pruning with non-reversible P = expm(Q t) (abstract.py:88-94): the kernels assume no
    symmetry of P; loglik and d/dP against the oracle."""
    print(f"case: non-reversible drop-in S={S} C={C} {dtype}", flush=True)
    from torchtree_amd import likelihood as hl

    T, N, K = 12, 150, 2
    post, tips, codes, w, pi, ex, mats, rates, props, bl = _random_case(S, T, N, K, C, 5 * S + C)
    g = torch.Generator().manual_seed(S + C)
    nonsym = torch.rand(S * (S - 1), generator=g, dtype=F64) * 2 + 0.05
    q = omodels.general_nonsymmetric_q(nonsym, torch.randperm(S * (S - 1), generator=g), pi)
    mats = omodels.p_t_expm(q, pi, bl.unsqueeze(-1) * rates.reshape(1, 1, -1))
    m = mats.detach().clone().requires_grad_(True)
    f = pi.reshape(1, -1).clone().requires_grad_(True)
    p = props.reshape(-1, 1, 1).clone().requires_grad_(True)
    ref = opruning.treelik_discrete_rescaled(tips + [None] * (T - 1), w, post, m, f, p)
    ref.sum().backward()
    m2 = mats.detach().to(gpu, dtype).requires_grad_(True)
    f2 = pi.reshape(1, -1).to(gpu, dtype).requires_grad_(True)
    p2 = props.reshape(-1, 1, 1).to(gpu, dtype).requires_grad_(True)
    got = hl.calculate_treelikelihood_discrete(tips + [None] * (T - 1), w, post, m2, f2, p2)
    rel = float(((got.detach().double().cpu() - ref.detach()).abs() / ref.detach().abs()).max())
    assert rel < rl, rel
    got.sum().backward()
    for a, b, name in ((m2.grad, m.grad, "mats"), (f2.grad, f.grad, "freqs"), (p2.grad, p.grad, "props")):
        err = float((a.double().cpu() - b).norm() / b.norm())
        assert err < rg, (name, err)


@pytest.mark.parametrize("dtype,rl,rg", [(F64, 1e-12, 1e-9), (torch.float32, 1e-6, 1e-5)])
def test_engine_aa(gpu, dtype, rl, rg):
    print(f"case: engine AA S=20 gamma4 {dtype}", flush=True)
    from torchtree_amd import models as hm
    from torchtree_amd import synthetic
    from torchtree_amd.engine import TreeJob

    wl = synthetic.make_workload(16, 120, 2, model="aa", site="gamma", shape=0.7, bl_mean=0.08,
                                 seed=21, gap_fraction=0.05)
    post = [tuple(int(x) for x in r) for r in wl.topo.postorder]
    bls = wl.bls.clone().requires_grad_(True)
    ex = wl.exch.clone().requires_grad_(True)
    fr = wl.freqs.clone().requires_grad_(True)
    q = omodels.q_from_exchangeabilities(ex, fr)
    mats = omodels.p_t_eigh(q, fr, bls.unsqueeze(-1) * wl.rates.reshape(1, 1, -1))
    ref = opruning.treelik_discrete_rescaled(wl.tip_partials() + [None] * (wl.T - 1), wl.weights,
                                             post, mats, fr.reshape(1, -1),
                                             wl.props.reshape(-1, 1, 1)).reshape(-1)
    ref.sum().backward()
    job = TreeJob(post, list(wl.codes), wl.weights, 20, dtype=dtype, device=gpu, tip_kind="states")
    b2 = wl.bls.clone().to(gpu).requires_grad_(True)
    ex2 = wl.exch.clone().requires_grad_(True)
    fr2 = wl.freqs.clone().requires_grad_(True)
    qn = hm.normalize(hm.q_from_exchangeabilities(ex2, fr2), fr2)
    got = job.loglik(b2, wl.rates.reshape(1, -1), wl.props.reshape(1, -1), qn, fr2)
    rel = float(((got.detach().cpu() - ref.detach()).abs() / ref.detach().abs()).max())
    assert rel < rl, rel
    got.sum().backward()
    for a, b, name in ((b2.grad.cpu(), bls.grad, "bls"), (ex2.grad, ex.grad, "exch"), (fr2.grad, fr.grad, "freqs")):
        err = float((a - b).norm() / b.norm())
        assert err < rg, (name, err)


@pytest.mark.parametrize("dtype,rl,rg", [(F64, 1e-12, 1e-9), (torch.float32, 1e-6, 1e-5)])
def test_engine_mg94(gpu, dtype, rl, rg):
    print(f"case: engine MG94 S=61 {dtype}", flush=True)
    from torchtree_amd import models as hm
    from torchtree_amd import synthetic
    from torchtree_amd.engine import TreeJob

    wl = synthetic.make_workload(12, 80, 2, model="mg94", site="gamma", shape=0.5, bl_mean=0.1,
                                 seed=22)
    post = [tuple(int(x) for x in r) for r in wl.topo.postorder]
    pi = wl.freqs
    kappa = torch.tensor([2.0], dtype=F64, requires_grad=True)
    omega = torch.tensor([0.3], dtype=F64, requires_grad=True)
    bls = wl.bls.clone().requires_grad_(True)
    one = torch.ones(1, dtype=F64)
    q = omodels.mg94_q(one, omega, kappa, pi)
    mats = omodels.p_t_eigh(q, pi, bls.unsqueeze(-1) * wl.rates.reshape(1, 1, -1))
    ref = opruning.treelik_discrete_rescaled(wl.tip_partials() + [None] * (wl.T - 1), wl.weights,
                                             post, mats, pi.reshape(1, -1),
                                             wl.props.reshape(-1, 1, 1)).reshape(-1)
    ref.sum().backward()
    job = TreeJob(post, list(wl.codes), wl.weights, 61, dtype=dtype, device=gpu, tip_kind="states")
    b2 = wl.bls.clone().to(gpu).requires_grad_(True)
    k2 = torch.tensor([2.0], dtype=F64, requires_grad=True)
    o2 = torch.tensor([0.3], dtype=F64, requires_grad=True)
    qn = hm.normalize(hm.mg94_q(one, o2, k2, pi), pi)
    got = job.loglik(b2, wl.rates.reshape(1, -1), wl.props.reshape(1, -1), qn, pi)
    rel = float(((got.detach().cpu() - ref.detach()).abs() / ref.detach().abs()).max())
    assert rel < rl, rel
    got.sum().backward()
    assert float((b2.grad.cpu() - bls.grad).norm() / bls.grad.norm()) < rg
    assert abs(float(k2.grad) - float(kappa.grad)) < rg * abs(float(kappa.grad)) + 1e-9
    assert abs(float(o2.grad) - float(omega.grad)) < rg * abs(float(omega.grad)) + 1e-9


def test_mfma_lane_layout(gpu):
    """The MFMA operand/result lane layouts the kernels assume (16x16x4 FP32/FP64, BF16, 4x4x1)."""
    print("case: mfma layout selftest", flush=True)
    import ctypes

    from torchtree_amd import _lib

    lib = _lib.load()
    m = ctypes.c_int(-1)
    _lib.check(lib.hx_selftest_mfma_layout(ctypes.byref(m), _lib.stream_handle(None)))
    torch.cuda.synchronize()
    assert m.value == 0


@pytest.mark.parametrize("S,C,N", [(20, 3, 1000), (61, 2, 333), (17, 1, 515), (33, 4, 70)])
def test_dropin_wide_many_tiles(gpu, S, C, N):
    """FP32 dense kernels over many pattern tiles and chunks, ragged last tile."""
    print(f"case: wide many tiles S={S} C={C} N={N}", flush=True)
    from torchtree_amd import likelihood as hl

    T, K = 12, 3
    post, tips, codes, w, pi, ex, mats, rates, props, bl = _random_case(S, T, N, K, C, S + N + C)
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
    assert rel < 2e-6, rel
    got.sum().backward()
    for a, b in ((m2, m), (f2, f), (p2, p)):
        ga, gb = a.grad.double().cpu(), b.grad
        err = float((ga - gb).norm() / gb.norm())
        assert err < 2e-5, err
