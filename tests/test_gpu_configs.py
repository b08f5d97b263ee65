"""Every benchmark workload at its own shape on the GPU, against the FP64 oracle.

* c1: JC69, 64 taxa, ~1,000 patterns, K = 1, C = 1, full size: log-likelihood and the
  branch-length gradient. The oracle uses the reference's JC69 closed form
  (``nucleotide.py:102-113``); the engine goes through the eigensystem, whose spectrum is
  degenerate (a triple eigenvalue), so the gradient exercises the repeated-eigenvalue limit of
  the divided differences.
* c2: GTR + Weibull4, 512 taxa, K = 32 on a 2,048-pattern slice (256 workgroups of the FP32
  pre-order walk); the oracle evaluates 4 of the 32 samples.
* c3: GTR + Weibull4, 1,024 taxa, K = 16 on a 4,096-pattern slice, in the reference
  harness's unconstrained coordinates (``benchmarks/benchmark.py:919-936``: 5 stick-breaking
  rate coordinates and 3 frequency coordinates); the oracle evaluates 2 of the samples, and FP64
  central differences (step 1e-6) check 20 branch lengths and all 8 model coordinates.
* c4a: 20 states (Gamma(1,1) exchangeabilities, Dirichlet(1) frequencies, gamma4), 256 taxa,
  20,000 patterns, full size.
* c4b: MG94 codons (kappa 2, omega 0.3), gamma4, 128 taxa, 10,000 patterns, full size.
* c5: 23 resident jobs of the serving mix, including its largest shapes: each request's result
  is bitwise the same alone and in the concurrent mix.

Tolerances (FP32 with rescaling against FP64): log-likelihood relative 1e-6, gradients relative
1e-5 norm-wise per parameter group; FP64 against FP64 1e-12 and 1e-10 (1e-9 through the
model coordinates); central differences relative 1e-5 norm-wise, or three times their own
rounding-noise floor where that is larger (a 1e-6 step on a log-likelihood of ~10^6).
"""
import math
import time

import numpy as np
import pytest
import torch

from oracle import models as omodels
from oracle import pruning as opruning

pytestmark = pytest.mark.gpu
F64 = torch.float64


def _flush(msg):
    print(msg, flush=True)


def _rel(a, b):
    a, b = a.detach().cpu().double(), b.detach().cpu().double()
    return float((a - b).norm() / b.norm())


def _post(wl):
    return [tuple(int(x) for x in r) for r in wl.topo.postorder]


def _engine_job(wl, gpu, dtype):
    from torchtree_amd.engine import TreeJob
    from torchtree_amd.likelihood import TipData

    return TreeJob(_post(wl), TipData.from_states(list(wl.codes), wl.S, gpu), wl.weights, wl.S,
                   dtype=dtype, device=gpu)


def _oracle_batched(wl, idx, q_fn, model_params, bls=None):
    """Oracle logL [len(idx)] of samples ``idx`` with autograd; ``model_params`` per sample."""
    b = (wl.bls if bls is None else bls)[idx].clone().requires_grad_(True)
    q, fr = q_fn(*model_params)
    mats = omodels.p_t_eigh(q, fr, b.unsqueeze(-1) * wl.rates.reshape(1, 1, -1))
    ll = opruning.treelik_discrete_rescaled(wl.tip_partials() + [None] * (wl.T - 1), wl.weights,
                                            _post(wl), mats, fr.unsqueeze(-2),
                                            wl.props.reshape(-1, 1, 1)).reshape(-1)
    ll.sum().backward()
    return ll.detach(), b.grad


def test_config1_full_with_gradient(gpu):
    _flush("case: c1 JC69 64 taxa full size, loglik + d/dbls")
    from torchtree_amd import synthetic

    wl = synthetic.config1()
    post = _post(wl)
    b = wl.bls.clone().requires_grad_(True)
    mats = omodels.jc69_p_t(b)  # [1, B, 4, 4], the reference's closed form
    tips = wl.tip_partials()
    ref = opruning.treelik(tips + [None] * (wl.T - 1), wl.weights, post, mats,
                           torch.full((1, 4), 0.25, dtype=F64)).reshape(-1)
    ref.sum().backward()
    for dtype, rl, rg in ((F64, 1e-12, 1e-10), (torch.float32, 1e-6, 1e-5)):
        job = _engine_job(wl, gpu, dtype)
        bd = wl.bls.to(gpu).requires_grad_(True)
        got = job.loglik(bd, wl.rates.reshape(1, -1), wl.props.reshape(1, -1), wl.q_norm, wl.freqs)
        got.sum().backward()
        assert math.isclose(float(got[0].detach()), float(ref[0].detach()), rel_tol=rl), dtype
        assert torch.isfinite(bd.grad).all()
        assert _rel(bd.grad, b.grad) < rg, (dtype, _rel(bd.grad, b.grad))


def _gtr_engine(job, wl, gpu, exch_k, freqs_k):
    from torchtree_amd import models as hm

    bd = wl.bls.to(gpu).requires_grad_(True)
    qn = hm.normalize(hm.gtr_q(exch_k, freqs_k), freqs_k)
    ll = job.loglik(bd, wl.rates.reshape(1, -1), wl.props.reshape(1, -1), qn, freqs_k)
    ll.sum().backward()
    return ll.detach().cpu(), bd.grad.cpu()


def test_config2_512_taxa_k32(gpu):
    _flush("case: c2 GTR+W4 512 taxa K=32 N=2048 fp32 vs oracle")
    from torchtree_amd import synthetic

    wl = synthetic.config2(N=2048)
    assert (wl.T, wl.K, wl.C) == (512, 32, 4)
    K = wl.K
    job = _engine_job(wl, gpu, torch.float32)
    ex = wl.exch.expand(K, 6).clone().requires_grad_(True)
    fr = wl.freqs.expand(K, 4).clone().requires_grad_(True)
    ll, gb = _gtr_engine(job, wl, gpu, ex, fr)
    idx = [0, 9, 21, 31]
    ex_o = wl.exch.expand(len(idx), 6).clone().requires_grad_(True)
    fr_o = wl.freqs.expand(len(idx), 4).clone().requires_grad_(True)
    t0 = time.time()
    ref, rgb = _oracle_batched(wl, idx, lambda e, f: (omodels.gtr_q(e, f), f), (ex_o, fr_o))
    _flush(f"  oracle {time.time() - t0:.1f}s")
    rel = float(((ll[idx] - ref).abs() / ref.abs()).max())
    assert rel < 1e-6, rel
    assert _rel(gb[idx], rgb) < 1e-5
    assert _rel(ex.grad[idx], ex_o.grad) < 1e-5
    assert _rel(fr.grad[idx], fr_o.grad) < 1e-5


def _stick(x):
    return torch.distributions.transforms.StickBreakingTransform()(x)


def test_config3_1024_taxa_k16_and_central_differences(gpu):
    _flush("case: c3 GTR+W4 1024 taxa K=16 N=4096, stick-breaking coordinates, FD")
    from torchtree_amd import models as hm
    from torchtree_amd import synthetic

    wl = synthetic.config3(N=4096)
    assert (wl.T, wl.K, wl.C) == (1024, 16, 4)
    K = wl.K
    sb = torch.distributions.transforms.StickBreakingTransform()
    x_r0 = sb.inv(wl.exch / wl.exch.sum())  # [5]
    x_f0 = sb.inv(wl.freqs)                 # [3]

    def model(job, x_r, x_f, bls):
        rates6, freqs = _stick(x_r), _stick(x_f)
        qn = hm.normalize(hm.gtr_q(rates6, freqs), freqs)
        return job.loglik(bls, wl.rates.reshape(1, -1), wl.props.reshape(1, -1), qn, freqs)

    grads = {}
    for dtype in (torch.float32, F64):
        job = _engine_job(wl, gpu, dtype)
        x_r = x_r0.expand(K, 5).clone().requires_grad_(True)
        x_f = x_f0.expand(K, 3).clone().requires_grad_(True)
        bd = wl.bls.to(gpu).requires_grad_(True)
        ll = model(job, x_r, x_f, bd)
        ll.sum().backward()
        grads[dtype] = (ll.detach().cpu(), bd.grad.cpu(), x_r.grad.clone(), x_f.grad.clone())
    job64 = _engine_job(wl, gpu, F64)

    # oracle on two samples, in the same coordinates
    idx = [0, 15]
    xr_o = x_r0.expand(len(idx), 5).clone().requires_grad_(True)
    xf_o = x_f0.expand(len(idx), 3).clone().requires_grad_(True)
    t0 = time.time()
    ref, rgb = _oracle_batched(
        wl, idx, lambda a, b: (omodels.gtr_q(_stick(a), _stick(b)), _stick(b)), (xr_o, xf_o))
    _flush(f"  oracle {time.time() - t0:.1f}s")
    for dtype, rl, rg in ((torch.float32, 1e-6, 1e-5), (F64, 1e-12, 1e-9)):
        ll, gb, gxr, gxf = grads[dtype]
        rel = float(((ll[idx] - ref).abs() / ref.abs()).max())
        assert rel < rl, (dtype, rel)
        assert _rel(gb[idx], rgb) < rg, (dtype, _rel(gb[idx], rgb))
        assert _rel(gxr[idx], xr_o.grad) < rg, (dtype, _rel(gxr[idx], xr_o.grad))
        assert _rel(gxf[idx], xf_o.grad) < rg, (dtype, _rel(gxf[idx], xf_o.grad))

    # FP64 central differences on the device; every sample is perturbed at once (samples are
    # independent), so each evaluation gives K difference quotients
    h = 1e-6
    rng = np.random.default_rng(0)
    branches = sorted(rng.choice(wl.topo.B - 1, 20, replace=False).tolist())

    def f(x_r, x_f, bls):
        with torch.no_grad():
            return model(job64, x_r, x_f, bls.to(gpu)).cpu()

    xr, xf = x_r0.expand(K, 5).clone(), x_f0.expand(K, 3).clone()
    fd_b = torch.empty(K, len(branches), dtype=F64)
    for j, b in enumerate(branches):
        up, dn = wl.bls.clone(), wl.bls.clone()
        up[:, b] += h
        dn[:, b] -= h
        fd_b[:, j] = (f(xr, xf, up) - f(xr, xf, dn)) / (2 * h)
    fd_r = torch.empty(K, 5, dtype=F64)
    for j in range(5):
        up, dn = xr.clone(), xr.clone()
        up[:, j] += h
        dn[:, j] -= h
        fd_r[:, j] = (f(up, xf, wl.bls) - f(dn, xf, wl.bls)) / (2 * h)
    fd_f = torch.empty(K, 3, dtype=F64)
    for j in range(3):
        up, dn = xf.clone(), xf.clone()
        up[:, j] += h
        dn[:, j] -= h
        fd_f[:, j] = (f(xr, up, wl.bls) - f(xr, dn, wl.bls)) / (2 * h)
    # A central difference carries the rounding of two FP64 evaluations of logL divided by 2h:
    # about 200 ulp of |logL| (the pattern sums) over 2e-6. Its relative accuracy on a group of
    # coordinates is that noise over the group's gradient norm, written out here; the
    # tolerance is the larger of 1e-5 and three times it.
    f0 = f(xr, xf, wl.bls)
    noise = 200 * 2.2e-16 * float(f0.abs().max()) / (2 * h)
    for dtype in (F64, torch.float32):
        _, gb, gxr, gxf = grads[dtype]
        for name, g, fd in (("bls", gb[:, branches], fd_b), ("rates", gxr, fd_r),
                            ("freqs", gxf, fd_f)):
            err = _rel(g, fd)
            tol = max(1e-5, 3 * noise * math.sqrt(fd.numel()) / float(fd.norm()))
            _flush(f"  {dtype} vs central differences, {name}: {err:.2e} (tolerance {tol:.1e}, "
                   f"|logL| {float(f0.abs().max()):.3e})")
            assert err < tol, (dtype, name, err, tol)


def test_config4a_amino_acid_full_size(gpu):
    _flush("case: c4a amino acid 256 taxa 20000 patterns fp32 vs oracle")
    from torchtree_amd import models as hm
    from torchtree_amd import synthetic

    wl = synthetic.make_workload(256, 20_000, 1, model="aa", site="gamma", shape=0.5,
                                 bl_mean=0.05, seed=41)
    assert (wl.T, wl.N, wl.S, wl.C) == (256, 20_000, 20, 4)
    job = _engine_job(wl, gpu, torch.float32)
    ex = wl.exch.clone().requires_grad_(True)
    fr = wl.freqs.clone().requires_grad_(True)
    bd = wl.bls.to(gpu).requires_grad_(True)
    ll = job.loglik(bd, wl.rates.reshape(1, -1), wl.props.reshape(1, -1),
                    hm.normalize(hm.q_from_exchangeabilities(ex, fr), fr), fr)
    ll.sum().backward()
    ex_o = wl.exch.clone().requires_grad_(True)
    fr_o = wl.freqs.clone().requires_grad_(True)
    t0 = time.time()
    ref, rgb = _oracle_batched(wl, [0], lambda e, f: (omodels.q_from_exchangeabilities(e, f), f),
                               (ex_o, fr_o))
    _flush(f"  oracle {time.time() - t0:.1f}s")
    assert math.isclose(float(ll[0].detach()), float(ref[0]), rel_tol=1e-6)
    assert _rel(bd.grad, rgb) < 1e-5
    assert _rel(ex.grad, ex_o.grad) < 1e-5
    assert _rel(fr.grad, fr_o.grad) < 1e-5


def test_config4b_codon_full_size(gpu):
    _flush("case: c4b MG94 128 taxa 10000 patterns fp32 vs oracle")
    from torchtree_amd import models as hm
    from torchtree_amd import synthetic

    wl = synthetic.make_workload(128, 10_000, 1, model="mg94", site="gamma", shape=0.5,
                                 bl_mean=0.05, seed=42)
    assert (wl.T, wl.N, wl.S, wl.C) == (128, 10_000, 61, 4)
    job = _engine_job(wl, gpu, torch.float32)
    abk = torch.tensor([1.0, 0.3, 2.0], dtype=F64)
    p = abk.clone().requires_grad_(True)
    fr = wl.freqs.clone().requires_grad_(True)
    bd = wl.bls.to(gpu).requires_grad_(True)
    qn = hm.normalize(hm.mg94_q(p[0:1], p[1:2], p[2:3], fr), fr)
    ll = job.loglik(bd, wl.rates.reshape(1, -1), wl.props.reshape(1, -1), qn, fr)
    ll.sum().backward()
    p_o = abk.clone().requires_grad_(True)
    fr_o = wl.freqs.clone().requires_grad_(True)
    t0 = time.time()
    ref, rgb = _oracle_batched(
        wl, [0], lambda a, f: (omodels.mg94_q(a[0:1], a[1:2], a[2:3], f), f), (p_o, fr_o))
    _flush(f"  oracle {time.time() - t0:.1f}s")
    assert math.isclose(float(ll[0].detach()), float(ref[0]), rel_tol=1e-6)
    assert _rel(bd.grad, rgb) < 1e-5
    assert _rel(p.grad, p_o.grad) < 1e-5
    assert _rel(fr.grad, fr_o.grad) < 1e-5


def test_config5_mix_alone_equals_mix(gpu):
    """23 resident jobs (20 drawn from the serving mix plus its largest DNA, amino-acid and
    codon shapes), 45 requests: bitwise the same alone and in the concurrent mix."""
    _flush("case: c5 serving mix, 23 jobs, alone == mix")
    from torchtree_amd import serving, synthetic

    rng = np.random.default_rng(44)
    specs = synthetic.serving_mix(20, seed=21)
    specs += [synthetic.random_job_spec("dna", 1024, 100_000, rng, name="dna-max"),
              synthetic.random_job_spec("aa", 1024, 100_000, rng, name="aa-max"),
              synthetic.random_job_spec("codon", 1024, 100_000, rng, name="codon-max")]
    srv = serving.Server(200 << 30, n_streams=8, pool_bytes=160 << 30,
                         device=gpu.index or 0)
    jobs = []
    for sp in specs:
        post = [tuple(int(x) for x in r) for r in sp["topo"].postorder]
        T, N, S, C = sp["topo"].T, sp["codes"].shape[1], sp["S"], len(sp["rates"])
        jb = serving.ResidentJob(sp["name"], post, sp["codes"], sp["table"], sp["weights"],
                                 sp["q_norm"], sp["freqs"], sp["rates"], sp["props"],
                                 dtype=torch.float32,
                                 slice_patterns=srv.slice_for(T, N, C, S, torch.float32),
                                 device=gpu.index or 0, kind=sp["kind"])
        srv.register(sp["name"], jb)
        jobs.append(jb)
    stream = synthetic.request_stream(40, 20, rate=500.0, seed=31)
    # the three largest jobs, with the K their kind allows in a test's time
    stream += [(20, 32, 101, 0.01), (21, 8, 102, 0.02), (22, 1, 103, 0.03)]
    stream.sort(key=lambda r: r[3])

    def requests():
        return [serving.Request(i, jobs[j], synthetic.draw_bls(specs[j], K, s), arrival=t)
                for i, (j, K, s, t) in enumerate(stream)]

    alone = requests()
    t0 = time.time()
    serving.run_sequential(alone, device=gpu.index or 0)
    _flush(f"  alone {time.time() - t0:.1f}s")
    mixed = requests()
    stats = srv.run(mixed)
    _flush(f"  mix {stats}")
    assert stats["requests"] == len(mixed)
    for a, b in zip(alone, mixed):
        assert torch.isfinite(a.loglik).all() and torch.isfinite(a.dbls).all(), a.job.name
        assert torch.equal(a.loglik, b.loglik), a.job.name
        assert torch.equal(a.dbls, b.dbls), a.job.name
