"""Served requests on the GPU: results are bitwise the same alone and in a
concurrent mix, pattern slicing gives the unsliced answer, and an FP64
job matches the oracle."""
import numpy as np
import pytest
import torch

from oracle import models as omodels
from oracle import pruning as opruning

pytestmark = pytest.mark.gpu
F64 = torch.float64


def _jobs(gpu, dtype, slice_for=None):
    from torchtree_amd import serving, synthetic

    rng = np.random.default_rng(3)
    specs = [synthetic.random_job_spec(k, T, N, rng, name=f"{k}{T}")
             for k, T, N in (("dna", 40, 700), ("dna", 17, 150), ("aa", 24, 300),
                             ("codon", 12, 90), ("dna", 90, 2500))]
    jobs = []
    for sp in specs:
        post = [tuple(int(x) for x in r) for r in sp["topo"].postorder]
        sl = slice_for(sp) if slice_for else 0
        jobs.append(serving.ResidentJob(sp["name"], post, sp["codes"], sp["table"], sp["weights"],
                                        sp["q_norm"], sp["freqs"], sp["rates"], sp["props"],
                                        dtype=dtype, slice_patterns=sl, device=gpu.index or 0,
                                        kind=sp["kind"]))
    return specs, jobs


def test_alone_equals_mix_and_slicing(gpu):
    print("case: serving alone == mix", flush=True)
    from torchtree_amd import serving, synthetic

    specs, jobs = _jobs(gpu, torch.float32)
    stream = synthetic.request_stream(30, len(jobs), rate=2000.0, seed=9, ks=(1, 3, 8))
    def requests():
        return [serving.Request(i, jobs[j], synthetic.draw_bls(specs[j], K, s), arrival=t)
                for i, (j, K, s, t) in enumerate(stream)]
    alone = requests()
    serving.run_sequential(alone, device=gpu.index or 0)
    srv = serving.Server(32 << 30, n_streams=4, pool_bytes=16 << 30, device=gpu.index or 0)
    for jb in jobs:
        srv.register(jb.name, jb)
    mixed = requests()
    stats = srv.run(mixed)
    assert stats["requests"] == len(mixed)
    for a, b in zip(alone, mixed):
        assert torch.equal(a.loglik, b.loglik), a.rid
        assert torch.equal(a.dbls, b.dbls), a.rid
    # pattern slices of 64: per-pattern log-likelihoods are unchanged and summed in
    # FP64, so logL agrees to reassociation; gradients accumulate FP32 partial sums
    # per chunk, so they agree to FP32 accuracy (1e-5, the engine's FP32 target)
    specs2, sliced = _jobs(gpu, torch.float32, slice_for=lambda sp: 64)
    again = [serving.Request(r.rid, sliced[jobs.index(r.job)], r.bls) for r in alone]
    serving.run_sequential(again, device=gpu.index or 0)
    for a, b in zip(alone, again):
        assert torch.allclose(a.loglik, b.loglik, rtol=1e-12, atol=0)
        assert float((a.dbls - b.dbls).norm() / a.dbls.norm()) < 1e-5


@pytest.mark.parametrize("dtype,rl,rg", [(F64, 1e-12, 1e-9), (torch.float32, 1e-6, 1e-5)])
def test_served_matches_oracle(gpu, dtype, rl, rg):
    print(f"case: served vs oracle {dtype}", flush=True)
    from torchtree_amd import serving, synthetic

    specs, jobs = _jobs(gpu, dtype)
    for sp, jb in zip(specs, jobs):
        bls = synthetic.draw_bls(sp, 2, 11)
        req = serving.Request(0, jb, bls)
        serving.run_sequential([req], device=gpu.index or 0)
        b = bls.clone().requires_grad_(True)
        tau = b.unsqueeze(-1) * torch.as_tensor(sp["rates"]).reshape(1, 1, -1)
        mats = omodels.p_t_eigh(sp["q_norm"], torch.as_tensor(sp["freqs"]), tau)
        tbl = torch.as_tensor(sp["table"])
        tips = [tbl[torch.as_tensor(c).long()].t().contiguous() for c in sp["codes"]]
        post = [tuple(int(x) for x in r) for r in sp["topo"].postorder]
        ref = opruning.treelik_discrete_rescaled(tips + [None] * (len(tips) - 1),
                                                 torch.as_tensor(sp["weights"]), post, mats,
                                                 torch.as_tensor(sp["freqs"]).reshape(1, -1),
                                                 torch.as_tensor(sp["props"]).reshape(-1, 1, 1)).reshape(-1)
        ref.sum().backward()
        rel = float(((req.loglik.cpu() - ref.detach()).abs() / ref.detach().abs()).max())
        assert rel < rl, (sp["name"], rel)
        g = b.grad[:, :-1]  # slot 2T-3 is a padded (fixed) length
        err = float((req.dbls.cpu()[:, :-1] - g).norm() / g.norm())
        assert err < rg, (sp["name"], err)


def test_served_walk_matches_oracle(gpu):
    """A DNA request large enough for the FP32 pre-order walk (d loglik / d tau
    variant: 24 taxa, 33,000 patterns, K = 2, 258 workgroups) against the oracle."""
    print("case: served DNA walk vs oracle", flush=True)
    from torchtree_amd import serving, synthetic

    rng = np.random.default_rng(77)
    sp = synthetic.random_job_spec("dna", 24, 33000, rng)
    post = [tuple(int(x) for x in r) for r in sp["topo"].postorder]
    jb = serving.ResidentJob("walk", post, sp["codes"], sp["table"], sp["weights"], sp["q_norm"],
                             sp["freqs"], sp["rates"], sp["props"], dtype=torch.float32,
                             device=gpu.index or 0, kind="dna")
    bls = synthetic.draw_bls(sp, 2, 13)
    req = serving.Request(0, jb, bls)
    serving.run_sequential([req], device=gpu.index or 0)
    b = bls.clone().requires_grad_(True)
    tau = b.unsqueeze(-1) * torch.as_tensor(sp["rates"]).reshape(1, 1, -1)
    mats = omodels.p_t_eigh(sp["q_norm"], torch.as_tensor(sp["freqs"]), tau)
    tbl = torch.as_tensor(sp["table"])
    tips = [tbl[torch.as_tensor(c).long()].t().contiguous() for c in sp["codes"]]
    ref = opruning.treelik_discrete_rescaled(tips + [None] * (len(tips) - 1),
                                             torch.as_tensor(sp["weights"]), post, mats,
                                             torch.as_tensor(sp["freqs"]).reshape(1, -1),
                                             torch.as_tensor(sp["props"]).reshape(-1, 1, 1)).reshape(-1)
    ref.sum().backward()
    rel = float(((req.loglik.cpu() - ref.detach()).abs() / ref.detach().abs()).max())
    assert rel < 1e-6, rel
    g = b.grad[:, :-1]  # slot 2T-3 is a padded (fixed) length
    err = float((req.dbls.cpu()[:, :-1] - g).norm() / g.norm())
    assert err < 1e-5, err


@pytest.mark.parametrize("dtype", [torch.float32, F64])
def test_workspace_size_without_a_job(gpu, dtype):
    """hx_job_workspace_bytes_for (host only, used to choose pattern slices before a job
    exists) gives exactly what the registered job asks for, at every state count and K."""
    print(f"case: job workspace from its shape {dtype}", flush=True)
    from torchtree_amd import _lib

    lib = _lib.load()
    code = _lib.HX_F64 if dtype == F64 else _lib.HX_F32
    specs, jobs = _jobs(gpu, dtype, slice_for=lambda sp: 100)
    for sp, job in zip(specs, jobs):
        for K in (1, 8, 32):
            for grad in (0, 1):
                got = int(lib.hx_job_workspace_bytes_for(job.T, job.S, job.slice, job.C, K, code, grad))
                assert got == job.workspace_bytes(K, bool(grad)), (sp["name"], K, grad)


@pytest.mark.parametrize("dtype,rl,rg", [(F64, 1e-12, 1e-9), (torch.float32, 1e-6, 1e-5)])
@pytest.mark.parametrize("batched", [False, True])
def test_served_model_request_matches_oracle(gpu, dtype, rl, rg, batched):
    """A resident DNA job evaluated under each request's own GTR model (hx_job_eval_model),
    with gradients for the exchangeabilities, frequencies, Weibull shape and branch lengths by
    autograd through the served d/dQ, d/dpi, d/drates, d/dprops; against the oracle. Batched:
    one model per sample; otherwise one model for all samples. The job is registered with
    another model and two pattern slices, which the request's model must override."""
    print(f"case: served per-request model {dtype} batched={batched}", flush=True)
    from torchtree_amd import models as hm
    from torchtree_amd import serving, synthetic

    rng = np.random.default_rng(91)
    sp = synthetic.random_job_spec("dna", 30, 900, rng)
    post = [tuple(int(x) for x in r) for r in sp["topo"].postorder]
    job = serving.ResidentJob("m", post, sp["codes"], sp["table"], sp["weights"], sp["q_norm"],
                              sp["freqs"], sp["rates"], sp["props"], dtype=dtype,
                              slice_patterns=500, device=gpu.index or 0)
    K = 3
    bls = synthetic.draw_bls(sp, K, 4)
    lead = (K,) if batched else ()
    g = torch.Generator().manual_seed(3)
    exch0 = torch.rand(lead + (6,), generator=g, dtype=F64) * 2 + 0.3
    fr0 = torch.rand(lead + (4,), generator=g, dtype=F64) + 0.5
    fr0 = fr0 / fr0.sum(-1, keepdim=True)
    shape0 = torch.tensor([0.8], dtype=F64)

    def run(engine):
        ex = exch0.clone().requires_grad_(True)
        fr = fr0.clone().requires_grad_(True)
        sh = shape0.clone().requires_grad_(True)
        b = bls.clone().requires_grad_(True)
        if engine:
            rates, props = hm.weibull_rates(sh, 4)
            qn = hm.normalize(hm.gtr_q(ex, fr), fr)
            ll = serving.served_loglik(job, b.to(gpu), rates, props, qn, fr)
        else:
            rates, props = omodels.weibull_rates_probs(sh, 4)
            q = omodels.gtr_q(ex, fr)
            mats = omodels.p_t_eigh(q, fr, b.unsqueeze(-1) * rates.reshape(1, 1, -1))
            tbl = torch.as_tensor(sp["table"])
            tips = [tbl[torch.as_tensor(c).long()].t().contiguous() for c in sp["codes"]]
            ll = opruning.treelik_discrete_rescaled(
                tips + [None] * (len(tips) - 1), torch.as_tensor(sp["weights"]), post, mats,
                fr.reshape(-1, 1, 4), props.reshape(-1, 1, 1)).reshape(-1)
        (ll * torch.arange(1, K + 1, dtype=F64, device=ll.device)).sum().backward()
        return ll.detach().cpu(), b.grad, ex.grad, fr.grad, sh.grad

    got, ref = run(True), run(False)
    rel = float(((got[0] - ref[0]).abs() / ref[0].abs()).max())
    assert rel < rl, rel
    for name, a, b in zip(("bls", "exch", "freqs", "shape"), got[1:], ref[1:]):
        a = a.detach().cpu()
        b2 = b[:, :-1] if name == "bls" else b
        a2 = a[:, :-1] if name == "bls" else a
        err = float((a2 - b2).norm() / b2.norm())
        assert err < rg, (name, err)


def test_graph_replays_match_eager(gpu):
    """Small requests served by replaying captured HIP graphs (one per job, K and gradient
    flag): every replay gives the eager result bit for bit. The server runs three times over
    the same graphs, with eager requests of larger jobs in the mix and the program's own eager
    work and host synchronisations between the runs."""
    print("case: serving graph replays == eager", flush=True)
    from torchtree_amd import serving, synthetic

    specs, jobs = _jobs(gpu, torch.float32)
    K_of = [1, 3, 2, 1, 3]
    sizes = sorted(jb.workspace_bytes(K_of[i], True) for i, jb in enumerate(jobs))
    limit = sizes[2]  # the three smallest are graphed, the other two run eagerly
    srv = serving.Server(32 << 30, n_streams=4, pool_bytes=16 << 30, device=gpu.index or 0,
                         graph_max_bytes=limit)
    for jb in jobs:
        srv.register(jb.name, jb)
    x = torch.randn(512, 512, device=gpu)
    graphed = 0
    for rep in range(3):
        reqs = [serving.Request(i, jobs[j], synthetic.draw_bls(specs[j], K_of[j], 100 * rep + i),
                                arrival=0.0005 * i)
                for i, j in enumerate([0, 1, 2, 3, 4] * 4)]
        stats = srv.run(reqs)
        assert stats["requests"] == len(reqs)
        graphed += stats["graphed"]
        ref = [serving.Request(r.rid, r.job, r.bls) for r in reqs]
        serving.run_sequential(ref, device=gpu.index or 0)
        for a, b in zip(reqs, ref):
            assert torch.equal(a.loglik, b.loglik), (rep, a.rid, a.graphed)
            assert torch.equal(a.dbls, b.dbls), (rep, a.rid, a.graphed)
        # eager work of the program itself between the runs
        x = torch.tanh(x @ x.t() / 512)
        torch.cuda.synchronize(gpu)
    assert graphed >= 6, graphed
    assert len(srv.executor.graphs) >= 3
    # unregistering a job drops its graphs and gives their workspace back
    n_before = len(srv.executor.graphs)
    in_use = srv.pool.in_use
    small = min(range(len(jobs)), key=lambda i: jobs[i].workspace_bytes(K_of[i], True))
    srv.unregister(jobs[small].name)
    assert len(srv.executor.graphs) < n_before and srv.pool.in_use < in_use


def test_server_model_requests(gpu):
    """Requests that carry their own GTR model (exchangeabilities, frequencies, Weibull rates)
    go through the scheduler with plain requests in the mix; each equals the same request run
    alone (hx_job_eval_model), bit for bit, including d/dQ, d/dfreqs, d/drates and d/dprops."""
    print("case: server with per-request models", flush=True)
    from torchtree_amd import models as hm
    from torchtree_amd import serving, synthetic

    specs, jobs = _jobs(gpu, torch.float32)
    dev = torch.device("cuda", gpu.index or 0)
    g = torch.Generator().manual_seed(5)

    def model(K, batched, C):
        lead = (K,) if batched else (1,)
        ex = torch.rand(lead + (6,), generator=g, dtype=F64) + 0.2
        fr = torch.rand(lead + (4,), generator=g, dtype=F64) + 0.5
        fr = fr / fr.sum(-1, keepdim=True)
        qn = hm.normalize(hm.gtr_q(ex, fr), fr)
        rates, props = hm.weibull_rates(torch.tensor([0.7], dtype=F64), C)
        R = lead[0]
        eig = hm.Eigen(qn if batched else qn[0], fr if batched else fr[0], dev)
        return {"eig": eig, "rates": rates.reshape(1, C).expand(R, C).contiguous().to(dev),
                "props": props.reshape(1, C).expand(R, C).contiguous().to(dev),
                "freqs": fr.to(dev).contiguous()}

    dna = [i for i, sp in enumerate(specs) if sp["kind"] == "dna"]
    reqs = []
    for i in range(16):
        j = dna[i % len(dna)] if i % 2 == 0 else i % len(jobs)
        K = 1 + i % 3
        m = model(K, (i % 4 == 0), jobs[j].C) if i % 2 == 0 else None
        reqs.append(serving.Request(i, jobs[j], synthetic.draw_bls(specs[j], K, i),
                                    arrival=0.0003 * i, model=m))
    srv = serving.Server(32 << 30, n_streams=4, pool_bytes=16 << 30, device=gpu.index or 0)
    for jb in jobs:
        srv.register(jb.name, jb)
    stats = srv.run(reqs)
    assert stats["requests"] == len(reqs)
    ref = [serving.Request(r.rid, r.job, r.bls, model=r.model) for r in reqs]
    serving.run_sequential(ref, device=gpu.index or 0)
    for a, b in zip(reqs, ref):
        assert torch.isfinite(a.loglik).all()
        assert torch.equal(a.loglik, b.loglik), a.rid
        assert torch.equal(a.dbls, b.dbls), a.rid
        if a.model is not None:
            for key in ("drates", "dq", "dfreqs", "dprops"):
                assert torch.equal(a.grads[key], b.grads[key]), (a.rid, key)


@pytest.mark.parametrize("dtype,rl,rg", [(F64, 1e-12, 1e-8), (torch.float32, 1e-6, 1e-5)])
@pytest.mark.parametrize("kind,S", [("dna", 4), ("aa", 20)])
def test_served_general_model_matches_oracle(gpu, dtype, rl, rg, kind, S):
    """A resident job with a non-reversible rate matrix (hx_job_create_general: P = exp(Q tau)
    per request) in two pattern slices, against the oracle with P from torch.matrix_exp in FP64
    (NonSymmetricSubstitutionModel.p_t, abstract.py:88-94) and autograd for d/dbls; then in a
    server mix with reversible jobs, bitwise the same alone and in the mix. The FP64 gradient
    bound is 1e-8: the engine forms dP/dtau as Q P, whose rows cancel (Q's rows sum to zero),
    against autograd through the block exponential (2e-9 measured at 20 states)."""
    print(f"case: served general Q {kind} {dtype}", flush=True)
    from torchtree_amd import serving, synthetic

    rng = np.random.default_rng(5 + S)
    sp = synthetic.random_job_spec(kind, 18, 700, rng)
    post = [tuple(int(x) for x in r) for r in sp["topo"].postorder]
    pi = torch.as_tensor(sp["freqs"], dtype=F64)
    exch = torch.tensor(rng.gamma(1.5, 1.0, S * (S - 1)), dtype=F64)
    q = omodels.normalize_q(omodels.general_nonsymmetric_q(exch, torch.arange(S * (S - 1)), pi), pi)
    jb = serving.ResidentJob("gen", post, sp["codes"], sp["table"], sp["weights"], q, pi,
                             sp["rates"], sp["props"], dtype=dtype, slice_patterns=400,
                             device=gpu.index or 0, kind=kind, general=True)
    K = 3
    bls = synthetic.draw_bls(sp, K, 21)
    req = serving.Request(0, jb, bls)
    serving.run_sequential([req], device=gpu.index or 0)
    b = bls.clone().requires_grad_(True)
    tau = b.unsqueeze(-1) * torch.as_tensor(sp["rates"], dtype=F64).reshape(1, 1, -1)
    mats = torch.matrix_exp(q * tau[..., None, None])
    tbl = torch.as_tensor(sp["table"])
    tips = [tbl[torch.as_tensor(c).long()].t().contiguous() for c in sp["codes"]]
    ref = opruning.treelik_discrete_rescaled(tips + [None] * (len(tips) - 1),
                                             torch.as_tensor(sp["weights"]), post, mats,
                                             pi.reshape(1, -1),
                                             torch.as_tensor(sp["props"]).reshape(-1, 1, 1)).reshape(-1)
    ref.sum().backward()
    rel = float(((req.loglik.cpu() - ref.detach()).abs() / ref.detach().abs()).max())
    assert rel < rl, rel
    g = b.grad[:, :-1]  # slot 2T-3 is a padded (fixed) length
    err = float((req.dbls.cpu()[:, :-1] - g).norm() / g.norm())
    assert err < rg, err
    # a per-request eigensystem is refused for such a job
    from torchtree_amd import models as hm
    with pytest.raises(serving.HipTreeError):
        jb.eval_model(bls.to(gpu), hm.Eigen(sp["q_norm"], pi, gpu),
                      torch.ones(1, jb.C, dtype=F64, device=gpu),
                      torch.full((1, jb.C), 1.0 / jb.C, dtype=F64, device=gpu),
                      pi.reshape(1, -1).to(gpu), torch.empty(jb.model_workspace_bytes(K) + 256,
                                                             dtype=torch.uint8, device=gpu),
                      torch.cuda.current_stream(gpu).cuda_stream)
    # alone == mix with reversible jobs
    specs, jobs = _jobs(gpu, dtype)
    alljobs = jobs + [jb]
    allspecs = specs + [sp]
    stream = synthetic.request_stream(24, len(alljobs), rate=3000.0, seed=4, ks=(1, 3))
    def reqs():
        return [serving.Request(i, alljobs[j], synthetic.draw_bls(allspecs[j], Kq, s), arrival=t)
                for i, (j, Kq, s, t) in enumerate(stream)]
    alone = reqs()
    serving.run_sequential(alone, device=gpu.index or 0)
    srv = serving.Server(32 << 30, n_streams=4, pool_bytes=16 << 30, device=gpu.index or 0)
    for i, jj in enumerate(alljobs):
        srv.register(f"j{i}", jj)
    mixed = reqs()
    srv.run(mixed)
    assert any(r.job is jb for r in mixed)
    for a, b2 in zip(alone, mixed):
        assert torch.equal(a.loglik, b2.loglik) and torch.equal(a.dbls, b2.dbls), a.rid
