"""hx_job_eval_noalloc: the job evaluation with the library's static device workspace.

Called through the C ABI and compared with hx_job_eval (caller workspace) on the same job and
inputs, bit for bit (the same kernels run in the same order), on two calls in a row, the second
with a smaller input; a request larger than the static workspace is refused before any launch.
"""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
F64 = torch.float64


def _job(gpu, kind, T, N, rng, dtype=torch.float32, slice_patterns=0):
    from torchtree_amd import serving, synthetic

    sp = synthetic.random_job_spec(kind, T, N, rng, name=f"{kind}{T}")
    post = [tuple(int(x) for x in r) for r in sp["topo"].postorder]
    job = serving.ResidentJob(sp["name"], post, sp["codes"], sp["table"], sp["weights"],
                              sp["q_norm"], sp["freqs"], sp["rates"], sp["props"], dtype=dtype,
                              slice_patterns=slice_patterns, device=gpu.index or 0, kind=kind)
    return sp, job


def _both(lib, job, bls, gpu, grad=True):
    from torchtree_amd import _lib

    K = bls.shape[0]
    st = torch.cuda.current_stream(gpu).cuda_stream
    out = []
    for noalloc in (False, True):
        ll = torch.full((K,), float("nan"), dtype=F64, device=gpu)
        db = torch.full((K, job.B), float("nan"), dtype=F64, device=gpu) if grad else None
        if noalloc:
            rc = lib.hx_job_eval_noalloc(job.handle, K, _lib.ptr(bls), _lib.ptr(ll), _lib.ptr(db), st)
        else:
            ws = torch.empty(job.workspace_bytes(K, grad), dtype=torch.uint8, device=gpu)
            rc = lib.hx_job_eval(job.handle, K, _lib.ptr(bls), _lib.ptr(ll), _lib.ptr(db),
                                 _lib.ptr(ws), ws.numel(), st)
        _lib.check(rc, "hx_job_eval_noalloc" if noalloc else "hx_job_eval")
        torch.cuda.synchronize(gpu)
        out.append((ll.cpu(), None if db is None else db.cpu()))
    return out


def test_noalloc_equals_job_eval_twice(gpu):
    print("case: hx_job_eval_noalloc vs hx_job_eval, two calls, second smaller", flush=True)
    from torchtree_amd import _lib, synthetic

    lib = _lib.load()
    rng = np.random.default_rng(71)
    sp, job = _job(gpu, "dna", 48, 1500, rng)
    need = job.workspace_bytes(8, True)
    assert need <= _lib.HX_NOALLOC_WORKSPACE_BYTES, need
    bls = synthetic.draw_bls(sp, 8, 5).to(gpu)
    (a_ll, a_db), (b_ll, b_db) = _both(lib, job, bls, gpu)
    assert torch.isfinite(a_ll).all() and torch.isfinite(a_db).all()
    assert torch.equal(a_ll, b_ll) and torch.equal(a_db, b_db)
    # second call: a smaller request on another job (20 states), over what the first left behind
    sp2, job2 = _job(gpu, "aa", 20, 300, rng)
    bls2 = synthetic.draw_bls(sp2, 2, 6).to(gpu)
    (c_ll, c_db), (d_ll, d_db) = _both(lib, job2, bls2, gpu)
    assert torch.isfinite(c_ll).all() and torch.isfinite(c_db).all()
    assert torch.equal(c_ll, d_ll) and torch.equal(c_db, d_db)
    # loglik only, smaller still
    (e_ll, _), (f_ll, _) = _both(lib, job2, bls2[:1].contiguous(), gpu, grad=False)
    assert torch.equal(e_ll, f_ll) and torch.equal(e_ll, c_ll[:1])


def test_noalloc_refuses_oversized_request(gpu):
    print("case: hx_job_eval_noalloc oversized request refused", flush=True)
    from torchtree_amd import _lib, synthetic

    lib = _lib.load()
    rng = np.random.default_rng(72)
    sp, job = _job(gpu, "dna", 200, 40000, rng)
    K = 8
    assert job.workspace_bytes(K, True) > _lib.HX_NOALLOC_WORKSPACE_BYTES
    bls = synthetic.draw_bls(sp, K, 1).to(gpu)
    ll = torch.zeros(K, dtype=F64, device=gpu)
    db = torch.zeros(K, job.B, dtype=F64, device=gpu)
    rc = lib.hx_job_eval_noalloc(job.handle, K, _lib.ptr(bls), _lib.ptr(ll), _lib.ptr(db),
                                 torch.cuda.current_stream(gpu).cuda_stream)
    assert rc == 4  # HX_ERR_WORKSPACE_TOO_SMALL
    assert b"static workspace" in lib.hx_last_error()
    torch.cuda.synchronize(gpu)
    assert bool((ll == 0).all()) and bool((db == 0).all())
