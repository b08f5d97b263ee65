"""Device-resident tree-likelihood jobs and the native autograd path.

A ``TreeJob`` keeps one alignment's tips (uint8 codes or dense), weights and
topology plan resident on the device. ``TreeJob.loglik`` evaluates
``sum_n w_n log L_n`` for K samples of branch lengths through

    Q, pi (host FP64 eigen) -> hx_pmatrix -> hx_loglik            (forward)
    hx_loglik_grad_mats -> hx_pmatrix_grad -> d bls, d rates, d Q  (backward)

which is the computation of ``TreeLikelihoodModel._call``
(``evolution/tree_likelihood.py:313-356``) with the P matrices of
``SymmetricSubstitutionModel.p_t`` (``substitution_model/abstract.py:57-79``)
and the gradient that torch autograd derives in the reference.
"""
from __future__ import annotations

import torch

from . import _lib
from .likelihood import Pruning, TipData, get_plan
from .models import Eigen, pmatrix, pmatrix_grad

F64 = torch.float64


class _TreeLikFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, bls, rates, props, q_norm, freqs, job):
        eig = Eigen(q_norm, freqs, bls.device)
        mats = pmatrix(eig, bls, rates, job.dtype)
        K = bls.shape[0]
        fr = freqs.detach().to(bls.device, job.dtype).reshape(-1, eig.S).contiguous()
        pp = props.detach().to(bls.device, job.dtype).reshape(-1, props.shape[-1]).contiguous()
        need = any(ctx.needs_input_grad[:5])
        out, dfreqs, dprops, ws = job.pruning.forward(mats, fr, pp, want_root_grads=need)
        if need:
            ctx.job, ctx.eig, ctx.ws = job, eig, ws
            ctx.mats, ctx.fr, ctx.pp = mats, fr, pp
            ctx.dfreqs, ctx.dprops = dfreqs, dprops
            ctx.save_for_backward(bls, rates, props, q_norm, freqs)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        bls, rates, props, q_norm, freqs = ctx.saved_tensors
        if ctx.ws is None:
            raise _lib.HipTreeError("the pruning workspace was consumed by an earlier backward")
        K = bls.shape[0]
        g = grad_out.to(F64)
        dmats = ctx.job.pruning.backward_mats(ctx.mats, ctx.fr, ctx.pp, ctx.ws)
        ctx.ws = None
        dmats = dmats * g.view(K, 1, 1, 1, 1)
        dbls, drates, dq = pmatrix_grad(ctx.eig, bls, rates, dmats)
        if rates.shape[0] != K or K == 1:
            drates = drates.sum(0, keepdim=True)
        if q_norm.dim() == 2:
            dq = dq.sum(0)
        df = ctx.dfreqs * g.view(K, 1)
        if freqs.dim() == 1 or freqs.shape[0] != K or K == 1:
            df = df.sum(0)
            df = df.reshape(freqs.shape)
        dp = ctx.dprops * g.view(K, 1)
        if props.dim() == 1 or props.shape[0] != K or K == 1:
            dp = dp.sum(0).reshape(props.shape)
        return (dbls.to(bls.device, bls.dtype),
                drates.reshape(rates.shape).to(rates.device, rates.dtype),
                dp.to(props.device, props.dtype),
                dq.reshape(q_norm.shape).to(q_norm.device, q_norm.dtype),
                df.to(freqs.device, freqs.dtype), None)


class TreeJob:
    """One resident likelihood job: topology, tips and weights on the device.

    ``post_indexing``: (node, left, right) triples (tree_model.py:187-195).
    ``tips``: list of T tensors, [S, N] partials or [N] integer codes in 0..S
    (``tip_kind`` "partials" or "states"), or a ready ``TipData``.
    ``dtype``: torch.float32 (rescaled FP32 partials) or torch.float64.
    ``products``: for FP32, "fp32" or "bf16x3" (dense-state products as split BF16 on the
    matrix cores, FP32 accumulation).
    """

    def __init__(self, post_indexing, tips, weights, S, dtype=torch.float32, device=None,
                 tip_kind="partials", products="fp32"):
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        device = torch.device(device)
        self.device = device
        self.dtype = dtype
        self.S = S
        T = len(post_indexing) + 1
        self.T = T
        self.B = 2 * T - 2
        idx = device.index if device.index is not None else torch.cuda.current_device()
        with torch.cuda.device(idx):
            self.plan = get_plan(T, post_indexing, idx)
            if isinstance(tips, TipData):
                td = tips
            elif tip_kind == "states":
                td = TipData.from_states(tips, S, device)
            else:
                td = TipData.from_partials(tips, device)
        self.tips = td
        self.N = td.N
        self.products = products
        self.pruning = Pruning(self.plan, td, torch.as_tensor(weights).to(device),
                               products=products)

    def loglik(self, bls, rates, props, q_norm, freqs):
        """loglik [K] (FP64) with autograd w.r.t. every tensor argument.

        bls [K,B]; rates [K|1,C]; props [K|1,C]; q_norm [S,S] or [K,S,S]
        (normalised, mean rate 1); freqs [S] or [K,S].
        """
        bls = bls.to(self.device)
        rates = rates.to(self.device)
        if rates.dim() == 1:
            rates = rates.unsqueeze(0)
        return _TreeLikFn.apply(bls, rates, props.to(self.device), q_norm, freqs.to(self.device),
                                self)

    def workspace_bytes(self, C, K):
        from ._lib import HX_F32, HX_F64

        code = HX_F64 if self.dtype == F64 else HX_F32
        return self.plan.workspace_bytes(self.S, self.N, C, K, code, _lib.HX_GRAD_MATS)
