"""Substitution and site models feeding the HIP pruning kernels.

Host-side model code (small S x S matrices, category rates) that builds the
inputs of the hot path, following torchtree 1.0.2:

* ``gtr_q`` / ``hky_q`` / ``jc69_q`` / ``general_symmetric_q`` / ``mg94_q`` ->
  ``GTR.q`` (``evolution/substitution_model/nucleotide.py:328-374``), ``HKY.q``
  (238-264), ``JC69.q`` (115-125), ``GeneralSymmetricSubstitutionModel.q``
  (``general.py:170-184``), ``MG94.q`` (``codon.py:57-77``)
* ``normalize`` -> ``AbstractSubstitutionModel.norm`` (``abstract.py:49-50``)
* ``Eigen`` -> the symmetrised eigendecomposition of
  ``SymmetricSubstitutionModel.p_t`` (``abstract.py:57-79``), kept in FP64
* ``weibull_rates`` / ``invariant_rates`` -> ``site_model.py:90-100,173-195,237-247``
* ``gamma_rates`` -> new (torchtree has no gamma model), same quantile-midpoint
  contract, Gamma(shape, rate=shape)

``Eigen``, ``pmatrix`` and ``pmatrix_grad`` are the transition-matrix stage of
the native path (driven by ``engine._TreeLikFn``): FP64 eigen on the host,
``hx_pmatrix`` for P on the device, and ``hx_pmatrix_grad`` for the chain rule
back to branch lengths, category rates and Q.
"""
from __future__ import annotations

import itertools

import torch

from . import _lib, timing
from ._lib import HX_F32, HX_F64, HipTreeError

F64 = torch.float64


def _sym_from_upper(vals, s):
    iu = torch.triu_indices(s, s, 1)
    r = vals.new_zeros(vals.shape[:-1] + (s, s))
    r[..., iu[0], iu[1]] = vals
    r[..., iu[1], iu[0]] = vals
    return r


def q_from_exchangeabilities(exch, pi):
    """Q_ij = R_ij pi_j off the diagonal, rows summing to zero."""
    s = pi.shape[-1]
    lead = torch.broadcast_shapes(exch.shape[:-1], pi.shape[:-1])
    exch = exch.expand(lead + exch.shape[-1:])
    pi = pi.expand(lead + (s,))
    q = _sym_from_upper(exch, s) * pi.unsqueeze(-2)
    return q - torch.diag_embed(q.sum(-1))


def gtr_q(rates, pi):
    """Exchangeabilities (AC, AG, AT, CG, CT, GT)."""
    return q_from_exchangeabilities(rates, pi)


def hky_q(kappa, pi):
    one = torch.ones_like(kappa)
    return q_from_exchangeabilities(torch.cat((one, kappa, one, one, kappa, one), -1), pi)


def jc69_q(dtype=F64):
    pi = torch.full((4,), 0.25, dtype=dtype)
    return q_from_exchangeabilities(torch.ones(6, dtype=dtype), pi), pi


def general_symmetric_q(rates, mapping, pi):
    return q_from_exchangeabilities(rates[..., mapping], pi)


def general_nonsymmetric_q(rates, mapping, pi):
    """Non-reversible Q (general.py:253-268): ``rates[mapping]`` with the first half of the
    mapping on the upper triangle (row by row) and the second half on the mirrored lower
    positions; Q = R diag(pi) with rows summing to zero."""
    s = pi.shape[-1]
    iu = torch.triu_indices(s, s, 1, device=rates.device)
    half = mapping.shape[-1] // 2
    lead = torch.broadcast_shapes(rates.shape[:-1], pi.shape[:-1])
    upper = rates[..., mapping[:half]].expand(lead + (half,))
    lower = rates[..., mapping[half:]].expand(lead + (half,))
    r = torch.zeros(lead + (s, s), dtype=rates.dtype, device=rates.device)
    r[..., iu[0], iu[1]] = upper
    r[..., iu[1], iu[0]] = lower
    q = r * pi[..., None, :]
    return q - torch.diag_embed(q.sum(-1))


class _ExpmFn(torch.autograd.Function):
    """mats [K, B, C, S, S] = exp(Q_e bls[k, b] rates[k, c]) on the device (hx_pmatrix_expm),
    with the backward to Q, bls and rates through the Frechet derivative (hx_pmatrix_expm_grad)."""

    @staticmethod
    def forward(ctx, q, bls, rates, dtype):
        lib = _lib.load()
        K, B = bls.shape
        E, S, _ = q.shape
        C = rates.shape[-1]
        dev = bls.device
        qd = q.detach().to(dev, F64).contiguous()
        bd = bls.detach().to(F64).contiguous()
        rd = rates.detach().to(dev, F64).contiguous()
        out = torch.empty(K, B, C, S, S, dtype=dtype, device=dev)
        code = HX_F64 if dtype == F64 else HX_F32
        with timing.phase("pmatrix_expm"):
            rc = lib.hx_pmatrix_expm(S, K, B, C, int(E == K and K > 1), _lib.ptr(qd), _lib.ptr(bd),
                                     _lib.ptr(rd), int(rd.shape[0] == K and K > 1), _lib.ptr(out),
                                     code, _lib.stream_handle(dev))
        _lib.check(rc, "hx_pmatrix_expm")
        ctx.save_for_backward(qd, bd, rd)
        ctx.shapes = (q.shape, q.device, q.dtype, rates.shape, rates.device, rates.dtype)
        return out

    @staticmethod
    def backward(ctx, g):
        lib = _lib.load()
        qd, bd, rd = ctx.saved_tensors
        K, B = bd.shape
        E, S, _ = qd.shape
        C = rd.shape[-1]
        dev = bd.device
        gd = g.to(F64).contiguous()
        scratch = torch.empty(int(lib.hx_pmatrix_expm_grad_scratch_bytes(S, K, B, C)),
                              dtype=torch.uint8, device=dev)
        dbls = torch.empty(K, B, dtype=F64, device=dev)
        drates = torch.empty(K, C, dtype=F64, device=dev)
        dq = torch.empty(E, S, S, dtype=F64, device=dev)
        rb = int(rd.shape[0] == K and K > 1)
        with timing.phase("pmatrix_expm_grad"):
            rc = lib.hx_pmatrix_expm_grad(S, K, B, C, int(E == K and K > 1), _lib.ptr(qd),
                                          _lib.ptr(bd), _lib.ptr(rd), rb, _lib.ptr(gd),
                                          _lib.ptr(dbls), _lib.ptr(drates), _lib.ptr(dq),
                                          _lib.ptr(scratch), _lib.stream_handle(dev))
        _lib.check(rc, "hx_pmatrix_expm_grad")
        qshape, qdev, qdt, rshape, rdev, rdt = ctx.shapes
        if not rb:
            drates = drates.sum(0, keepdim=True)
        return (dq.reshape(qshape).to(qdev, qdt), dbls.to(bd.dtype),
                drates.reshape(rshape).to(rdev, rdt), None)


def pmatrix_expm(q_norm, bls, rates, dtype):
    """P = exp(Q tau) of a normalised rate matrix of any symmetry, tau = bls[k, b] rates[., c]
    (NonSymmetricSubstitutionModel.p_t, abstract.py:88-94), on the device.

    ``q_norm`` [S, S] or [E, S, S] (E = 1 or K), ``bls`` [K, B] (device), ``rates`` [K|1, C];
    returns mats [K, B, C, S, S] in ``dtype``, differentiable with respect to all three."""
    q = q_norm if q_norm.dim() == 3 else q_norm.unsqueeze(0)
    if rates.dim() == 1:
        rates = rates.unsqueeze(0)
    return _ExpmFn.apply(q, bls, rates, dtype)


UNIVERSAL_CODE = "KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSS*CWCLFLF"


def mg94_flags(table=UNIVERSAL_CODE):
    """Per upper-triangle codon pair: (transition, synonymous, non-synonymous) flags.

    Pairs that differ at more than one position keep all flags at zero and
    so get exchangeability 1, as in ``MG94.__init__`` / ``MG94.q`` (codon.py:33-77).
    """
    trip = ["".join(t) for t in itertools.product("ACGT", repeat=3)]
    sense = [i for i in range(64) if table[i] != "*"]
    ts, syn, non = [], [], []
    for a, b in itertools.combinations(sense, 2):
        diff = [p for p in range(3) if trip[a][p] != trip[b][p]]
        t = s_ = n_ = 0.0
        if len(diff) == 1:
            p = diff[0]
            pair = {trip[a][p], trip[b][p]}
            t = 1.0 if pair in ({"A", "G"}, {"C", "T"}) else 0.0
            if table[a] == table[b]:
                s_ = 1.0
            else:
                n_ = 1.0
        ts.append(t)
        syn.append(s_)
        non.append(n_)
    return (torch.tensor(ts, dtype=F64), torch.tensor(syn, dtype=F64),
            torch.tensor(non, dtype=F64))


def mg94_q(alpha, beta, kappa, pi, flags=None):
    ts, syn, non = flags if flags is not None else mg94_flags()
    one = torch.ones_like(kappa)
    ex = (torch.where(ts == 1.0, kappa, one) * torch.where(syn == 1.0, alpha, one)
          * torch.where(non == 1.0, beta, one))
    return q_from_exchangeabilities(ex, pi)


def normalize(q, pi):
    """Scale Q to mean rate 1: Q / (-sum_i pi_i Q_ii)."""
    norm = -(torch.diagonal(q, dim1=-2, dim2=-1) * pi).sum(-1)
    return q / norm[..., None, None]


# ------------------------------------------------------------- site models


def _midpoints(c, dtype=F64):
    return (2.0 * torch.arange(c, dtype=dtype) + 1.0) / (2.0 * c)


def weibull_rates(shape, categories, invariant=None, mu=None):
    """Rates and proportions [..., C] of the discretised Weibull model."""
    q = _midpoints(categories, shape.dtype)
    base = torch.pow(-torch.log1p(-q), 1.0 / shape)
    if invariant is not None:
        props = torch.cat((invariant, ((1.0 - invariant) / categories).expand(
            invariant.shape[:-1] + (categories,))), -1)
        rates = torch.cat((torch.zeros_like(invariant),
                           base.expand(invariant.shape[:-1] + (categories,))), -1)
    else:
        props = torch.full(base.shape, 1.0 / categories, dtype=base.dtype)
        rates = base
    rates = rates / (rates * props).sum(-1, keepdim=True)
    if mu is not None:
        rates = rates * mu
    return rates, props


def gamma_rates(shape: float, categories: int):
    """Gamma(shape, rate=shape) quantile midpoints, normalised to mean 1."""
    from scipy.special import gammaincinv

    q = _midpoints(categories).numpy()
    base = torch.tensor(gammaincinv(shape, q) / shape, dtype=F64)
    props = torch.full((categories,), 1.0 / categories, dtype=F64)
    return base / (base * props).sum(), props


def invariant_rates(invariant, mu=None):
    props = torch.cat((invariant, 1.0 - invariant), -1)
    rates = torch.cat((torch.zeros_like(invariant), 1.0 / (1.0 - invariant)), -1)
    if mu is not None:
        rates = rates * mu
    return rates, props


# -------------------------------------------------------------- eigen + P


class Eigen:
    """FP64 eigensystem of normalised reversible Q matrices, device resident.

    For Q reversible with respect to pi, Pi^1/2 Q Pi^-1/2 is symmetric; with
    its eigenpairs (lambda, V): P(t) = (Pi^-1/2 V) diag(e^{lambda t}) (V^T Pi^1/2).
    ``left`` = Pi^-1/2 V and ``right`` = V^T Pi^1/2 (its inverse).
    """

    def __init__(self, q_norm: torch.Tensor, pi: torch.Tensor, device):
        qh = q_norm.detach().to("cpu", F64)
        ph = pi.detach().to("cpu", F64).expand(qh.shape[:-1])
        sq = ph.sqrt()
        sym = sq.unsqueeze(-1) * qh / sq.unsqueeze(-2)
        sym = 0.5 * (sym + sym.transpose(-1, -2))
        lam, vec = torch.linalg.eigh(sym)
        left = vec / sq.unsqueeze(-1)
        right = vec.transpose(-1, -2) * sq.unsqueeze(-2)
        self.batched = qh.dim() == 3 and qh.shape[0] > 1
        E = qh.shape[0] if qh.dim() == 3 else 1
        S = qh.shape[-1]
        self.S = S
        self.E = E
        self.lam = lam.reshape(E, S).to(device).contiguous()
        self.left = left.reshape(E, S, S).to(device).contiguous()
        self.right = right.reshape(E, S, S).to(device).contiguous()


def pmatrix(eig: Eigen, bls: torch.Tensor, rates: torch.Tensor, dtype) -> torch.Tensor:
    """mats [K,B,C,S,S] in ``dtype`` from bls [K,B] and rates [K|1,C] (FP64, device)."""
    lib = _lib.load()
    K, B = bls.shape
    C = rates.shape[-1]
    if eig.batched and eig.E != K:
        raise HipTreeError("batched eigensystem must have one entry per sample")
    bls = bls.to(F64).contiguous()
    rates = rates.to(F64).contiguous()
    out = torch.empty(K, B, C, eig.S, eig.S, dtype=dtype, device=bls.device)
    code = HX_F64 if dtype == F64 else HX_F32
    with timing.phase("pmatrix"):
        rc = lib.hx_pmatrix(eig.S, K, B, C, int(eig.batched), _lib.ptr(eig.lam), _lib.ptr(eig.left),
                        _lib.ptr(eig.right), _lib.ptr(bls), _lib.ptr(rates),
                        int(rates.shape[0] == K and K > 1), _lib.ptr(out), code,
                        _lib.stream_handle(bls.device))
    _lib.check(rc, "hx_pmatrix")
    return out


def pmatrix_grad(eig: Eigen, bls, rates, dmats):
    """(dbls [K,B], drates [K,C], dq [E,S,S]) from dmats [K,B,C,S,S] (FP64)."""
    lib = _lib.load()
    K, B = bls.shape
    C = rates.shape[-1]
    dev = bls.device
    bls = bls.to(F64).contiguous()
    rates = rates.to(F64).contiguous()
    dmats = dmats.to(F64).contiguous()
    scratch = torch.empty(int(lib.hx_pmatrix_grad_scratch_bytes(eig.S, K, B, C)), dtype=torch.uint8,
                          device=dev)
    dbls = torch.empty(K, B, dtype=F64, device=dev)
    drates = torch.empty(K, C, dtype=F64, device=dev)
    dq = torch.empty(eig.E, eig.S, eig.S, dtype=F64, device=dev)
    with timing.phase("pmatrix_grad", 4):
        rc = lib.hx_pmatrix_grad(eig.S, K, B, C, int(eig.batched), _lib.ptr(eig.lam),
                             _lib.ptr(eig.left), _lib.ptr(eig.right), _lib.ptr(bls),
                             _lib.ptr(rates), int(rates.shape[0] == K and K > 1), _lib.ptr(dmats),
                             _lib.ptr(dbls), _lib.ptr(drates), _lib.ptr(dq), _lib.ptr(scratch),
                             _lib.stream_handle(dev))
    _lib.check(rc, "hx_pmatrix_grad")
    return dbls, drates, dq
