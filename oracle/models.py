"""Substitution and site models, FP64 CPU restatement (test infrastructure only).

Follows torchtree 1.0.2:

* ``normalize_q`` / ``p_t_eigh``  -> ``AbstractSubstitutionModel.norm`` and
  ``SymmetricSubstitutionModel.p_t`` (``substitution_model/abstract.py:49-50,57-79``)
* ``jc69_p_t``                    -> ``JC69.p_t`` (``nucleotide.py:102-113``),
  ``general_jc69_p_t``            -> ``GeneralJC69.p_t`` (``general.py:51-69``)
* ``hky_q``                       -> ``HKY.q`` (``nucleotide.py:238-264``)
* ``gtr_q``                       -> ``GTR.q`` (``nucleotide.py:328-374``)
* ``general_symmetric_q``         -> ``GeneralSymmetricSubstitutionModel.q`` (``general.py:170-184``)
* ``mg94_masks`` / ``mg94_q``     -> ``MG94.__init__`` / ``MG94.q`` (``codon.py:33-77``)
* ``general_nonsymmetric_q``      -> ``GeneralNonSymmetricSubstitutionModel.q`` (``general.py:253-268``),
  ``p_t_expm``                    -> ``NonSymmetricSubstitutionModel.p_t`` (``abstract.py:88-94``)
* ``weibull_rates_probs``         -> ``UnivariateDiscretizedSiteModel.update_rates`` +
  ``WeibullSiteModel.inverse_cdf`` (``site_model.py:173-195,237-247``)
* ``invariant_rates_probs``       -> ``InvariantSiteModel.update_rates_probs`` (``site_model.py:90-100``)
* ``gamma_rates_probs``           -> no reference counterpart (torchtree has no gamma
  model); same quantile-midpoint contract as Weibull, Gamma(shape, rate=shape).

Pinned by: the R-derived GTR P(0.1) table and the HKY table
(``test/test_substitution_model.py:53-213``), the JC69 closed form, general ==
GTR, the MG94 row checks (``:298-310``), the Weibull tables
(``test/test_site_model.py``), ``scipy.linalg.expm`` on the same Q, and the
CTMC identities (rows sum to 1, P(0)=I, P(s)P(t)=P(s+t), detailed balance).
The gamma quantiles are pinned by ``scipy.special.gammaincinv`` and mean 1.
The non-symmetric Q and its P are pinned by the two-state closed form, a
hand-written three-state mapping, and the reduction to the symmetric model
(and to ``p_t_eigh``) when both triangles share their rates.
"""
from __future__ import annotations

import itertools
import math

import torch

F64 = torch.float64


def normalize_q(q: torch.Tensor, pi: torch.Tensor) -> torch.Tensor:
    """Q / (-sum_i pi_i Q_ii): mean substitution rate 1 (abstract.py:49-50,58)."""
    norm = -(torch.diagonal(q, dim1=-2, dim2=-1) * pi).sum(-1)
    return q / norm[..., None, None]


def p_t_eigh(q_unnorm: torch.Tensor, pi: torch.Tensor, tau: torch.Tensor) -> torch.Tensor:
    """P(tau) through the symmetrised eigensystem (abstract.py:57-79).

    ``q_unnorm`` [..., S, S] and ``pi`` [..., S] share leading (sample) dims;
    ``tau`` [..., B, C] (branch length times category rate). Returns
    [..., B, C, S, S].
    """
    q = normalize_q(q_unnorm, pi)
    sq = pi.sqrt()
    sym = sq[..., :, None] * q / sq[..., None, :]
    evals, evecs = torch.linalg.eigh(sym)
    left = evecs / sq[..., :, None]            # Pi^-1/2 V
    right = torch.linalg.inv(evecs) * sq[..., None, :]  # V^-1 Pi^1/2
    expo = torch.exp(evals[..., None, None, :] * tau[..., None])  # [..., B, C, S]
    lb = left[..., None, None, :, :]
    rb = right[..., None, None, :, :]
    return torch.matmul(lb * expo[..., None, :], rb)


def jc69_p_t(tau: torch.Tensor) -> torch.Tensor:
    """Closed form a = 1/4 + 3/4 e^{-4d/3}, b = 1/4 - 1/4 e^{-4d/3} (nucleotide.py:102-113)."""
    return general_jc69_p_t(tau, 4)


def general_jc69_p_t(tau: torch.Tensor, s: int) -> torch.Tensor:
    """JC model over ``s`` states (general.py:51-69)."""
    e = torch.exp(-s / (s - 1.0) * tau)
    same = 1.0 / s + (s - 1.0) / s * e
    diff = 1.0 / s - e / s
    eye = torch.eye(s, dtype=tau.dtype)
    return diff[..., None, None] * (1 - eye) + same[..., None, None] * eye


def _sym_from_upper(vals: torch.Tensor, s: int) -> torch.Tensor:
    """Symmetric matrix, upper triangle filled row by row from ``vals``."""
    r = torch.zeros(vals.shape[:-1] + (s, s), dtype=vals.dtype)
    iu = torch.triu_indices(s, s, 1)
    r[..., iu[0], iu[1]] = vals
    r[..., iu[1], iu[0]] = vals
    return r


def q_from_exchangeabilities(exch: torch.Tensor, pi: torch.Tensor) -> torch.Tensor:
    """Q_ij = R_ij pi_j (i != j), Q_ii = -sum_j Q_ij.  R given upper-triangle row-major."""
    s = pi.shape[-1]
    r = _sym_from_upper(exch, s)
    q = r * pi[..., None, :]
    q = q - torch.diag_embed(q.sum(-1))
    return q


def gtr_q(rates6: torch.Tensor, pi: torch.Tensor) -> torch.Tensor:
    """GTR: exchangeabilities a..f on (AC, AG, AT, CG, CT, GT) (nucleotide.py:328-374)."""
    lead = torch.broadcast_shapes(rates6.shape[:-1], pi.shape[:-1])
    return q_from_exchangeabilities(rates6.expand(lead + (6,)), pi.expand(lead + (4,)))


def hky_q(kappa: torch.Tensor, pi: torch.Tensor) -> torch.Tensor:
    """HKY: kappa on the transitions A<->G and C<->T (nucleotide.py:238-264)."""
    one = torch.ones_like(kappa)
    ex = torch.cat((one, kappa, one, one, kappa, one), -1)
    lead = torch.broadcast_shapes(ex.shape[:-1], pi.shape[:-1])
    return q_from_exchangeabilities(ex.expand(lead + (6,)), pi.expand(lead + (4,)))


def general_symmetric_q(rates: torch.Tensor, mapping: torch.Tensor, pi: torch.Tensor) -> torch.Tensor:
    """R_ij = rates[mapping[k]] for the k-th upper-triangle pair (general.py:170-184)."""
    ex = rates[..., mapping]
    lead = torch.broadcast_shapes(ex.shape[:-1], pi.shape[:-1])
    return q_from_exchangeabilities(ex.expand(lead + ex.shape[-1:]), pi.expand(lead + pi.shape[-1:]))


def general_nonsymmetric_q(rates: torch.Tensor, mapping: torch.Tensor, pi: torch.Tensor) -> torch.Tensor:
    """R from ``rates[mapping]``: the first half fills the upper triangle row by row, the second
    half the mirrored lower positions; Q = R diag(pi), rows summing to 0 (general.py:253-268)."""
    s = pi.shape[-1]
    iu = torch.triu_indices(s, s, 1)
    half = mapping.shape[-1] // 2
    lead = torch.broadcast_shapes(rates.shape[:-1], pi.shape[:-1])
    r = torch.zeros(lead + (s, s), dtype=rates.dtype)
    r[..., iu[0], iu[1]] = rates[..., mapping[:half]].expand(lead + (half,))
    r[..., iu[1], iu[0]] = rates[..., mapping[half:]].expand(lead + (half,))
    q = r * pi[..., None, :]
    return q - torch.diag_embed(q.sum(-1))


def p_t_expm(q_unnorm: torch.Tensor, pi: torch.Tensor, tau: torch.Tensor) -> torch.Tensor:
    """P(tau) = expm(Q tau) of the normalised Q (abstract.py:88-94); tau [..., B, C]."""
    q = normalize_q(q_unnorm, pi)
    return torch.matrix_exp(q[..., None, None, :, :] * tau[..., None, None])


UNIVERSAL = "KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSS*CWCLFLF"


def codon_triplets():
    """The 64 triplets in AAA, AAC, ... TTT order (base order A, C, G, T)."""
    return ["".join(t) for t in itertools.product("ACGT", repeat=3)]


def mg94_masks(table: str = UNIVERSAL):
    """Transition / synonymous / non-synonymous flags per upper-triangle pair.

    Pairs run over ``itertools.combinations`` of the sense codons, which is the
    row-major upper-triangle order (codon.py:33-55). Only codon pairs that
    differ at exactly one position get flags; every other pair keeps all three
    flags at 0 and therefore rate 1 in ``mg94_q``, as in the reference.
    """
    trip = codon_triplets()
    sense = [i for i in range(64) if table[i] != "*"]
    n = len(sense)
    m = n * (n - 1) // 2
    ts = torch.zeros(m, dtype=F64)
    syn = torch.zeros(m, dtype=F64)
    nonsyn = torch.zeros(m, dtype=F64)
    for k, (a, b) in enumerate(itertools.combinations(sense, 2)):
        ca, cb = trip[a], trip[b]
        diff = [p for p in range(3) if ca[p] != cb[p]]
        if len(diff) != 1:
            continue
        p = diff[0]
        if {ca[p], cb[p]} in ({"A", "G"}, {"C", "T"}):
            ts[k] = 1.0
        if table[a] == table[b]:
            syn[k] = 1.0
        else:
            nonsyn[k] = 1.0
    return ts, syn, nonsyn


def mg94_q(alpha, beta, kappa, pi, masks=None):
    """MG94 Q (codon.py:57-77): off-diagonal = kappa^ts * alpha^syn * beta^nonsyn * pi_j."""
    ts, syn, nonsyn = masks if masks is not None else mg94_masks()
    one = torch.ones_like(kappa)
    ex = (
        torch.where(ts == 1.0, kappa, one)
        * torch.where(syn == 1.0, alpha, one)
        * torch.where(nonsyn == 1.0, beta, one)
    )
    lead = torch.broadcast_shapes(ex.shape[:-1], pi.shape[:-1])
    return q_from_exchangeabilities(ex.expand(lead + ex.shape[-1:]), pi.expand(lead + pi.shape[-1:]))


def _quantile_midpoints(c: int) -> torch.Tensor:
    return (2.0 * torch.arange(c, dtype=F64) + 1.0) / (2.0 * c)


def weibull_rates_probs(shape: torch.Tensor, categories: int, invariant=None, mu=None):
    """Weibull rates at quantile midpoints, normalised to mean 1 (site_model.py:173-195,237-247).

    ``shape`` [..., 1]; returns rates [..., C'] and probs [..., C'] where
    C' = categories (+1 with an invariant class in front).
    """
    q = _quantile_midpoints(categories)
    base = torch.pow(-torch.log(1.0 - q), 1.0 / shape)
    if invariant is not None:
        probs = torch.cat((invariant, ((1.0 - invariant) / categories).expand(
            invariant.shape[:-1] + (categories,))), -1)
        rates = torch.cat((torch.zeros_like(invariant), base.expand(invariant.shape[:-1] + (categories,))), -1)
    else:
        probs = torch.full(base.shape, 1.0 / categories, dtype=base.dtype)
        rates = base
    rates = rates / (rates * probs).sum(-1, keepdim=True)
    if mu is not None:
        rates = rates * mu
    return rates, probs


def gamma_rates_probs(shape: float, categories: int):
    """Gamma(shape, rate=shape) quantile-midpoint rates, normalised to mean 1.

    No reference counterpart; the quantile contract matches
    ``UnivariateDiscretizedSiteModel.update_rates`` (site_model.py:173-195).
    """
    from scipy.special import gammaincinv

    q = _quantile_midpoints(categories).numpy()
    base = torch.tensor(gammaincinv(shape, q) / shape, dtype=F64)
    probs = torch.full((categories,), 1.0 / categories, dtype=F64)
    rates = base / (base * probs).sum()
    return rates, probs


def invariant_rates_probs(invariant: torch.Tensor, mu=None):
    """Two classes: rate 0 with p_inv, 1/(1-p_inv) with 1-p_inv (site_model.py:90-100)."""
    probs = torch.cat((invariant, 1.0 - invariant), -1)
    rates = torch.cat((torch.zeros_like(invariant), 1.0 / (1.0 - invariant)), -1)
    if mu is not None:
        rates = rates * mu
    return rates, probs
