"""Pins the oracle's P(t), Q builders and site models.

Goldens from the reference tests: the R-derived GTR P(0.1) table and its
batched form (test/test_substitution_model.py:53-136), the HKY tables
(:139-213), JC69 closed form (:216-230), general symmetric == GTR (:31-50),
MG94 rows (:298-310), Weibull rate tables (test/test_site_model.py:26-137).
Independent checks: scipy.linalg.expm of the same Q, CTMC identities
(rows sum to 1, P(0)=I, P(s)P(t)=P(s+t), detailed balance, pi P = pi) and
scipy.special.gammaincinv for the gamma quantiles. The non-symmetric Q and its
expm are pinned by the two-state closed form, a hand-written three-state
mapping and the reduction to the symmetric model (general.py:202-291).
"""
import numpy as np
import pytest
import torch

from oracle import models

F64 = torch.float64
RATES = torch.tensor([0.060602, 0.402732, 0.028230, 0.047910, 0.407249, 0.053277], dtype=F64)
PI = torch.tensor([0.479367, 0.172572, 0.140933, 0.207128], dtype=F64)
GTR_P01 = torch.tensor([
    [0.93717830, 0.009506685, 0.047505899, 0.005809115],
    [0.02640748, 0.894078744, 0.006448058, 0.073065722],
    [0.16158572, 0.007895626, 0.820605951, 0.009912704],
    [0.01344433, 0.060875872, 0.006744752, 0.918935042]], dtype=F64)
HKY_P = torch.tensor([
    [[0.93211187, 0.01511617, 0.03462891, 0.01814305],
     [0.04198939, 0.89405292, 0.01234480, 0.05161289],
     [0.11778615, 0.01511617, 0.84895463, 0.01814305],
     [0.04198939, 0.04300210, 0.01234480, 0.90266370]],
    [[0.9992649548, 0.0001581235, 0.0003871353, 0.0001897863],
     [0.0004392323, 0.9988625812, 0.0001291335, 0.0005690531],
     [0.0013167952, 0.0001581235, 0.9983352949, 0.0001897863],
     [0.0004392323, 0.0004741156, 0.0001291335, 0.9989575186]]], dtype=F64)


def test_gtr_table():
    P = models.p_t_eigh(models.gtr_q(RATES, PI), PI, torch.tensor([[0.1]], dtype=F64))
    assert torch.allclose(P.squeeze(), GTR_P01, rtol=1e-5)


def test_gtr_batch_table():
    rates = torch.stack((RATES, torch.tensor([1.0, 3.0, 1.0, 1.0, 3.0, 1.0], dtype=F64)))
    pi = PI.expand(2, 4)
    P = models.p_t_eigh(models.gtr_q(rates, pi), pi, torch.tensor([[[0.1]], [[0.001]]], dtype=F64))
    assert torch.allclose(P[0, 0, 0], GTR_P01, rtol=1e-5)
    assert torch.allclose(P[1, 0, 0], HKY_P[1], rtol=1e-5)


def test_hky_table():
    """HKY kappa=3 at t=0.1 and t=0.001 (test_substitution_model.py:139-175)."""
    q = models.hky_q(torch.tensor([3.0], dtype=F64), PI)
    P = models.p_t_eigh(q, PI, torch.tensor([[0.1], [0.001]], dtype=F64))
    assert torch.allclose(P[0, 0], HKY_P[0], atol=1e-6)
    assert torch.allclose(P[1, 0], HKY_P[1], atol=1e-6)


@pytest.mark.parametrize("t", [0.001, 0.1, 1.7])
def test_jc69_closed_form(t):
    P = models.jc69_p_t(torch.tensor([t], dtype=F64))[0]
    ii = 0.25 + 0.75 * np.exp(-4.0 / 3.0 * t)
    ij = 0.25 - 0.25 * np.exp(-4.0 / 3.0 * t)
    assert abs(float(P[0, 0]) - ii) < 1e-15 and abs(float(P[0, 1]) - ij) < 1e-15
    # and equal to the eigen route with equal rates, uniform pi
    Pe = models.p_t_eigh(models.gtr_q(torch.ones(6, dtype=F64), torch.full((4,), 0.25, dtype=F64)),
                         torch.full((4,), 0.25, dtype=F64), torch.tensor([[t]], dtype=F64))[0, 0]
    assert torch.allclose(P, Pe, atol=1e-14)


def test_general_symmetric_equals_gtr():
    qg = models.general_symmetric_q(RATES, torch.arange(6), PI)
    assert torch.allclose(qg, models.gtr_q(RATES, PI), rtol=0, atol=0)


def test_general_jc69_equals_jc69():
    t = torch.tensor([[0.1], [0.5]], dtype=F64)
    assert torch.allclose(models.general_jc69_p_t(t, 4), models.jc69_p_t(t), rtol=1e-15)


def test_mg94_rows():
    """test_substitution_model.py:298-310: alpha=beta=kappa=1, uniform pi."""
    one = torch.tensor([1.0], dtype=F64)
    pi = torch.full((61,), 1.0 / 61, dtype=F64)
    q = models.mg94_q(one, one, one, pi)
    assert torch.allclose(q[0, 1:], torch.full((60,), 1 / 61, dtype=F64))
    assert torch.allclose(torch.diagonal(q), torch.full((61,), -60 / 61, dtype=F64))


def test_mg94_flags_counts():
    """Ordered one-step sense-to-sense mutations of the universal code: 61*9 - 23 = 526."""
    ts, syn, non = models.mg94_masks()
    tab, trip = models.UNIVERSAL, models.codon_triplets()
    ordered = 0
    for a in range(64):
        if tab[a] == "*":
            continue
        for p in range(3):
            for b in "ACGT":
                t = trip[a][:p] + b + trip[a][p + 1:]
                if b != trip[a][p] and tab[trip.index(t)] != "*":
                    ordered += 1
    assert ordered == 526
    assert int((syn + non).sum()) * 2 == ordered
    assert torch.all((ts == 0) | ((syn + non) == 1))
    assert torch.all(syn * non == 0)


# Hand-derived MG94 entries (codon.py:33-77) under the universal code. Sense-state indices are
# 16*b1 + 4*b2 + b3 (A=0, C=1, G=2, T=3) minus the stops TAA (48), TAG (50), TGA (56) before it.
# Each entry is the exchangeability times pi_j, with kappa on transitions (A<->G, C<->T), alpha
# on synonymous and beta on non-synonymous one-step changes, and 1 for multi-step pairs.
_A, _B, _K = 1.3, 0.3, 2.5
MG94_ENTRIES = [
    # (from, to, state i, state j, exchangeability)
    ("AAA", "AAG", 0, 2, _K * _A),    # K -> K, A<->G: synonymous transition
    ("AAA", "AAC", 0, 1, _B),         # K -> N, A<->C: non-synonymous transversion
    ("AAA", "AGA", 0, 8, _K * _B),    # K -> R, A<->G: non-synonymous transition
    ("AAA", "CCC", 0, 21, 1.0),       # three positions differ
    ("AAA", "ACC", 0, 5, 1.0),        # two positions differ
    ("CTA", "CTC", 28, 29, _A),       # L -> L, A<->C: synonymous transversion
    ("TTA", "CTA", 57, 28, _K * _A),  # L -> L, T<->C at the first position
    ("ATG", "ATA", 14, 12, _K * _B),  # M -> I, G<->A: non-synonymous transition
    ("TAC", "TAT", 48, 49, _K * _A),  # Y -> Y across the TAA stop
    ("TGG", "TGC", 55, 54, _B),       # W -> C, G<->C across three stops
    ("TGG", "CGG", 55, 26, _K * _B),  # W -> R, T<->C
    ("GTT", "TTT", 47, 60, _B),       # V -> F, G<->T: transversion
]


def _mg94_qs():
    from torchtree_amd import models as hm

    g = torch.Generator().manual_seed(5)
    pi = torch.rand(61, generator=g, dtype=F64) + 0.2
    pi = pi / pi.sum()
    a, b, k = (torch.tensor([v], dtype=F64) for v in (_A, _B, _K))
    return pi, models.mg94_q(a, b, k, pi), hm.mg94_q(a, b, k, pi)


@pytest.mark.parametrize("which", ["oracle", "product"])
def test_mg94_hand_derived_entries(which):
    """Q_ij of MG94 with alpha, beta and kappa all different, against entries derived by hand
    from the universal code; catches a wrong transition set, a swapped synonymous test or a
    misplaced stop-codon shift (both the oracle and the engine's Q builder)."""
    pi, qo, qp = _mg94_qs()
    q = qo if which == "oracle" else qp
    trip = models.codon_triplets()
    sense = [c for i, c in enumerate(trip) if models.UNIVERSAL[i] != "*"]
    for c1, c2, i, j, ex in MG94_ENTRIES:
        assert sense[i] == c1 and sense[j] == c2
        assert abs(float(q[i, j]) - ex * float(pi[j])) < 1e-15, (c1, c2)
        assert abs(float(q[j, i]) - ex * float(pi[i])) < 1e-15, (c2, c1)
    assert torch.allclose(q.sum(-1), torch.zeros(61, dtype=F64), atol=1e-14)


def test_mg94_pair_totals():
    """Unordered one-step sense pairs of the universal code, counted by mutating each codon
    position: 263 in all, 67 of them synonymous. Transitions: each of the 64 codons has one
    transition neighbour per position, 96 unordered pairs, less the 7 that touch a stop codon
    (TAA-TAG, TAA-TGA, TAA-CAA, TAG-TGG, TAG-CAG, TGA-TGG, TGA-CGA): 89."""
    ts, syn, non = models.mg94_masks()
    assert int((syn + non).sum()) == 263
    assert int(ts.sum()) == 89
    assert int(syn.sum()) == 67
    tab, trip = models.UNIVERSAL, models.codon_triplets()
    n_ts = n_syn = 0
    for a in range(64):
        for p in range(3):
            for base in "ACGT":
                t = trip[a][:p] + base + trip[a][p + 1:]
                b = trip.index(t)
                if b <= a or tab[a] == "*" or tab[b] == "*":
                    continue
                n_ts += {trip[a][p], base} in ({"A", "G"}, {"C", "T"})
                n_syn += tab[a] == tab[b]
    assert (n_ts, n_syn) == (89, 67)


@pytest.mark.parametrize("seed,S", [(0, 4), (1, 20), (2, 61)])
def test_expm_and_identities(seed, S):
    from scipy.linalg import expm

    rng = np.random.default_rng(seed)
    pi = torch.tensor(rng.dirichlet(np.ones(S) * 3), dtype=F64)
    ex = torch.tensor(rng.gamma(1.0, 1.0, S * (S - 1) // 2), dtype=F64)
    q = models.q_from_exchangeabilities(ex, pi)
    qn = models.normalize_q(q, pi)
    assert abs(float(-(torch.diagonal(qn) * pi).sum()) - 1.0) < 1e-14
    ts = torch.tensor([[0.0, 0.05, 0.3, 2.0]], dtype=F64)
    P = models.p_t_eigh(q, pi, ts)[0]
    for i, t in enumerate(ts[0]):
        ref = expm(qn.numpy() * float(t))
        assert np.allclose(P[i].numpy(), ref, atol=1e-12)
    assert torch.allclose(P.sum(-1), torch.ones(4, S, dtype=F64), atol=1e-12)
    assert torch.allclose(P[0], torch.eye(S, dtype=F64), atol=1e-12)
    Pst = models.p_t_eigh(q, pi, torch.tensor([[0.35]], dtype=F64))[0, 0]
    assert torch.allclose(P[1] @ P[2], Pst, atol=1e-12)
    db = pi[:, None] * P[2]
    assert torch.allclose(db, db.t(), atol=1e-13)
    assert torch.allclose(pi @ P[3], pi, atol=1e-12)


def test_weibull_tables():
    r1, p1 = models.weibull_rates_probs(torch.tensor([[1.0], [0.1]], dtype=F64), 4)
    exp = torch.tensor([[0.1457844, 0.5131316, 1.0708310, 2.2702530],
                        [4.766392e-12, 1.391131e-06, 2.179165e-03, 3.997819]], dtype=F64)
    assert torch.allclose(r1, exp, rtol=1e-6)
    assert torch.allclose((r1 * p1).sum(-1), torch.ones(2, dtype=F64), rtol=1e-14)
    r2, p2 = models.weibull_rates_probs(torch.tensor([1.0], dtype=F64), 3,
                                        invariant=torch.tensor([0.0], dtype=F64))
    assert torch.allclose(r2, torch.tensor([0.0, 0.20506860315799713, 0.7796264290809631,
                                            2.0153050422668457], dtype=F64), rtol=1e-6)
    r3, _ = models.weibull_rates_probs(torch.tensor([1.0], dtype=F64), 4,
                                       mu=torch.tensor([2.0], dtype=F64))
    assert torch.allclose(r3, exp[0] * 2, rtol=1e-6)


def test_invariant():
    r, p = models.invariant_rates_probs(torch.tensor([0.2], dtype=F64),
                                        mu=torch.tensor([2.0], dtype=F64))
    assert abs(float(r[1]) - 2 / 0.8) < 1e-15
    assert torch.allclose(p, torch.tensor([0.2, 0.8], dtype=F64))


def test_gamma_quantiles():
    from scipy.special import gammaincinv
    from scipy.stats import gamma as gdist

    r, p = models.gamma_rates_probs(0.5, 4)
    q = (2 * np.arange(4) + 1) / 8
    raw = gdist.ppf(q, 0.5, scale=1 / 0.5)
    assert np.allclose(r.numpy(), raw / raw.mean(), rtol=1e-12)
    assert abs(float((r * p).sum()) - 1.0) < 1e-14


def test_nonsymmetric_two_state_closed_form():
    """Two states: Q01 = a = r01 pi1, Q10 = b = r10 pi0, normalised by pi0 a + pi1 b; then
    P(t) = [[b + a e, a (1 - e)], [b (1 - e), a + b e]] / (a + b) with e = exp(-(a + b) t)."""
    rates = torch.tensor([0.7, 2.3], dtype=F64)
    pi = torch.tensor([0.35, 0.65], dtype=F64)
    mapping = torch.tensor([0, 1])
    q = models.general_nonsymmetric_q(rates, mapping, pi)
    a, b = 0.7 * 0.65, 2.3 * 0.35
    assert torch.allclose(q, torch.tensor([[-a, a], [b, -b]], dtype=F64), atol=1e-15)
    norm = 0.35 * a + 0.65 * b
    a, b = a / norm, b / norm
    for t in (0.01, 0.3, 2.0):
        e = np.exp(-(a + b) * t)
        want = torch.tensor([[b + a * e, a * (1 - e)], [b * (1 - e), a + b * e]], dtype=F64) / (a + b)
        got = models.p_t_expm(q, pi, torch.tensor([[t]], dtype=F64))[0, 0]
        assert torch.allclose(got, want, atol=1e-14)


def test_nonsymmetric_three_state_mapping():
    """The first half of the mapping fills (0,1), (0,2), (1,2); the second (1,0), (2,0), (2,1)."""
    rates = torch.tensor([1.0, 2.0, 3.0, 4.0, 5.0, 6.0], dtype=F64)
    pi = torch.tensor([0.2, 0.3, 0.5], dtype=F64)
    q = models.general_nonsymmetric_q(rates, torch.tensor([5, 4, 3, 2, 1, 0]), pi)
    r = torch.tensor([[0.0, 6.0, 5.0], [3.0, 0.0, 4.0], [2.0, 1.0, 0.0]], dtype=F64)
    want = r * pi
    want = want - torch.diag(want.sum(-1))
    assert torch.allclose(q, want, atol=1e-15)


def test_nonsymmetric_reduces_to_symmetric():
    """Equal rates in both triangles give the symmetric model's Q, and expm gives p_t_eigh."""
    g = torch.Generator().manual_seed(4)
    s, d = 5, 10
    rates = torch.rand(d, generator=g, dtype=F64) + 0.1
    pi = torch.rand(s, generator=g, dtype=F64) + 0.2
    pi = pi / pi.sum()
    mapping = torch.cat((torch.arange(d), torch.arange(d)))
    qn = models.general_nonsymmetric_q(rates, mapping, pi)
    qs = models.general_symmetric_q(rates, torch.arange(d), pi)
    assert torch.allclose(qn, qs, atol=1e-15)
    tau = torch.tensor([[0.05, 0.4], [1.1, 0.0]], dtype=F64)
    assert torch.allclose(models.p_t_expm(qn, pi, tau), models.p_t_eigh(qs, pi, tau), atol=1e-12)


def _empirical_fixture():
    import json
    from pathlib import Path

    d = json.loads((Path(__file__).parent / "golden" / "empirical_aa.json").read_text())
    return {k: (torch.tensor(v["rates"], dtype=F64), torch.tensor(v["frequencies"], dtype=F64))
            for k, v in d.items()}


@pytest.mark.parametrize("name", ["LG", "WAG"])
def test_empirical_tables_and_identities(name):
    """LG / WAG (amino_acid.py:12-105): the engine's tables equal the reference's (fixture made
    from its source text), and P(t) from them is a proper reversible transition matrix."""
    from torchtree_amd import empirical

    rates, pi = _empirical_fixture()[name]
    r2, f2 = empirical.empirical(name)
    assert torch.equal(rates, r2) and torch.equal(pi, f2)
    assert rates.shape == (190,) and pi.shape == (20,) and bool((rates > 0).all())
    assert abs(float(pi.sum()) - 1.0) < 2e-6
    q = models.q_from_exchangeabilities(rates, pi)
    qn = models.normalize_q(q, pi)
    assert abs(float(-(torch.diagonal(qn) * pi).sum()) - 1.0) < 1e-14
    t = torch.tensor([[0.0, 0.1, 0.35, 2.0]], dtype=F64)
    P = models.p_t_eigh(q, pi, t)[0]  # [4, 20, 20]
    assert torch.allclose(P[0], torch.eye(20, dtype=F64), atol=1e-13)
    assert torch.allclose(P.sum(-1), torch.ones(4, 20, dtype=F64), atol=1e-13)
    assert torch.allclose(pi @ P, pi.expand(4, 20), atol=1e-13)
    assert torch.allclose(pi[:, None] * P, (pi[:, None] * P).transpose(-1, -2), atol=1e-14)
    P45 = models.p_t_eigh(q, pi, torch.tensor([[0.45]], dtype=F64))[0, 0]
    assert torch.allclose(P[1] @ P[2], P45, atol=1e-13)
    from scipy.linalg import expm

    assert np.allclose(P[3].numpy(), expm(qn.numpy() * 2.0), atol=1e-12)


@pytest.mark.parametrize("inp,size", [([[1.0]], [1, 1, 20, 20]), ([[1.0, 1.0]], [1, 2, 20, 20]),
                                      ([[1.0], [1.0]], [2, 1, 20, 20]),
                                      ([[1.0, 1.0], [1.0, 1.0], [1.0, 1.0]], [3, 2, 20, 20])])
def test_empirical_p_t_shapes(inp, size):
    """test/test_substitution_model.py:283-295 (LG)."""
    rates, pi = _empirical_fixture()["LG"]
    P = models.p_t_eigh(models.q_from_exchangeabilities(rates, pi), pi, torch.tensor(inp, dtype=F64))
    assert list(P.shape) == size
