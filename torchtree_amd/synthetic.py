"""Seeded synthetic workloads: random trees, CTMC simulation, unique patterns.

Node numbering follows the torchtree convention (``setup_indexes``,
``evolution/tree_model.py:37-53``): tips 0..T-1, internal nodes T..2T-2 in
post-order with the root last, branch b = node b. Unrooted trees carry
2T-3 free lengths; slot 2T-3 (a child of the root) is held at zero, as
``TreeLikelihoodModel._call`` pads it (``tree_likelihood.py:323-337``).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np
import torch

from . import models


@dataclasses.dataclass
class Topology:
    T: int
    postorder: np.ndarray    # [T-1, 3] int32 (node, left, right)
    parent: np.ndarray       # [2T-1] int32, -1 at the root

    @property
    def B(self):
        return 2 * self.T - 2

    def preorder_edges(self):
        """(parent, child) for every non-root node, parents before children."""
        out = []
        for v, a, b in self.postorder[::-1]:
            out.append((int(v), int(a)))
            out.append((int(v), int(b)))
        return out

    def newick(self, lengths=None, names=None):
        kids = {int(v): (int(a), int(b)) for v, a, b in self.postorder}

        def rec(n):
            if n < self.T:
                s = names[n] if names else f"t{n}"
            else:
                a, b = kids[n]
                s = f"({rec(a)},{rec(b)})"
            if lengths is not None and n != 2 * self.T - 2:
                s += f":{float(lengths[n]):.17g}"
            return s

        return rec(2 * self.T - 2) + ";"


def random_topology(T: int, rng: np.random.Generator) -> Topology:
    """Random rooted binary topology by repeated joining of two random lineages."""
    if T < 2:
        raise ValueError("need at least two taxa")
    lineages = list(range(T))
    children = {}
    nxt = T
    while len(lineages) > 1:
        i, j = rng.choice(len(lineages), size=2, replace=False)
        a, b = lineages[i], lineages[j]
        for idx in sorted((i, j), reverse=True):
            lineages.pop(idx)
        children[nxt] = (a, b)
        lineages.append(nxt)
        nxt += 1
    root = lineages[0]
    # renumber internal nodes in post-order, root last
    order = []
    stack = [(root, False)]
    while stack:
        n, done = stack.pop()
        if n < T:
            continue
        if done:
            order.append(n)
        else:
            stack.append((n, True))
            a, b = children[n]
            stack.append((b, False))
            stack.append((a, False))
    remap = {old: T + i for i, old in enumerate(order)}
    remap.update({t: t for t in range(T)})
    post = np.array([(remap[n], remap[children[n][0]], remap[children[n][1]]) for n in order],
                    dtype=np.int32)
    parent = np.full(2 * T - 1, -1, dtype=np.int32)
    for v, a, b in post:
        parent[a] = v
        parent[b] = v
    return Topology(T, post, parent)


def unrooted_lengths(topo: Topology, rng, mean: float) -> np.ndarray:
    """2T-2 branch lengths ~ Exp(mean) with slot 2T-3 (a root child) set to 0."""
    bl = rng.exponential(mean, size=topo.B)
    bl[topo.B - 1] = 0.0
    return bl


@dataclasses.dataclass
class Workload:
    topo: Topology
    S: int
    codes: torch.Tensor       # [T, N] uint8, S = unknown
    weights: torch.Tensor     # [N] float64
    bls: torch.Tensor         # [K, B] float64
    rates: torch.Tensor       # [C] float64
    props: torch.Tensor       # [C] float64
    q_norm: torch.Tensor      # [S, S] float64, mean rate 1
    freqs: torch.Tensor       # [S] float64
    exch: torch.Tensor        # exchangeabilities (upper triangle) float64
    name: str = ""

    @property
    def T(self):
        return self.topo.T

    @property
    def N(self):
        return self.codes.shape[1]

    @property
    def K(self):
        return self.bls.shape[0]

    @property
    def C(self):
        return self.rates.shape[0]

    def tip_table(self):
        return torch.cat((torch.eye(self.S, dtype=torch.float64),
                          torch.ones(1, self.S, dtype=torch.float64)))

    def tip_partials(self):
        """Dense tips [S, N] per taxon (the reference's compress_alignment layout)."""
        tbl = self.tip_table()
        return [tbl[self.codes[t].long()].t().contiguous() for t in range(self.T)]

    def updates(self, grad: bool) -> int:
        """Site-partial updates of one evaluation (post-order, plus pre-order with grad)."""
        per = self.N * self.C * self.K
        u = (self.T - 1) * per
        if grad:
            u += (2 * self.T - 2) * per
        return u


def simulate_codes(topo, bl, q_norm, freqs, rates, props, n_sites, rng):
    """CTMC simulation of ``n_sites`` columns down the tree (root ~ freqs)."""
    S = freqs.shape[0]
    C = rates.shape[0]
    lam, vec = np.linalg.eigh(_sym(q_norm, freqs))
    sq = np.sqrt(freqs)
    left = vec / sq[:, None]
    right = vec.T * sq[None, :]
    cat = rng.choice(C, size=n_sites, p=props)
    states = np.empty((2 * topo.T - 1, n_sites), dtype=np.int64)
    root = 2 * topo.T - 2
    states[root] = rng.choice(S, size=n_sites, p=freqs)
    for par, ch in topo.preorder_edges():
        # cumulative P rows per category for this branch
        P = np.einsum("im,cm,mj->cij", left, np.exp(np.outer(rates * bl[ch], lam)), right)
        P = np.clip(P, 0.0, None)
        P /= P.sum(-1, keepdims=True)
        cum = np.cumsum(P, -1)  # [C, S, S]
        rows = cum[cat, states[par]]  # [n, S]
        u = rng.random(n_sites)
        states[ch] = np.minimum((u[:, None] > rows).sum(-1), S - 1)
    return states[: topo.T]


def _sym(q, pi):
    sq = np.sqrt(pi)
    s = sq[:, None] * q / sq[None, :]
    return 0.5 * (s + s.T)


def unique_patterns(tip_states: np.ndarray, n_patterns: int):
    """First ``n_patterns`` distinct columns in order of appearance."""
    T, n = tip_states.shape
    cols = np.ascontiguousarray(tip_states.T.astype(np.uint8))
    view = cols.view(np.dtype((np.void, T)))
    _, first = np.unique(view, return_index=True)
    first = np.sort(first)[:n_patterns]
    return tip_states[:, first]


def make_workload(T, n_patterns, K, *, S=4, model="gtr", categories=4, site="weibull",
                  shape=0.5, bl_mean=0.02, sigma=0.1, seed=0, weights="poisson3",
                  gap_fraction=0.0, name=""):
    """Seeded workload of the kind the benchmark configurations describe.

    model: "jc69" | "gtr" (Dirichlet(10) freqs, Gamma(2, rate .5) exchangeabilities) |
           "aa" (Gamma(1,1) exchangeabilities, Dirichlet(1) freqs) |
           "mg94" (kappa=2, omega=0.3, uniform codon freqs).
    site:  "weibull" | "gamma" | "constant".
    weights: "poisson3" (Poisson(3)+1) or "counts" (pattern multiplicities).
    """
    rng = np.random.default_rng(seed)
    topo = random_topology(T, rng)
    bl = unrooted_lengths(topo, rng, bl_mean)
    if model == "jc69":
        S = 4
        freqs = np.full(4, 0.25)
        exch = np.ones(6)
        q = models.gtr_q(torch.tensor(exch), torch.tensor(freqs))
    elif model == "gtr":
        S = 4
        freqs = rng.dirichlet(np.full(4, 10.0))
        exch = rng.gamma(2.0, 1.0 / 0.5, size=6)
        q = models.gtr_q(torch.tensor(exch), torch.tensor(freqs))
    elif model == "aa":
        S = 20
        freqs = rng.dirichlet(np.ones(20))
        exch = rng.gamma(1.0, 1.0, size=190)
        q = models.q_from_exchangeabilities(torch.tensor(exch), torch.tensor(freqs))
    elif model == "mg94":
        S = 61
        freqs = np.full(61, 1.0 / 61)
        q = models.mg94_q(torch.tensor([1.0], dtype=torch.float64),
                          torch.tensor([0.3], dtype=torch.float64),
                          torch.tensor([2.0], dtype=torch.float64), torch.tensor(freqs))
        exch = torch.tensor([1.0, 0.3, 2.0], dtype=torch.float64).numpy()
    else:
        raise ValueError(model)
    q_norm = models.normalize(q, torch.tensor(freqs))
    if site == "weibull":
        rates, props = models.weibull_rates(torch.tensor([shape], dtype=torch.float64), categories)
    elif site == "gamma":
        rates, props = models.gamma_rates(shape, categories)
    else:
        rates, props = torch.ones(1, dtype=torch.float64), torch.ones(1, dtype=torch.float64)
    n_sim = int(n_patterns * 1.3) + 64
    states = None
    for _ in range(6):
        sim = simulate_codes(topo, bl, q_norm.numpy(), freqs, rates.numpy(), props.numpy(), n_sim,
                             rng)
        states = sim if states is None else np.concatenate((states, sim), 1)
        pats = unique_patterns(states, n_patterns)
        if pats.shape[1] >= n_patterns:
            break
        n_sim *= 2
    if gap_fraction > 0:
        mask = rng.random(pats.shape) < gap_fraction
        pats = np.where(mask, S, pats)
    N = pats.shape[1]
    if weights == "poisson3":
        w = rng.poisson(3.0, size=N) + 1
    else:
        w = np.ones(N)
    draws = bl[None, :] * np.exp(sigma * rng.standard_normal((K, topo.B)))
    draws[:, topo.B - 1] = 0.0
    return Workload(
        topo=topo, S=S,
        codes=torch.tensor(pats.astype(np.uint8)),
        weights=torch.tensor(w, dtype=torch.float64),
        bls=torch.tensor(draws, dtype=torch.float64),
        rates=rates.reshape(-1).to(torch.float64), props=props.reshape(-1).to(torch.float64),
        q_norm=q_norm.to(torch.float64), freqs=torch.tensor(freqs, dtype=torch.float64),
        exch=torch.as_tensor(exch, dtype=torch.float64), name=name)


def config1(seed=1):
    """64 taxa, JC69, 1,000 simulated sites compressed to unique patterns, K=1, C=1."""
    rng = np.random.default_rng(seed)
    topo = random_topology(64, rng)
    bl = unrooted_lengths(topo, rng, 0.05)
    freqs = np.full(4, 0.25)
    q = models.normalize(models.gtr_q(torch.ones(6, dtype=torch.float64),
                                      torch.tensor(freqs)), torch.tensor(freqs))
    sim = simulate_codes(topo, bl, q.numpy(), freqs, np.ones(1), np.ones(1), 1000, rng)
    cols, counts = np.unique(sim.T, axis=0, return_counts=True)
    pats = cols.T
    return Workload(
        topo=topo, S=4, codes=torch.tensor(pats.astype(np.uint8)),
        weights=torch.tensor(counts, dtype=torch.float64),
        bls=torch.tensor(bl[None, :], dtype=torch.float64),
        rates=torch.ones(1, dtype=torch.float64), props=torch.ones(1, dtype=torch.float64),
        q_norm=q, freqs=torch.tensor(freqs), exch=torch.ones(6, dtype=torch.float64),
        name="config1-jc69-64taxa")


def config2(seed=2, N=100_000, K=32, T=512):
    """GTR+Weibull4(0.5), 512 taxa, 100,000 unique patterns, K=32 log-normal draws."""
    return make_workload(T, N, K, model="gtr", categories=4, site="weibull", shape=0.5,
                         bl_mean=0.02, seed=seed, name=f"config2-gtr-w4-{T}taxa-{N}pat-K{K}")


def config3(seed=3, N=50_000, K=16, T=1024):
    """GTR+Weibull4, 1,024 taxa, 50,000 patterns, K=16."""
    return make_workload(T, N, K, model="gtr", categories=4, site="weibull", shape=0.5,
                         bl_mean=0.02, seed=seed, name=f"config3-gtr-w4-{T}taxa-{N}pat-K{K}")


# ------------------------------------------------------------ serving mix


def random_job_spec(kind, T, N, rng, name=""):
    """A resident-job spec with random tip states (2% unknown) for serving.

    kind: "dna" (GTR + Weibull4(0.5)), "aa" (20 states, Gamma(1,1)
    exchangeabilities, Dirichlet(1) frequencies, gamma4(0.5)), "codon"
    (MG94 kappa=2, omega=0.3, uniform frequencies, gamma4(0.5)).
    """
    topo = random_topology(T, rng)
    if kind == "dna":
        S = 4
        freqs = rng.dirichlet(np.full(4, 10.0))
        q = models.gtr_q(torch.tensor(rng.gamma(2.0, 2.0, 6)), torch.tensor(freqs))
        rates, props = models.weibull_rates(torch.tensor([0.5], dtype=torch.float64), 4)
    elif kind == "aa":
        S = 20
        freqs = rng.dirichlet(np.ones(20))
        q = models.q_from_exchangeabilities(torch.tensor(rng.gamma(1.0, 1.0, 190)), torch.tensor(freqs))
        rates, props = models.gamma_rates(0.5, 4)
    elif kind == "codon":
        S = 61
        freqs = np.full(61, 1.0 / 61)
        one = torch.ones(1, dtype=torch.float64)
        q = models.mg94_q(one, 0.3 * one, 2.0 * one, torch.tensor(freqs))
        rates, props = models.gamma_rates(0.5, 4)
    else:
        raise ValueError(kind)
    q_norm = models.normalize(q, torch.tensor(freqs))
    codes = rng.integers(0, S, size=(T, N), dtype=np.uint8)
    unknown = rng.random((T, N)) < 0.02
    codes[unknown] = S
    table = np.concatenate((np.eye(S), np.ones((1, S))))
    bl = unrooted_lengths(topo, rng, 0.05)
    return dict(name=name, kind=kind, topo=topo, S=S, codes=codes, table=table,
                weights=rng.poisson(3.0, N) + 1.0, q_norm=q_norm, freqs=freqs,
                rates=rates.reshape(-1).numpy(), props=props.reshape(-1).numpy(), bl=bl)


def serving_mix(n_jobs=200, seed=5, t_range=(16, 1024), n_range=(100, 100_000),
                mix=(("dna", 0.80), ("aa", 0.15), ("codon", 0.05))):
    """Job specs of the serving configuration: T ~ U{16..1024}, N log-uniform, 80/15/5 types."""
    rng = np.random.default_rng(seed)
    kinds, probs = zip(*mix)
    specs = []
    for j in range(n_jobs):
        kind = kinds[rng.choice(len(kinds), p=probs)]
        T = int(rng.integers(t_range[0], t_range[1] + 1))
        N = int(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1]))))
        specs.append(random_job_spec(kind, T, N, rng, name=f"job{j}-{kind}-T{T}-N{N}"))
    return specs


def request_stream(n_requests, n_jobs, rate, seed=6, ks=(1, 8, 32), sigma=0.1):
    """(job index, K, log-normal jitter seed, arrival time) of Poisson arrivals at ``rate``/s."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    return [(int(rng.integers(n_jobs)), int(rng.choice(ks)), int(rng.integers(1 << 30)), float(t[i]))
            for i in range(n_requests)]


def draw_bls(spec, K, seed, sigma=0.1):
    rng = np.random.default_rng(seed)
    bl = spec["bl"]
    draws = bl[None, :] * np.exp(sigma * rng.standard_normal((K, bl.shape[0])))
    draws[:, -1] = 0.0
    return torch.tensor(draws, dtype=torch.float64)
