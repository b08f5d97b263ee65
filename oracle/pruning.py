"""Felsenstein pruning, FP64 CPU restatement (test infrastructure only).

Follows ``torchtree/evolution/tree_likelihood.py`` (torchtree 1.0.2):

* ``treelik``                    -> ``calculate_treelikelihood`` (14-37)
* ``treelik_discrete``           -> ``calculate_treelikelihood_discrete`` (40-75)
* ``treelik_tip_states``         -> ``calculate_treelikelihood_tip_states_discrete`` (78-131)
* ``treelik_discrete_safe``      -> ``calculate_treelikelihood_discrete_safe`` (134-183)
* ``treelik_discrete_rescaled``  -> ``calculate_treelikelihood_discrete_rescaled`` (186-221)
* ``treelik_tip_states_rescaled``-> ``calculate_treelikelihood_tip_states_discrete_rescaled`` (224-278)

Shapes follow the reference: tips are ``[S, N]`` (or ``[N]`` integer codes for
the tip-states variants), internal partials ``[..., C, S, N]``, ``mats``
``[..., B, C, S, S]`` (``[..., B, S, S]`` for ``treelik``), ``freqs``
``[..., 1, S]``, ``props`` ``[..., C, 1, 1]``; every function returns
``[..., 1]``. Unlike the reference, the caller's ``partials`` list is not
mutated: a private copy holds the internal nodes.

Pinned by: the reference goldens -83.329016 (tiny, JC69) and
-4618.2062529058 (fluA, JC69+Weibull4+clock), brute-force enumeration of
internal states on small trees, the two-taxon closed form, rescaled ==
unscaled, and batch == single (``tests/test_oracle_pruning.py``).
"""
from __future__ import annotations

import torch


def _child_term(mats_b: torch.Tensor, partial: torch.Tensor) -> torch.Tensor:
    # P_b . L_b  for one child: [..., (C,) S, S] @ [..., (C,) S, N]
    return torch.matmul(mats_b, partial)


def treelik(partials, weights, post_indexing, mats, freqs):
    """Plain pruning, no rate categories (tree_likelihood.py:14-37)."""
    part = list(partials)
    for node, left, right in post_indexing:
        part[node] = _child_term(mats[..., left, :, :], part[left]) * _child_term(
            mats[..., right, :, :], part[right]
        )
    root = part[post_indexing[-1][0]]
    site = torch.log(torch.matmul(freqs, root))  # [..., 1, N]
    return (site * weights).sum(-1)


def _root_site_lik(root, freqs, props):
    # sum_c p_c L_root^c  then  pi . (.)   -> [..., 1, N]
    mixed = (props * root).sum(-3)
    return torch.matmul(freqs, mixed)


def treelik_discrete(partials, weights, post_indexing, mats, freqs, props):
    """Pruning with rate categories (tree_likelihood.py:40-75)."""
    part = list(partials)
    for node, left, right in post_indexing:
        part[node] = _child_term(mats[..., left, :, :, :], part[left]) * _child_term(
            mats[..., right, :, :, :], part[right]
        )
    site = torch.log(_root_site_lik(part[post_indexing[-1][0]], freqs, props))
    return (site * weights).sum(-1)


def _scale_node(partial):
    # scaler = max over (C, S) per (sample, pattern): tree_likelihood.py:168-174
    flat = partial.reshape(*partial.shape[:-3], -1, partial.shape[-1])
    scaler = flat.max(-2, keepdim=True).values  # [..., 1, N]
    return partial / scaler.unsqueeze(-2), scaler


def treelik_discrete_rescaled(partials, weights, post_indexing, mats, freqs, props):
    """Every node divided by its (C,S)-max (tree_likelihood.py:186-221)."""
    part = list(partials)
    log_scale = 0.0
    for node, left, right in post_indexing:
        p = _child_term(mats[..., left, :, :, :], part[left]) * _child_term(
            mats[..., right, :, :, :], part[right]
        )
        part[node], scaler = _scale_node(p)
        log_scale = log_scale + torch.log(scaler)
    site = torch.log(_root_site_lik(part[post_indexing[-1][0]], freqs, props))
    return ((site + log_scale) * weights).sum(-1)


def treelik_discrete_safe(partials, weights, post_indexing, mats, freqs, props,
                          threshold):
    """Selective rescaling after a first underflow (tree_likelihood.py:134-183).

    ``partials`` must already hold the internal nodes of an unscaled pass
    (the reference calls this after ``treelik_discrete`` filled them); use
    ``fill_internal`` to obtain them. A node is recomputed and rescaled when
    a child was rescaled or its own (C,S)-max falls below ``threshold``.
    """
    part = list(partials)
    rescaled = [False] * (post_indexing[-1][0] + 1)
    log_scale = 0.0
    for node, left, right in post_indexing:
        low = bool((part[node].max(-2, keepdim=True).values < threshold).any())
        if rescaled[left] or rescaled[right] or low:
            p = _child_term(mats[..., left, :, :, :], part[left]) * _child_term(
                mats[..., right, :, :, :], part[right]
            )
            part[node], scaler = _scale_node(p)
            log_scale = log_scale + torch.log(scaler)
            rescaled[node] = True
    site = torch.log(_root_site_lik(part[post_indexing[-1][0]], freqs, props))
    return ((site + log_scale) * weights).sum(-1)


def fill_internal(partials, post_indexing, mats):
    """Unscaled internal partials, as left behind by ``treelik_discrete``."""
    part = list(partials)
    for node, left, right in post_indexing:
        part[node] = _child_term(mats[..., left, :, :, :], part[left]) * _child_term(
            mats[..., right, :, :, :], part[right]
        )
    return part


def _tip_gather(mats_b, codes):
    # mats with an extra all-ones column, selected by code: tree_likelihood.py:106-116
    ones = torch.ones(mats_b.shape[:-1] + (1,), dtype=mats_b.dtype)
    ext = torch.cat((mats_b, ones), -1)  # [..., C, S, S+1]
    return ext[..., codes]  # [..., C, S, N]


def _tip_states_child(mats, b, part, tip_count):
    if b < tip_count:
        return _tip_gather(mats[..., b, :, :, :], part[b])
    return _child_term(mats[..., b, :, :, :], part[b])


def treelik_tip_states(partials, weights, post_indexing, mats, freqs, props):
    """Tips given as integer codes (tree_likelihood.py:78-131)."""
    part = list(partials)
    tip_count = len(post_indexing) + 1
    for node, left, right in post_indexing:
        part[node] = _tip_states_child(mats, left, part, tip_count) * _tip_states_child(
            mats, right, part, tip_count
        )
    site = torch.log(_root_site_lik(part[post_indexing[-1][0]], freqs, props))
    return (site * weights).sum(-1)


def treelik_tip_states_rescaled(partials, weights, post_indexing, mats, freqs, props):
    """Tip codes plus per-node rescaling (tree_likelihood.py:224-278)."""
    part = list(partials)
    tip_count = len(post_indexing) + 1
    log_scale = 0.0
    for node, left, right in post_indexing:
        p = _tip_states_child(mats, left, part, tip_count) * _tip_states_child(
            mats, right, part, tip_count
        )
        part[node], scaler = _scale_node(p)
        log_scale = log_scale + torch.log(scaler)
    site = torch.log(_root_site_lik(part[post_indexing[-1][0]], freqs, props))
    return ((site + log_scale) * weights).sum(-1)


def brute_force_loglik(tips, weights, post_indexing, mats, freqs, props):
    """Sum over every assignment of internal states (small trees only).

    ``tips`` are ``[S, N]`` partials, ``mats`` ``[B, C, S, S]`` (one sample),
    ``freqs`` ``[S]``, ``props`` ``[C]``. Enumerates S^(T-1) assignments;
    independent of the recursion used by every other function here.
    """
    import itertools

    T = len(tips)
    S, N = tips[0].shape
    root = post_indexing[-1][0]
    parent = {}
    for node, left, right in post_indexing:
        parent[left] = node
        parent[right] = node
    internals = [n for n, _, _ in post_indexing]
    C = mats.shape[1]
    site = torch.zeros(N, dtype=torch.float64)
    for c in range(C):
        acc = torch.zeros(N, dtype=torch.float64)
        for assign in itertools.product(range(S), repeat=len(internals)):
            st = dict(zip(internals, assign))
            w = float(freqs[st[root]])
            for node, left, right in post_indexing:
                for ch in (left, right):
                    if ch >= T:
                        w *= float(mats[ch, c, st[node], st[ch]])
            if w == 0.0:
                continue
            term = torch.full((N,), w, dtype=torch.float64)
            for tip in range(T):
                term = term * torch.matmul(mats[tip, c, st[parent[tip]], :], tips[tip])
            acc = acc + term
        site = site + props[c] * acc
    return float((torch.log(site) * weights).sum())
