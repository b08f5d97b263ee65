"""Swap the pruning functions of a torchtree module for the HIP engine's.

``TreeLikelihoodModel`` (``torchtree/evolution/tree_likelihood.py:281-450``)
calls the six ``calculate_treelikelihood*`` functions through its module's
globals (``tree_likelihood.py:358-420``), so replacing those module attributes
routes every existing torchtree model, JSON file and optimiser loop through the
HIP kernels with no other change. ``install`` does that and returns the
previous functions; ``uninstall`` puts them back.
"""
from __future__ import annotations

from . import likelihood

NAMES = (
    "calculate_treelikelihood",
    "calculate_treelikelihood_discrete",
    "calculate_treelikelihood_tip_states_discrete",
    "calculate_treelikelihood_discrete_safe",
    "calculate_treelikelihood_discrete_rescaled",
    "calculate_treelikelihood_tip_states_discrete_rescaled",
)


def install(module) -> dict:
    """Replace the six functions on ``module`` (e.g. ``torchtree.evolution.tree_likelihood``).

    Returns ``{name: previous function}`` for ``uninstall``. Raises
    ``AttributeError`` if ``module`` lacks one of them (not a tree-likelihood
    module), before changing anything.
    """
    missing = [n for n in NAMES if not hasattr(module, n)]
    if missing:
        raise AttributeError(f"{getattr(module, '__name__', module)!r} has no {', '.join(missing)}")
    saved = {n: getattr(module, n) for n in NAMES}
    for n in NAMES:
        setattr(module, n, getattr(likelihood, n))
    return saved


def uninstall(module, saved: dict) -> None:
    """Restore the functions returned by ``install``."""
    for n, f in saved.items():
        setattr(module, n, f)
