"""Engine hook for ``torchtree-cli ... --engine torchtree_amd``.

torchtree's CLI imports ``<engine>.cli.evolution`` and calls its
``process_tree_likelihood(arg, data)`` on every ``TreeLikelihoodModel`` it writes
(``torchtree/cli/evolution.py:700-710,1739-1754``). Selecting this engine retargets the object
to the HIP ``TreeLikelihoodModel``; ``--hiptree_precision`` applies when the plug-in's
arguments are loaded too.
"""
from __future__ import annotations

from ..plugin import retarget


def process_tree_likelihood(arg, data):
    retarget(data, getattr(arg, "hiptree_precision", None))
