"""torchtree-cli plug-in: route the generated ``TreeLikelihoodModel`` to the HIP engine.

torchtree discovers plug-ins as top-level modules named ``torchtree_*`` that expose
``__plugin__`` (``torchtree/cli/plugin_manager.py:10-23``): here
``torchtree_amd.__plugin__ = "plugin.HipTreePlugin"``. The manager instantiates the class, calls
``load_arguments(subparsers)`` once (``cli/cli.py:59-60``) and, for every tree likelihood the CLI
writes, ``process_tree_likelihood(arg, data)`` (``cli/evolution.py:663-664,697-698``; the
interface is ``cli/plugin.py:4-16``). With ``--hiptree`` on the command line the plug-in
retargets the JSON object's ``type`` to this package's ``TreeLikelihoodModel``, which the
torchtree runner then loads by its dotted name (``core/utils.py:116-125``); the keys of the
object are unchanged, and ``precision`` picks FP32 (with rescaling) or FP64 partials.

The same retargeting is reachable through ``torchtree-cli ... --engine torchtree_amd``, which
imports ``torchtree_amd.cli.evolution`` (``cli/evolution.py:700-710,1739-1754``).
"""
from __future__ import annotations

ENGINE_TYPE = "torchtree_amd.treelikelihood_model.TreeLikelihoodModel"
PRECISIONS = ("float32", "float64", "bf16x3")

try:  # torchtree's abstract base when torchtree is importable; the interface otherwise
    from torchtree.cli.plugin import Plugin as _Base  # type: ignore
except Exception:  # pragma: no cover - torchtree (and DendroPy) absent
    class _Base:  # same three hooks as torchtree/cli/plugin.py:4-16
        def load_arguments(self, subparsers):
            raise NotImplementedError

        def process_tree_likelihood(self, arg, data):
            pass

        def process_coalescent(self, arg, data):
            pass

        def process_all(self, arg, data):
            pass


def retarget(data: dict, precision: str | None = None) -> dict:
    """Point one ``TreeLikelihoodModel`` JSON object (in place) at the HIP engine."""
    if data.get("type", "").split(".")[-1] != "TreeLikelihoodModel":
        raise ValueError(f"not a TreeLikelihoodModel object: {data.get('type')!r}")
    data["type"] = ENGINE_TYPE
    if precision is not None:
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {PRECISIONS}")
        data["precision"] = precision
    return data


class HipTreePlugin(_Base):
    """Adds ``--hiptree`` and ``--hiptree_precision`` to every torchtree-cli sub-command."""

    def load_arguments(self, subparsers):
        for parser in subparsers.choices.values():
            parser.add_argument("--hiptree", action="store_true",
                                help="evaluate tree likelihoods with the MI355X HIP engine")
            parser.add_argument("--hiptree_precision", choices=PRECISIONS, default=None,
                                help="partials precision of the HIP engine (default: the "
                                     "runner's default dtype)")

    def process_tree_likelihood(self, arg, data):
        if getattr(arg, "hiptree", False):
            retarget(data, getattr(arg, "hiptree_precision", None))
