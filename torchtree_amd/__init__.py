"""hiptree: MI355X-native tree-likelihood engine for torchtree's pruning path.

``__plugin__`` names the torchtree-cli plug-in class, relative to this package, as torchtree's
plug-in manager expects of ``torchtree_*`` modules (``torchtree/cli/plugin_manager.py:10-23``).
Importing the package loads nothing native; ``torchtree_amd._lib.load()`` maps the HIP library.
"""
__plugin__ = "plugin.HipTreePlugin"
