"""torchtree-cli engine entry (``--engine torchtree_amd``): see ``evolution``."""
