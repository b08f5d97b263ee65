"""CPU FP64 oracle for the tree-likelihood path.

TEST INFRASTRUCTURE ONLY. This package restates the reference method
(torchtree 1.0.2, ``torchtree/evolution/tree_likelihood.py`` and the model
code that feeds it) in plain PyTorch float64 on the CPU. It is the checker
for the HIP engine in ``torchtree_amd``. Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
may import it, and only to check or to time the baseline. The product path
never routes through it.

Modules
-------
pruning   the six ``calculate_treelikelihood*`` functions
models    Q builders, P(t), site models (rates / proportions)
data      data types, FASTA, pattern compression, Newick + node indexing,
          time-tree branch lengths

Gradients of the oracle are torch autograd through these functions (the
reference's own gradient mechanism) and central finite differences, both
taken in the tests.

How each function is pinned (golden values, brute force, closed forms,
library cross-checks) is recorded in the module docstrings and in
``tests/test_oracle_*.py``. Functions without such a check are marked
"parity unpinned" in their docstring.
"""
