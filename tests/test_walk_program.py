"""The post-order walk program of the four-state kernel (csrc/walk.hpp).

The program is compiled with g++ from the library's own header and executed
here as a stack machine: every internal node must find its children's values
in the slots the program names (tips read directly), each node must appear
once after both children, the root last, and the stack must stay within
``need`` <= floor(log2 T) + 1 entries.
"""
import math
import shutil
import subprocess

import numpy as np
import pytest

from torchtree_amd import synthetic
from torchtree_amd.build import CSRC

HERE = __import__("pathlib").Path(__file__).parent


@pytest.fixture(scope="module")
def walk_bin(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("walk") / "walk_dump"
    subprocess.run(["g++", "-O1", "-std=c++17", f"-I{CSRC}", str(HERE / "native" / "walk_dump.cpp"),
                    "-o", str(out)], check=True)
    return out


def _run(walk_bin, T, post):
    inp = f"{T}\n" + "\n".join(f"{v} {a} {b}" for v, a, b in post) + "\n"
    r = subprocess.run([str(walk_bin)], input=inp, capture_output=True, text=True, check=True)
    lines = [ln for ln in r.stdout.split("\n") if ln.strip()]
    need, nseg = (int(x) for x in lines[0].split())
    prog = [tuple(int(x) for x in ln.split()) for ln in lines[1:T]]
    segtab = [[int(x) for x in ln.split()] for ln in lines[T:]]
    assert len(segtab) == nseg
    return need, (prog, segtab)


def _check(T, post, need, walk):
    prog, segtab = walk
    kids = {v: (a, b) for v, a, b in post}
    assert len(prog) == T - 1
    assert prog[-1][0] == post[-1][0] == 2 * T - 2
    leaves = {t: frozenset([t]) for t in range(T)}  # value of a node = its leaf set
    stack = [None] * need
    done = set()
    seg = -1
    used = []
    prev = None
    for v, a, b, slots, tpos, sopen, kinds in prog:
        assert kids[v] == (a, b) and v not in done
        sa, sb, so = (slots & 0xFF) - 1, ((slots >> 8) & 0xFF) - 1, slots >> 16
        pa, pb = (tpos & 0xFF) - 1, ((tpos >> 8) & 0xFF) - 1
        if sopen >= 0:
            assert sopen == seg + 1  # segments open in order
            seg = sopen
        assert seg >= 0
        vals = []
        for ch, sl, pos, kind in ((a, sa, pa, kinds & 0xFF), (b, sb, pb, kinds >> 8)):
            # kind: 0 tip, 1 result of the previous entry, 2 stacked earlier result
            assert kind == (0 if ch < T else (1 if ch == prev else 2))
            if ch < T:
                assert sl == -1 and 0 <= pos < 16
                assert segtab[seg][pos] == ch  # the tip's codes are staged in that row
                used.append(ch)
                vals.append(leaves[ch])
            else:
                assert pos == -1
                assert 0 <= sl < need and ch in done
                assert stack[sl] == leaves[ch]  # the slot still holds that child's value
                vals.append(stack[sl])
        val = vals[0] | vals[1]
        assert 0 <= so < need
        # slots above the output are free again once the node is formed
        for s in range(so, need):
            stack[s] = None
        stack[so] = val
        leaves[v] = val
        done.add(v)
        prev = v
    assert leaves[2 * T - 2] == frozenset(range(T))
    assert sorted(used) == list(range(T))  # every tip staged and consumed once
    assert seg == len(segtab) - 1
    assert sorted(t for row in segtab for t in row if t >= 0) == list(range(T))
    assert need <= int(math.floor(math.log2(T))) + 1


def _caterpillar(T):
    post, nxt = [], T
    prev = 0
    for t in range(1, T):
        post.append((nxt, prev, t))
        prev = nxt
        nxt += 1
    return post


def _balanced(T):
    level = list(range(T))
    post, nxt = [], T
    while len(level) > 1:
        new = []
        for i in range(0, len(level) - 1, 2):
            post.append((nxt, level[i], level[i + 1]))
            new.append(nxt)
            nxt += 1
        if len(level) % 2:
            new.append(level[-1])
        level = new
    return post


@pytest.mark.parametrize("T", [2, 3, 5, 17, 64, 512, 1024])
def test_walk_random(walk_bin, T):
    rng = np.random.default_rng(T)
    topo = synthetic.random_topology(T, rng)
    post = [tuple(int(x) for x in r) for r in topo.postorder]
    need, prog = _run(walk_bin, T, post)
    _check(T, post, need, prog)


@pytest.mark.parametrize("T", [2, 9, 200])
def test_walk_caterpillar(walk_bin, T):
    post = _caterpillar(T)
    need, prog = _run(walk_bin, T, post)
    assert need == 1
    _check(T, post, need, prog)


@pytest.mark.parametrize("T", [4, 64, 256, 1000])
def test_walk_balanced(walk_bin, T):
    post = _balanced(T)
    need, prog = _run(walk_bin, T, post)
    if T & (T - 1) == 0:
        assert need == int(math.log2(T))  # perfectly balanced: log2 T entries
    _check(T, post, need, prog)
