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
    # drop bits: who consumes each node, and how
    consumed = {}
    for v, a, b, slots, tpos, sopen, kinds, drop in prog:
        for ch, sl, kind in ((a, (slots & 0xFF) - 1, kinds & 0xFF),
                             (b, ((slots >> 8) & 0xFF) - 1, kinds >> 8)):
            consumed[ch] = (kind, sl)
    for v, a, b, slots, tpos, sopen, kinds, drop in prog:
        cherry = a < T and b < T
        kind, sl = consumed.get(v, (None, None))
        in_regs = 2 <= sl < 5 if sl is not None else False
        # mode 0: only a cherry whose parent takes it as the previous result; mode 1 also
        # one waiting at a stack depth the post-order kernel keeps in registers (2..4)
        assert (drop & 1) == (1 if cherry and kind == 1 else 0), (v, drop)
        assert (drop >> 1) == (1 if cherry and (kind == 1 or (kind == 2 and in_regs)) else 0)
        assert drop & ~3 == 0
    for v, a, b, slots, tpos, sopen, kinds, drop in prog:
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


# ---------------------------------------------------------------- pre-order walk
@pytest.fixture(scope="module")
def pre_walk_bin(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("prewalk") / "pre_walk_dump"
    subprocess.run(["g++", "-O1", "-std=c++17", f"-I{CSRC}",
                    str(HERE / "native" / "pre_walk_dump.cpp"), "-o", str(out)], check=True)
    return out


def _run_pre(pre_walk_bin, T, post):
    inp = f"{T}\n" + "\n".join(f"{v} {a} {b}" for v, a, b in post) + "\n"
    r = subprocess.run([str(pre_walk_bin)], input=inp, capture_output=True, text=True, check=True)
    lines = [ln for ln in r.stdout.split("\n") if ln.strip()]
    return int(lines[0]), [tuple(int(x) for x in ln.split()) for ln in lines[1:]]


def _check_pre(T, post, depth, prog):
    """Structural check: every internal node is a parent once, root first; a
    continuation names a child of the previous entry; a pop finds its node on
    top of the stack at the depth it was pushed; depth <= floor(log2 T)."""
    kids = {v: (a, b) for v, a, b in post}
    root = post[-1][0]
    assert len(prog) == T - 1 and prog[0][0] == root and prog[0][3] == 0
    cherries = [v for v, a, b in post if a < T and b < T]  # rows T + i, in post-order
    seen, stack, nxt = set(), [], None
    for p, a, b, usrc, cont, push, tips, drop, row_a, row_b, a0, a1, b0, b1 in prog:
        # a dropped child (either register mode) is a cherry, rebuilt from its two tips,
        # whose codes are row T + its index among the cherries
        for h, (ch, row, g0, g1) in enumerate(((a, row_a, a0, a1), (b, row_b, b0, b1))):
            bits = (drop >> (2 * h)) & 3
            if ch < T:
                assert row == ch and bits == 0 and g0 == g1 == 0
            elif ch in cherries:
                assert row == T + cherries.index(ch) and (g0, g1) == kids[ch]
                assert bits in (0, 2, 3)  # mode 1 drops whatever mode 0 drops
            else:
                assert row == 0 and bits == 0 and g0 == g1 == 0
        assert drop & ~15 == 0
    for p, a, b, usrc, cont, push, tips, *_ in prog:
        assert kids[p] == (a, b) and p not in seen
        seen.add(p)
        if usrc == 1:
            assert nxt == p
        elif usrc != 0:
            assert usrc & 0xFF == 2 and nxt is None
            d = usrc >> 8
            assert stack and stack[-1] == p and len(stack) - 1 == d
            stack.pop()
        assert tips == (a < T) | ((b < T) << 1)
        internal = [h for h, x in enumerate((a, b)) if x >= T]
        if len(internal) == 2:
            assert push >= 0 and cont >= 0 and (push & 0xFF) == 1 - cont
            assert push >> 8 == len(stack)
            stack.append((a, b)[push & 0xFF])
        else:
            assert push == -1
            assert cont == (internal[0] if internal else -1)
        nxt = (a, b)[cont] if cont >= 0 else None
        assert len(stack) <= depth
    assert not stack and len(seen) == T - 1
    assert depth <= max(0, int(math.floor(math.log2(T))))


@pytest.mark.parametrize("T", [2, 3, 5, 17, 64, 512, 1024])
def test_pre_walk_random(pre_walk_bin, T):
    rng = np.random.default_rng(1000 + T)
    topo = synthetic.random_topology(T, rng)
    post = [tuple(int(x) for x in r) for r in topo.postorder]
    depth, prog = _run_pre(pre_walk_bin, T, post)
    _check_pre(T, post, depth, prog)


@pytest.mark.parametrize("shape,T", [("caterpillar", 2), ("caterpillar", 50), ("balanced", 64),
                                     ("balanced", 1000)])
def test_pre_walk_shapes(pre_walk_bin, shape, T):
    post = _caterpillar(T) if shape == "caterpillar" else _balanced(T)
    depth, prog = _run_pre(pre_walk_bin, T, post)
    _check_pre(T, post, depth, prog)
    if shape == "caterpillar":
        assert depth == 0
    elif T & (T - 1) == 0:
        assert depth == int(math.log2(T)) - 1


def _pre_walk_statistics(T, post, prog, tips, mats, freqs, props, weights, rs, mode=None):
    """The kernel's algorithm restated in FP64 numpy: post-order partials are
    overwritten in place by adjoints pushed at depth >= rs, shallower pushes
    live in `rs` register slots; returns d loglik / d mats [B, C, S, S].
    ``mode`` (0 or 1): the cherries whose drop bit of that register mode is set
    were never stored, and the walk rebuilds them from the two tips its entry
    names (reading such a slot fails)."""
    C, S = mats.shape[1], mats.shape[2]
    post_l = {}
    for t in range(T):
        post_l[t] = np.broadcast_to(tips[t], (C, S, tips.shape[-1]))
    for v, a, b in post:
        post_l[v] = np.einsum("cij,cjn->cin", mats[a], post_l[a]) * np.einsum(
            "cij,cjn->cin", mats[b], post_l[b])
    mem = {v: ("L", post_l[v]) for v in range(T, 2 * T - 1)}
    if mode is not None:
        for p, a, b, usrc, cont, push, tipf, drop, *_ in prog:
            for h, x in enumerate((a, b)):
                if (drop >> (2 * h + mode)) & 1:
                    mem[x] = ("never stored", None)
    regs = [None] * rs
    grad = np.zeros_like(mats)
    U = None
    for p, a, b, usrc, cont, push, tipf, drop, row_a, row_b, a0, a1, b0, b1 in prog:
        if usrc == 0:
            U = np.broadcast_to(freqs[:, None], (C, S, tips.shape[-1]))
        elif usrc != 1:
            d = usrc >> 8
            if d < rs:
                U = regs[d]
            else:
                kind, U = mem[p]
                assert kind == "U"  # the in-place slot holds the pushed adjoint
        L = []
        for h, (x, g) in enumerate(((a, (a0, a1)), (b, (b0, b1)))):
            if x < T:
                L.append(post_l[x])
            elif mode is not None and (drop >> (2 * h + mode)) & 1:
                # rebuilt from the tips the entry names (nothing was stored)
                L.append(np.einsum("cij,cjn->cin", mats[g[0]], post_l[g[0]]) *
                         np.einsum("cij,cjn->cin", mats[g[1]], post_l[g[1]]))
            else:
                kind, val = mem[x]
                assert kind == "L"  # read before any push overwrote it
                L.append(val)
        m = [np.einsum("cij,cjn->cin", mats[x], l) for x, l in zip((a, b), L)]
        lik = np.einsum("c,cin->n", props, U * m[0] * m[1])
        Ut = [U * m[1], U * m[0]]
        for h, x in enumerate((a, b)):
            grad[x] = np.einsum("n,c,cin,cjn->cij", weights / lik, props, Ut[h], L[h])

        def adjoint(h):
            un = np.einsum("cij,cin->cjn", mats[(a, b)[h]], Ut[h])
            e = np.frexp(un.max(axis=(0, 1)))[1]
            return np.ldexp(un, -e)

        if push >= 0:
            h, d = push & 0xFF, push >> 8
            if d < rs:
                regs[d] = adjoint(h)
            else:
                mem[(a, b)[h]] = ("U", adjoint(h))
        U = adjoint(cont) if cont >= 0 else None
    return grad


@pytest.mark.parametrize("mode", [None, 0, 1])
@pytest.mark.parametrize("T,rs", [(2, 2), (9, 0), (40, 1), (40, 2), (130, 2)])
def test_pre_walk_statistics_match_oracle(pre_walk_bin, T, rs, mode):
    """The walk's branch statistics equal the oracle's autograd gradient, also when
    the cherries a register mode drops are rebuilt from their tips."""
    import torch

    from oracle import pruning as orc

    rng = np.random.default_rng(7 * T + rs)
    topo = synthetic.random_topology(T, rng)
    post = [tuple(int(x) for x in r) for r in topo.postorder]
    _, prog = _run_pre(pre_walk_bin, T, post)
    C, S, N = 3, 4, 11
    B = 2 * T - 2
    mats = rng.dirichlet(np.ones(S), size=(B, C, S))  # rows sum to one
    freqs = rng.dirichlet(np.ones(S))
    props = rng.dirichlet(np.ones(C))
    tips = rng.uniform(0.05, 1.0, size=(T, S, N))
    weights = rng.integers(1, 5, size=N).astype(np.float64)
    got = _pre_walk_statistics(T, post, prog, tips, mats, freqs, props, weights, rs, mode)

    tm = torch.tensor(mats, requires_grad=True)
    parts = [torch.tensor(tips[t]) for t in range(T)] + [None] * (T - 1)
    ll = orc.treelik_discrete(parts, torch.tensor(weights), post, tm, torch.tensor(freqs)[None],
                              torch.tensor(props)[:, None, None])
    ll.sum().backward()
    want = tm.grad.numpy()
    np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-12)
