"""Drop-in replacements for torchtree's pruning functions, backed by HIP kernels.

Same names, arguments and return shapes as ``torchtree/evolution/tree_likelihood.py``
(torchtree 1.0.2):

* ``calculate_treelikelihood``                          (tree_likelihood.py:14-37)
* ``calculate_treelikelihood_discrete``                 (40-75)
* ``calculate_treelikelihood_tip_states_discrete``      (78-131)
* ``calculate_treelikelihood_discrete_safe``            (134-183)
* ``calculate_treelikelihood_discrete_rescaled``        (186-221)
* ``calculate_treelikelihood_tip_states_discrete_rescaled`` (224-278)

Differences a caller can observe:

* every node is rescaled by an exact power of two, so the plain, ``_safe``
  and ``_rescaled`` variants return the same value (the reference's
  variants agree to rounding; ``test_tree_likelihood.py:55-70``);
* the caller's ``partials`` list is not written to (the reference stores the
  internal partials into it);
* the result is differentiable with respect to ``mats``, ``freqs`` and
  ``props`` through a hand-written pre-order adjoint; the post-order
  workspace is consumed by the backward, so a graph can be backpropagated
  once (``retain_graph`` second passes raise);
* the tip-states variants read a tip of unknown state (code S) as P times a
  row of ones, which equals the reference's appended column of ones
  (``tree_likelihood.py:106-112``) because the rows of P sum to 1; d loglik /
  d mats therefore carries, for such tips, a component along each row's
  all-ones direction that the reference's does not. Every derivative through
  P(t) (branch lengths, rates, Q) is blind to that direction and agrees;
* CPU inputs are copied to the current HIP device and the result is
  returned on the input device. Without a HIP device these functions raise.
"""
from __future__ import annotations

import collections
import ctypes
import math
import weakref

import numpy as np
import torch

from . import _lib, timing
from ._lib import HX_F32, HX_F64, HX_TIPS_CODES, HX_TIPS_DENSE, HipTreeError


# ----------------------------------------------------------------- plans


class Plan:
    """One topology: post-order triples with device-resident level schedules."""

    def __init__(self, taxa: int, post_indexing, device: int | None = None):
        lib = _lib.load()
        arr = np.ascontiguousarray(np.asarray(post_indexing, dtype=np.int32).reshape(-1, 3))
        if arr.shape[0] != taxa - 1:
            raise HipTreeError(f"expected {taxa - 1} post-order triples, got {arr.shape[0]}")
        if device is None:
            device = torch.cuda.current_device()
        handle = ctypes.c_void_p()
        _lib.check(lib.hx_plan_create(taxa, arr.ctypes.data, device, ctypes.byref(handle)),
                   "hx_plan_create")
        self.handle = handle
        self.taxa = taxa
        self.device = device
        self.postorder = arr
        pl, dl, mw = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _lib.check(lib.hx_plan_info(handle, ctypes.byref(pl), ctypes.byref(dl), ctypes.byref(mw)))
        self.post_levels, self.pre_levels, self.max_width = pl.value, dl.value, mw.value

    def workspace_bytes(self, S, N, C, K, dtype_code, grad_mode=_lib.HX_GRAD_MATS) -> int:
        return int(_lib.load().hx_workspace_bytes(self.handle, S, N, C, K, dtype_code, grad_mode))

    def __del__(self):
        # at interpreter exit the module's globals may already be gone: the device memory
        # then goes with the process
        h = getattr(self, "handle", None)
        lib = getattr(_lib, "_lib", None) if _lib is not None else None
        if h is not None and lib is not None:
            lib.hx_plan_destroy(h)
            self.handle = None


_PLAN_CACHE: "collections.OrderedDict[tuple, Plan]" = collections.OrderedDict()


def get_plan(taxa: int, post_indexing, device: int) -> Plan:
    arr = np.asarray(post_indexing, dtype=np.int32).reshape(-1, 3)
    key = (taxa, device, arr.tobytes())
    plan = _PLAN_CACHE.get(key)
    if plan is None:
        plan = Plan(taxa, arr, device)
        _PLAN_CACHE[key] = plan
        while len(_PLAN_CACHE) > 64:
            _PLAN_CACHE.popitem(last=False)
    else:
        _PLAN_CACHE.move_to_end(key)
    return plan


# ------------------------------------------------------------------ tips


class TipData:
    """Device tip data: uint8 codes [T][N] plus a table of partial rows, or dense [T][N][S]."""

    def __init__(self, kind, tips, table, S, N, T):
        self.kind = kind
        self.tips = tips
        self.table = table
        self.S, self.N, self.T = S, N, T
        self.ncodes = 0 if table is None else table.shape[0]
        self._by_dtype = {}
        if kind == HX_TIPS_CODES and tips.numel() and int(tips.max()) >= self.ncodes:
            # the kernels index the table with the raw code
            raise HipTreeError(f"tip code {int(tips.max())} outside a table of {self.ncodes} rows")

    def table_as(self, dtype):
        if self.table is None:
            return None
        t = self._by_dtype.get(dtype)
        if t is None:
            t = self.table.to(dtype).contiguous()
            self._by_dtype[dtype] = t
        return t

    def dense_as(self, dtype):
        t = self._by_dtype.get(("dense", dtype))
        if t is None:
            t = self.tips.to(dtype).contiguous()
            self._by_dtype[("dense", dtype)] = t
        return t

    @classmethod
    def from_partials(cls, tips, device):
        """Tips [S, N] each; unique columns become a code table when there are <= 256."""
        stacked = torch.stack([t.to(torch.float64) for t in tips]).to(device)  # [T, S, N]
        T, S, N = stacked.shape
        rows = stacked.permute(0, 2, 1).reshape(T * N, S)
        uniq, inverse = torch.unique(rows, dim=0, return_inverse=True)
        if uniq.shape[0] <= 256:
            codes = inverse.to(torch.uint8).reshape(T, N).contiguous()
            return cls(HX_TIPS_CODES, codes, uniq.contiguous(), S, N, T)
        dense = stacked.permute(0, 2, 1).contiguous()  # [T, N, S]
        return cls(HX_TIPS_DENSE, dense, None, S, N, T)

    @classmethod
    def from_states(cls, tips, S, device):
        """Integer codes [N] each in 0..S (S = unknown state, all-ones row)."""
        codes = torch.stack([t.to(torch.int64) for t in tips]).to(device)
        if int(codes.min()) < 0 or int(codes.max()) > S:
            raise HipTreeError("tip state codes must lie in 0..S")
        table = torch.cat((torch.eye(S, dtype=torch.float64), torch.ones(1, S, dtype=torch.float64)))
        T, N = codes.shape
        return cls(HX_TIPS_CODES, codes.to(torch.uint8).contiguous(), table.to(device), S, N, T)


_TIP_CACHE: "collections.OrderedDict[tuple, tuple]" = collections.OrderedDict()


def _tip_key(tips, kind, device):
    return (kind, device) + tuple((id(t), t.data_ptr(), tuple(t.shape)) for t in tips)


def _tips_unchanged(refs, versions, tips):
    """The cached entry was built from these very tensor objects, unmodified since.

    An id or a storage address alone can be reused by a new tensor once the old
    one is freed; a weak reference cannot: it dies with its tensor. In-place
    writes (through the tensor or any view of it) bump ``_version``.
    """
    return all(r() is t and t._version == v for r, v, t in zip(refs, versions, tips))


def get_tips(tips, kind: str, S: int, device: int) -> TipData:
    key = _tip_key(tips, kind, device)
    hit = _TIP_CACHE.get(key)
    if hit is not None and _tips_unchanged(hit[0], hit[1], tips):
        _TIP_CACHE.move_to_end(key)
        return hit[2]
    if kind == "states":
        td = TipData.from_states(tips, S, device)
    else:
        td = TipData.from_partials(tips, device)
    _TIP_CACHE[key] = ([weakref.ref(t) for t in tips], [t._version for t in tips], td)
    _TIP_CACHE.move_to_end(key)
    while len(_TIP_CACHE) > 16:
        _TIP_CACHE.popitem(last=False)
    return td


# ---------------------------------------------------------- core function


PRODUCTS = ("fp32", "bf16x3")
_dense_products = "fp32"


def set_dense_products(mode: str) -> str:
    """How FP32 calls of the drop-in functions form their dense-state products (S != 4, at
    most 4 categories): "fp32" on the FP32 matrix cores, or "bf16x3", split BF16 (hi + lo
    halves of each operand, three products, FP32 accumulation). Returns the previous mode."""
    global _dense_products
    if mode not in PRODUCTS:
        raise HipTreeError(f"products must be one of {PRODUCTS}")
    prev, _dense_products = _dense_products, mode
    return prev


def _dtype_code(dtype, products=None):
    if dtype == torch.float64:
        return HX_F64
    if dtype == torch.float32:
        mode = _dense_products if products is None else products
        return _lib.HX_F32_BF16X3 if mode == "bf16x3" else HX_F32
    raise HipTreeError(f"unsupported dtype {dtype}")


class Pruning:
    """Bound inputs of one pruning call on the device (no autograd).

    ``products``: "fp32" or "bf16x3" for FP32 calls (dense states), or None to follow
    ``set_dense_products``."""

    def __init__(self, plan: Plan, tips: TipData, weights: torch.Tensor, products=None):
        self.plan = plan
        self.tips = tips
        self.weights = weights.to(torch.float64).contiguous()
        self.S, self.N, self.T = tips.S, tips.N, tips.T
        if products is not None and products not in PRODUCTS:
            raise HipTreeError(f"products must be one of {PRODUCTS}")
        self.products = products

    def _common(self, mats, freqs, props, dtype):
        K, B, C, S, _ = mats.shape
        code = _dtype_code(dtype, self.products)
        if self.tips.kind == HX_TIPS_CODES:
            tips_p, table_p = self.tips.tips, self.tips.table_as(dtype)
        else:
            tips_p, table_p = self.tips.dense_as(dtype), None
        return K, B, C, S, code, tips_p, table_p

    def _launches(self, which, S, C, code, K):
        """Kernel launches of one pass (the walks are one launch for the whole tree)."""
        if S == 4 and which == "post":
            return 2 + int(self.tips.kind == _lib.HX_TIPS_CODES)  # pad codes, walk, root sums
        if S == 4 and which == "pre" and code != _lib.HX_F64 and C <= 4:
            return 2  # walk, statistic reduction (mirrors use_pre_walk, csrc/layout.hpp)
        return (self.plan.post_levels if which == "post" else self.plan.pre_levels) + 1

    def forward(self, mats, freqs, props, want_root_grads=True):
        """mats [K,B,C,S,S]; freqs [K|1,S]; props [K|1,C] (all the same dtype, contiguous).

        Returns (loglik [K] f64, dfreqs [K,S] f64, dprops [K,C] f64, workspace).
        """
        lib = _lib.load()
        dtype = mats.dtype
        K, B, C, S, code, tips_p, table_p = self._common(mats, freqs, props, dtype)
        dev = mats.device
        mode = _lib.HX_GRAD_MATS if want_root_grads else _lib.HX_GRAD_NONE
        nbytes = self.plan.workspace_bytes(S, self.N, C, K, code, mode)
        ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        out = torch.empty(K, dtype=torch.float64, device=dev)
        dfreqs = torch.empty(K, S, dtype=torch.float64, device=dev) if want_root_grads else None
        dprops = torch.empty(K, C, dtype=torch.float64, device=dev) if want_root_grads else None
        with timing.phase("postorder", self._launches("post", S, C, code, K)):
            rc = lib.hx_loglik(
            self.plan.handle, S, self.N, C, K, code, self.tips.kind,
            _lib.ptr(tips_p), _lib.ptr(table_p), self.tips.ncodes,
            _lib.ptr(self.weights), _lib.ptr(mats),
            _lib.ptr(freqs), int(freqs.shape[0] == K and K > 1),
            _lib.ptr(props), int(props.shape[0] == K and K > 1),
            _lib.ptr(ws), nbytes, _lib.ptr(out), _lib.ptr(dfreqs), _lib.ptr(dprops),
            _lib.stream_handle(dev))
        _lib.check(rc, "hx_loglik")
        return out, dfreqs, dprops, ws

    def backward_mats(self, mats, freqs, props, ws):
        """d loglik / d mats [K,B,C,S,S] f64; consumes ``ws``."""
        lib = _lib.load()
        dtype = mats.dtype
        K, B, C, S, code, tips_p, table_p = self._common(mats, freqs, props, dtype)
        dev = mats.device
        dmats = torch.empty(K, B, C, S, S, dtype=torch.float64, device=dev)
        with timing.phase("preorder", self._launches("pre", S, C, code, K)):
            rc = lib.hx_loglik_grad_mats(
            self.plan.handle, S, self.N, C, K, code, self.tips.kind,
            _lib.ptr(tips_p), _lib.ptr(table_p), self.tips.ncodes,
            _lib.ptr(self.weights), _lib.ptr(mats),
            _lib.ptr(freqs), int(freqs.shape[0] == K and K > 1),
            _lib.ptr(props), int(props.shape[0] == K and K > 1),
            _lib.ptr(ws), ws.numel(), _lib.ptr(dmats), _lib.stream_handle(dev))
        _lib.check(rc, "hx_loglik_grad_mats")
        return dmats


class _PruneFn(torch.autograd.Function):
    """loglik [K] (f64) of canonical inputs mats [K,B,C,S,S], freqs [K|1,S], props [K|1,C]."""

    @staticmethod
    def forward(ctx, mats, freqs, props, pruning: Pruning):
        need = ctx.needs_input_grad[0] or ctx.needs_input_grad[1] or ctx.needs_input_grad[2]
        out, dfreqs, dprops, ws = pruning.forward(mats, freqs, props, want_root_grads=need)
        if need:
            ctx.pruning = pruning
            ctx.ws = ws
            ctx.dfreqs = dfreqs
            ctx.dprops = dprops
            ctx.save_for_backward(mats, freqs, props)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        mats, freqs, props = ctx.saved_tensors
        if ctx.ws is None:
            raise HipTreeError("the pruning workspace was consumed by an earlier backward")
        K = mats.shape[0]
        g = grad_out.to(torch.float64)
        dm = df = dp = None
        if ctx.needs_input_grad[0]:
            dmats = ctx.pruning.backward_mats(mats, freqs, props, ctx.ws)
            ctx.ws = None
            dm = (dmats * g.view(K, 1, 1, 1, 1)).to(mats.dtype)
        if ctx.needs_input_grad[1]:
            df = ctx.dfreqs * g.view(K, 1)
            if freqs.shape[0] != K or K == 1:
                df = df.sum(0, keepdim=True)
            df = df.to(freqs.dtype)
        if ctx.needs_input_grad[2]:
            dp = ctx.dprops * g.view(K, 1)
            if props.shape[0] != K or K == 1:
                dp = dp.sum(0, keepdim=True)
            dp = dp.to(props.dtype)
        return dm, df, dp, None


def _device_for(*tensors):
    for t in tensors:
        if isinstance(t, torch.Tensor) and t.is_cuda:
            return t.device
    if not torch.cuda.is_available():
        raise HipTreeError("no HIP device: the hiptree engine has no CPU path")
    return torch.device("cuda", torch.cuda.current_device())


def prune(tips, tip_kind, weights, post_indexing, mats, freqs, props):
    """Shared body: mats [..., B, C, S, S], freqs [..., 1, S], props [..., C, 1, 1] -> [..., 1]."""
    in_dev = mats.device
    dev = _device_for(mats, freqs, props, weights)
    dtype = mats.dtype
    if dtype not in (torch.float32, torch.float64):
        raise HipTreeError(f"unsupported dtype {dtype}")
    B, C, S = mats.shape[-4], mats.shape[-3], mats.shape[-1]
    T = len(post_indexing) + 1
    if B != 2 * T - 2:
        raise HipTreeError(f"mats has {B} branches, expected {2 * T - 2}")
    batch = torch.broadcast_shapes(mats.shape[:-4], freqs.shape[:-2], props.shape[:-3])
    K = int(math.prod(batch)) if len(batch) else 1
    mats_c = mats.to(dev).expand(batch + mats.shape[-4:]).reshape(K, B, C, S, S).contiguous()
    fb = freqs.shape[:-2]
    if len(fb) and math.prod(fb) > 1:
        freqs_c = freqs.to(dev, dtype).expand(batch + freqs.shape[-2:]).reshape(K, S)
    else:
        freqs_c = freqs.to(dev, dtype).reshape(1, S)
    pb = props.shape[:-3]
    if len(pb) and math.prod(pb) > 1:
        props_c = props.to(dev, dtype).expand(batch + props.shape[-3:]).reshape(K, C)
    else:
        props_c = props.to(dev, dtype).reshape(1, C)
    device_index = dev.index if dev.index is not None else torch.cuda.current_device()
    with torch.cuda.device(device_index):
        plan = get_plan(T, post_indexing, device_index)
        td = get_tips(tips, tip_kind, S, device_index)
        if td.N != weights.shape[-1]:
            raise HipTreeError("weights and tip partials disagree on the pattern count")
        pr = Pruning(plan, td, weights.to(dev))
        out = _PruneFn.apply(mats_c, freqs_c.contiguous(), props_c.contiguous(), pr)
    out = out.to(dtype).reshape(batch + (1,))
    return out.to(in_dev) if in_dev != dev else out


def _split(partials, post_indexing):
    T = len(post_indexing) + 1
    return list(partials[:T])


def calculate_treelikelihood(partials, weights, post_indexing, mats, freqs):
    """Plain pruning (tree_likelihood.py:14-37): mats [..., B, S, S], freqs [..., 1, S] -> [..., 1]."""
    S = mats.shape[-1]
    props = torch.ones((1, 1, 1), dtype=mats.dtype, device=mats.device)
    return prune(_split(partials, post_indexing), "partials", weights, post_indexing,
                 mats.unsqueeze(-3), freqs, props)


def calculate_treelikelihood_discrete(partials, weights, post_indexing, mats, freqs, props):
    """Rate categories (tree_likelihood.py:40-75)."""
    return prune(_split(partials, post_indexing), "partials", weights, post_indexing, mats,
                 freqs, props)


def calculate_treelikelihood_tip_states_discrete(partials, weights, post_indexing, mats, freqs,
                                                 props):
    """Tip codes [N] in 0..S (tree_likelihood.py:78-131)."""
    return prune(_split(partials, post_indexing), "states", weights, post_indexing, mats,
                 freqs, props)


def calculate_treelikelihood_discrete_safe(partials, weights, post_indexing, mats, freqs, props,
                                           threshold):
    """Selective rescaling in the reference (tree_likelihood.py:134-183); every node here.

    ``threshold`` is accepted for signature compatibility: power-of-two
    rescaling of every node makes the threshold test unnecessary.
    """
    return prune(_split(partials, post_indexing), "partials", weights, post_indexing, mats,
                 freqs, props)


def calculate_treelikelihood_discrete_rescaled(partials, weights, post_indexing, mats, freqs,
                                               props):
    """Per-node rescaling (tree_likelihood.py:186-221)."""
    return prune(_split(partials, post_indexing), "partials", weights, post_indexing, mats,
                 freqs, props)


def calculate_treelikelihood_tip_states_discrete_rescaled(partials, weights, post_indexing, mats,
                                                          freqs, props):
    """Tip codes plus rescaling (tree_likelihood.py:224-278)."""
    return prune(_split(partials, post_indexing), "states", weights, post_indexing, mats,
                 freqs, props)
