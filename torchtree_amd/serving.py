"""Continuous-batching loglik + gradient serving for many resident jobs on one GPU.

Jobs (alignment tips, weights, topology plan, FP64 eigensystem, site model)
are registered once and stay resident in HBM (``hx_job_create``); a
registration that would exceed the resident share of the HBM budget is
refused. Requests (job, K branch-length samples) arrive over time; the
scheduler admits them first-come first-served onto a fixed set of streams as
streams and their workspaces free up, so the work of requests that arrive
while others run overlaps on the device. Each request runs as one
``hx_job_eval`` on its stream: P matrices, then post-/pre-order over pattern
slices sized at registration so that the largest request fits the stream's
workspace. A request's result depends only on its own inputs (fixed slice
and chunk order, no cross-request reductions), so it is bitwise the same
alone and in any mix.

The executor is pluggable: ``GpuExecutor`` drives the library on HIP
streams; ``FakeExecutor`` simulates completion times on the CPU so the
scheduler's admission, budget and ordering logic is testable without a GPU.
The reference has no serving layer; the per-request computation is
``TreeLikelihoodModel._call`` (``evolution/tree_likelihood.py:313-356``) and
its branch-length gradient.
"""
from __future__ import annotations

import collections
import ctypes
import dataclasses
import math
import time

import numpy as np
import torch

from . import _lib
from ._lib import HX_F32, HX_F64, HipTreeError

F64 = torch.float64


class BudgetError(HipTreeError):
    pass


def _host_eigen(q_norm: torch.Tensor, freqs: torch.Tensor):
    q = q_norm.detach().to("cpu", F64)
    pi = freqs.detach().to("cpu", F64)
    sq = pi.sqrt()
    sym = sq[:, None] * q / sq[None, :]
    lam, vec = torch.linalg.eigh(0.5 * (sym + sym.t()))
    left = vec / sq[:, None]
    right = vec.t() * sq[None, :]
    return (np.ascontiguousarray(lam.numpy()), np.ascontiguousarray(left.numpy()),
            np.ascontiguousarray(right.numpy()))


class ResidentJob:
    """One alignment resident on the device (``hx_job_create``)."""

    def __init__(self, name, post_indexing, codes, table, weights, q_norm, freqs, rates, props,
                 dtype=torch.float32, slice_patterns=0, device=None, kind="dna", products="fp32",
                 general=False):
        """``products``: for FP32 jobs, "fp32" or "bf16x3" (dense-state products as split
        BF16 on the matrix cores, FP32 accumulation). ``general``: ``q_norm`` is a
        non-reversible normalised rate matrix (``hx_job_create_general``: P = exp(Q tau) per
        request, at most 32 states) instead of one reversible with respect to ``freqs``."""
        from .likelihood import get_plan

        lib = _lib.load()
        if device is None:
            device = torch.cuda.current_device()
        self.name, self.kind, self.dtype, self.device = name, kind, dtype, device
        self.products = products
        codes = np.ascontiguousarray(np.asarray(codes, dtype=np.uint8))
        self.T, self.N = codes.shape
        self.B = 2 * self.T - 2
        table = np.ascontiguousarray(np.asarray(table, dtype=np.float64))
        self.S = table.shape[1]
        rates = np.ascontiguousarray(np.asarray(rates, dtype=np.float64).reshape(-1))
        props = np.ascontiguousarray(np.asarray(props, dtype=np.float64).reshape(-1))
        self.C = rates.shape[0]
        weights = np.ascontiguousarray(np.asarray(weights, dtype=np.float64))
        self.general = bool(general)
        if self.general:
            qg = np.ascontiguousarray(torch.as_tensor(q_norm).detach().to("cpu", F64).numpy())
        else:
            lam, left, right = _host_eigen(torch.as_tensor(q_norm), torch.as_tensor(freqs))
        fr = np.ascontiguousarray(np.asarray(freqs, dtype=np.float64))
        self.plan = get_plan(self.T, post_indexing, device)
        h = ctypes.c_void_p()
        code = HX_F64 if dtype == F64 else (_lib.HX_F32_BF16X3 if products == "bf16x3" else HX_F32)
        prev = torch.cuda.current_device()
        torch.cuda.set_device(device)
        try:
            if self.general:
                rc = lib.hx_job_create_general(
                    self.plan.handle, self.S, self.N, self.C, code, int(slice_patterns),
                    codes.ctypes.data, table.ctypes.data, table.shape[0], weights.ctypes.data,
                    qg.ctypes.data, rates.ctypes.data, props.ctypes.data, fr.ctypes.data, device,
                    ctypes.byref(h))
            else:
                rc = lib.hx_job_create(self.plan.handle, self.S, self.N, self.C, code,
                                       int(slice_patterns), codes.ctypes.data, table.ctypes.data,
                                       table.shape[0], weights.ctypes.data, lam.ctypes.data,
                                       left.ctypes.data, right.ctypes.data, rates.ctypes.data,
                                       props.ctypes.data, fr.ctypes.data, device, ctypes.byref(h))
        finally:
            torch.cuda.set_device(prev)
        _lib.check(rc, "hx_job_create_general" if self.general else "hx_job_create")
        self.handle = h
        self.slice = min(slice_patterns, self.N) if slice_patterns > 0 else self.N  # as hx_job_create
        self.resident_bytes = int(lib.hx_job_resident_bytes(h))

    def workspace_bytes(self, K, grad=True):
        return int(_lib.load().hx_job_workspace_bytes(self.handle, K, int(grad)))

    def updates(self, K, grad=True):
        per = self.N * self.C * K
        return (self.T - 1) * per + ((2 * self.T - 2) * per if grad else 0)

    def eval(self, bls_dev, out_loglik, out_dbls, ws, stream_handle):
        K = bls_dev.shape[0]
        rc = _lib.load().hx_job_eval(self.handle, K, _lib.ptr(bls_dev), _lib.ptr(out_loglik),
                                     _lib.ptr(out_dbls), _lib.ptr(ws), ws.numel(), stream_handle)
        _lib.check(rc, "hx_job_eval")

    def model_workspace_bytes(self, K, eigen_batched=False):
        return int(_lib.load().hx_job_model_workspace_bytes(self.handle, K, int(eigen_batched)))

    def eval_model(self, bls_dev, eig, rates, props, freqs, ws, stream_handle):
        """One request with its own model (hx_job_eval_model).

        ``eig``: ``models.Eigen`` of the request's normalised Q (device FP64); ``rates``,
        ``props`` [R, C] and ``freqs`` [R, S] device FP64 with R = K or 1. Returns a dict of
        device FP64 tensors: loglik [K], dbls [K, B], drates [K, C], dq [E, S, S] (through P),
        dfreqs [K, S] (root), dprops [K, C].
        """
        K = bls_dev.shape[0]
        dev = bls_dev.device
        R = rates.shape[0]
        if R not in (1, K) or props.shape[0] != R or freqs.shape[0] != R:
            raise HipTreeError("rates, props and freqs need the same leading size, 1 or K")
        if eig.batched and eig.E != K:
            raise HipTreeError("a batched eigensystem needs one entry per sample")
        out = {"loglik": torch.empty(K, dtype=F64, device=dev),
               "dbls": torch.empty(K, self.B, dtype=F64, device=dev),
               "drates": torch.empty(K, self.C, dtype=F64, device=dev),
               "dq": torch.empty(eig.E, self.S, self.S, dtype=F64, device=dev),
               "dfreqs": torch.empty(K, self.S, dtype=F64, device=dev),
               "dprops": torch.empty(K, self.C, dtype=F64, device=dev)}
        rc = _lib.load().hx_job_eval_model(
            self.handle, K, _lib.ptr(bls_dev), int(eig.batched), _lib.ptr(eig.lam),
            _lib.ptr(eig.left), _lib.ptr(eig.right), _lib.ptr(rates), _lib.ptr(props),
            _lib.ptr(freqs), int(R == K and K > 1), _lib.ptr(out["loglik"]), _lib.ptr(out["dbls"]),
            _lib.ptr(out["drates"]), _lib.ptr(out["dq"]), _lib.ptr(out["dfreqs"]),
            _lib.ptr(out["dprops"]), _lib.ptr(ws), ws.numel(), stream_handle)
        _lib.check(rc, "hx_job_eval_model")
        return out

    def __del__(self):
        # at interpreter exit the module's globals may already be gone: the device memory
        # then goes with the process
        h = getattr(self, "handle", None)
        lib = getattr(_lib, "_lib", None) if _lib is not None else None
        if h is not None and lib is not None:
            lib.hx_job_destroy(h)
            self.handle = None


class _ServedModelFn(torch.autograd.Function):
    """loglik [K] of a resident job under a request's own model, with every gradient formed
    in the one device call (the backward only scales them by the incoming gradient)."""

    @staticmethod
    def forward(ctx, bls, rates, props, q_norm, freqs, job):
        from .models import Eigen

        dev = torch.device("cuda", job.device)
        K = bls.shape[0]
        eig = Eigen(q_norm, freqs, dev)
        if not eig.batched and K > 1:
            # one copy per sample, so that d/dQ comes back per sample and the backward can
            # weight each sample's share by its incoming gradient
            eig.lam, eig.left, eig.right = (x.expand(K, *x.shape[1:]).contiguous()
                                            for x in (eig.lam, eig.left, eig.right))
            eig.batched, eig.E = True, K
        fr = freqs.detach().to(dev, F64).reshape(-1, job.S).contiguous()
        r = rates.detach().to(dev, F64).reshape(-1, job.C).contiguous()
        p = props.detach().to(dev, F64).reshape(-1, job.C).contiguous()
        R = max(fr.shape[0], r.shape[0], p.shape[0])
        fr, r, p = (x.expand(R, -1).contiguous() for x in (fr, r, p))
        ws = torch.empty(job.model_workspace_bytes(K, eig.batched), dtype=torch.uint8, device=dev)
        out = job.eval_model(bls.detach().to(dev, F64).contiguous(), eig, r, p, fr, ws,
                             torch.cuda.current_stream(dev).cuda_stream)
        ctx.out = out
        ctx.shapes = (rates.shape, props.shape, q_norm.shape, freqs.shape)
        ctx.devices = (bls.device, rates.device, props.device, q_norm.device, freqs.device)
        return out["loglik"].to(bls.device)

    @staticmethod
    def backward(ctx, g):
        o = ctx.out
        K = o["loglik"].shape[0]
        g = g.to(o["loglik"].device, F64)
        rs, ps, qs, fs = ctx.shapes
        db = o["dbls"] * g.view(K, 1)

        def fold(t, shape):  # per-sample gradient onto a parameter of leading size K or 1
            t = t * g.view(K, *([1] * (t.dim() - 1)))
            return t.reshape(shape) if t.numel() == math.prod(shape) else t.sum(0).reshape(shape)

        dr = fold(o["drates"], rs)
        dp = fold(o["dprops"], ps)
        df = fold(o["dfreqs"], fs)
        dq = o["dq"] * g.view(-1, 1, 1)  # [K, S, S]: per sample (K = 1 included)
        dq = dq.reshape(qs) if dq.numel() == math.prod(qs) else dq.sum(0).reshape(qs)
        devs = ctx.devices
        return (db.to(devs[0]), dr.to(devs[1]), dp.to(devs[2]), dq.to(devs[3]), df.to(devs[4]),
                None)


def served_loglik(job, bls, rates, props, q_norm, freqs):
    """loglik [K] of resident ``job`` for branch lengths [K, B] under a per-request model:
    normalised Q [S, S] or [K, S, S], frequencies [S] or [K, S], category rates and
    proportions [C] or [K, C]. Differentiable with respect to all five through torch autograd,
    so gradients reach whatever parameters built Q (``models.gtr_q`` and the like)."""
    return _ServedModelFn.apply(bls, rates, props, q_norm, freqs, job)


@dataclasses.dataclass
class Request:
    rid: int
    job: object
    bls: torch.Tensor          # [K, B] float64 (host, pinned when possible)
    arrival: float = 0.0       # seconds from the start of the run
    grad: bool = True
    loglik: torch.Tensor | None = None
    dbls: torch.Tensor | None = None
    t_start: float = math.nan
    t_done: float = math.nan
    status: str = "new"        # new, queued, running, done, rejected, cancelled
    cancelled: bool = False    # set by Server.cancel; a queued request is then dropped
    bypassed: int = 0          # later requests admitted while this one waited for memory
    graphed: bool = False      # served by replaying a captured graph
    # A request's own model (hx_job_eval_model) instead of the job's registered one:
    # {"eig": models.Eigen on the job's device, "rates", "props" [R, C], "freqs" [R, S]}
    # device FP64 with R = 1 or K. Its results then also fill ``grads`` (drates, dq, dfreqs,
    # dprops, as ResidentJob.eval_model returns them).
    model: dict | None = None
    grads: dict | None = None

    @property
    def K(self):
        return self.bls.shape[0]

    def workspace_bytes(self):
        if self.model is not None:
            return self.job.model_workspace_bytes(self.K, self.model["eig"].batched)
        return self.job.workspace_bytes(self.K, self.grad)

    def run_on(self, bls_dev, ws, stream):
        """Evaluate on ``stream`` into fresh device outputs (``loglik``, ``dbls``, ``grads``)."""
        dev = bls_dev.device
        if self.model is not None:
            m = self.model
            out = self.job.eval_model(bls_dev, m["eig"], m["rates"], m["props"], m["freqs"], ws,
                                      stream.cuda_stream)
            self.loglik, self.dbls, self.grads = out["loglik"], out["dbls"], out
            return
        self.loglik = torch.empty(self.K, dtype=F64, device=dev)
        self.dbls = torch.empty(self.K, self.job.B, dtype=F64, device=dev) if self.grad else None
        self.job.eval(bls_dev, self.loglik, self.dbls, ws, stream.cuda_stream)

    @property
    def latency(self):
        return self.t_done - self.arrival

    def nonfinite(self):
        """Per-sample status flags: True where the log-likelihood or its gradient is NaN or
        infinite (the result is returned as computed; nothing is raised)."""
        bad = ~torch.isfinite(self.loglik)
        if self.dbls is not None:
            bad = bad | ~torch.isfinite(self.dbls).all(-1)
        return bad


class WorkspacePool:
    """First-fit allocator of request workspaces inside one device buffer.

    Offsets are 256-byte aligned; freed ranges coalesce with their neighbours. The server
    admits a request only when its workspace can be carved from here, so the device memory
    held by in-flight requests never exceeds the pool.
    """

    ALIGN = 256

    def __init__(self, nbytes, buf=None):
        self.size = int(nbytes) // self.ALIGN * self.ALIGN
        self.buf = buf
        self.free = [(0, self.size)]
        self.in_use = 0
        self.peak = 0

    def _round(self, n):
        return (int(n) + self.ALIGN - 1) // self.ALIGN * self.ALIGN

    def alloc(self, n):
        n = self._round(max(n, 1))
        for i, (off, sz) in enumerate(self.free):
            if sz >= n:
                if sz == n:
                    del self.free[i]
                else:
                    self.free[i] = (off + n, sz - n)
                self.in_use += n
                self.peak = max(self.peak, self.in_use)
                return off
        return None

    def release(self, off, n):
        n = self._round(max(n, 1))
        self.in_use -= n
        free = self.free
        i = 0
        while i < len(free) and free[i][0] < off:
            i += 1
        free.insert(i, (off, n))
        if i + 1 < len(free) and free[i][0] + free[i][1] == free[i + 1][0]:
            free[i] = (free[i][0], free[i][1] + free[i + 1][1])
            del free[i + 1]
        if i > 0 and free[i - 1][0] + free[i - 1][1] == free[i][0]:
            free[i - 1] = (free[i - 1][0], free[i - 1][1] + free[i][1])
            del free[i]

    def view(self, off, n):
        return None if self.buf is None else self.buf[off: off + int(n)]


class GpuExecutor:
    """Runs requests on HIP streams of one device (default stream priority).

    With ``graph_max_bytes`` > 0, a request whose workspace is at most that many bytes is
    served by replaying a captured HIP graph of its job's launch sequence (keyed by job, K and
    gradient flag), which replaces tens of kernel launches by one. A graph owns fixed input,
    output and workspace buffers; a replay copies the request's branch lengths in and its
    results out, on the request's stream. One replay of a graph is in flight at a time; a
    request whose graph is busy runs eagerly. The graphs' workspaces come from the pool.
    """

    def __init__(self, device, n_streams, graph_max_bytes=0, max_graphs=64):
        self.device = torch.device("cuda", device)
        self.n = n_streams
        self.streams = [torch.cuda.Stream(self.device) for _ in range(n_streams)]
        self.events = [None] * n_streams
        self.graph_max_bytes = int(graph_max_bytes)
        self.max_graphs = max_graphs
        self.graphs = collections.OrderedDict()  # key -> dict(graph, bls, ll, db, ws_off, busy)
        self.capture_stream = torch.cuda.Stream(self.device) if graph_max_bytes else None
        self.pool = None

    def make_pool(self, nbytes):
        self.pool = WorkspacePool(nbytes, torch.empty(int(nbytes), dtype=torch.uint8,
                                                      device=self.device))
        return self.pool

    # ---- graphs
    def graph_for(self, req):
        """The captured graph that will serve ``req`` (captured now on first sight of its job,
        K and gradient flag), or None when the request runs eagerly."""
        if not self.graph_max_bytes or req.model is not None:
            return None
        need = req.workspace_bytes()
        if need > self.graph_max_bytes:
            return None
        key = (id(req.job), req.K, req.grad)
        g = self.graphs.get(key)
        if g is not None:
            if g["busy"]:
                return None
            self.graphs.move_to_end(key)
            return g
        while len(self.graphs) >= self.max_graphs:
            old = next((k for k, v in self.graphs.items() if not v["busy"]), None)
            if old is None:
                return None
            ent = self.graphs.pop(old)
            self.pool.release(ent["ws_off"], ent["ws_n"])
        off = self.pool.alloc(need)
        if off is None:
            return None
        job = req.job
        ent = {"bls": torch.zeros(req.K, job.B, dtype=F64, device=self.device),
               "ll": torch.empty(req.K, dtype=F64, device=self.device),
               "db": torch.empty(req.K, job.B, dtype=F64, device=self.device) if req.grad else None,
               "ws_off": off, "ws_n": need, "busy": False, "job": job}
        ws = self.pool.view(off, need)
        st = self.capture_stream
        st.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(st):
            job.eval(ent["bls"], ent["ll"], ent["db"], ws, st.cuda_stream)  # first run eagerly
        st.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=st):
            job.eval(ent["bls"], ent["ll"], ent["db"], ws, st.cuda_stream)
        ent["graph"] = graph
        self.graphs[key] = ent
        return ent

    def start(self, slot, req: Request, ws_view, ent=None):
        st = self.streams[slot]
        with torch.cuda.stream(st):
            bls = req.bls.to(self.device, non_blocking=True)
            if ent is not None:
                ent["busy"] = True
                ent["bls"].copy_(bls)
                ent["graph"].replay()
                req.loglik = ent["ll"].clone()
                req.dbls = ent["db"].clone() if req.grad else None
                req.graphed = True
            else:
                req.run_on(bls, ws_view, st)
            ev = torch.cuda.Event()
            ev.record(st)
        self.events[slot] = (ev, bls, ent)

    def drop_job(self, job):
        """Forget the graphs captured for ``job``; their buffers go back to the pool."""
        for key in [k for k, v in self.graphs.items() if v["job"] is job]:
            ent = self.graphs.pop(key)
            if ent["busy"]:
                torch.cuda.synchronize(self.device)
            self.pool.release(ent["ws_off"], ent["ws_n"])

    def done(self, slot):
        ev = self.events[slot]
        if ev is None:
            return True
        if ev[0].query():
            if ev[2] is not None:
                ev[2]["busy"] = False
            self.events[slot] = None
            return True
        return False

    def wait_all(self):
        torch.cuda.synchronize(self.device)


class FakeExecutor:
    """CPU stand-in: a request finishes ``cost(req)`` seconds after it starts.

    ``order`` may reorder completions: when several requests are due, ``order(slots)`` picks
    which slot reports first (randomised completion order in the scheduler tests)."""

    def __init__(self, n_streams, cost, clock, order=None):
        self.n = n_streams
        self.cost = cost
        self.clock = clock
        self.order = order
        self.busy_until = [None] * n_streams
        self.log = []
        self.pool = None

    def make_pool(self, nbytes):
        self.pool = WorkspacePool(nbytes)
        return self.pool

    def graph_for(self, req):
        return None

    def drop_job(self, job):
        pass

    def start(self, slot, req, ws_view, ent=None):
        self.busy_until[slot] = self.clock() + self.cost(req)
        self.log.append(("start", slot, req.rid))

    def done(self, slot):
        t = self.busy_until[slot]
        if t is None:
            return True
        if self.clock() >= t:
            if self.order is not None:
                due = [i for i, u in enumerate(self.busy_until) if u is not None and self.clock() >= u]
                if self.order(due) != slot:
                    return False
            self.busy_until[slot] = None
            self.log.append(("done", slot))
            return True
        return False

    def wait_all(self):
        pass


class Server:
    """Continuous batching of loglik + gradient requests for resident jobs on one GPU.

    ``hbm_budget``: bytes this server may hold on the device: the resident job data plus a
    workspace pool of ``pool_bytes`` (default 80 % of the budget). A request is admitted when
    a stream is free and its workspace can be carved from the pool; it gives the memory back
    when it completes. Dispatch is first come first served with backfill: when the oldest
    waiting request does not fit the free memory, later requests that fit go first, until
    ``max_bypass`` of them have passed it; then nothing more starts until it fits. A request
    larger than ``max_request_bytes`` (default a quarter of the pool, the size pattern slices
    are chosen for) or than the whole pool is rejected. Requests can be cancelled while they
    wait. Each request's result depends only on its own inputs, so it is bitwise the same
    alone and in any mix.
    """

    def __init__(self, hbm_budget, n_streams=8, pool_bytes=None, executor=None, device=0,
                 clock=time.perf_counter, max_bypass=None, max_request_bytes=None,
                 graph_max_bytes=0):
        self.budget = int(hbm_budget)
        self.n_streams = n_streams
        pool_bytes = int(0.8 * self.budget) if pool_bytes is None else int(pool_bytes)
        self.resident_cap = self.budget - pool_bytes
        if self.resident_cap <= 0 or pool_bytes <= 0:
            raise BudgetError("HBM budget leaves no room for resident jobs and a workspace pool")
        self.max_request_bytes = int(max_request_bytes or pool_bytes // 4)
        self.resident = 0
        self.jobs = {}
        self.clock = clock
        self.executor = executor if executor is not None else GpuExecutor(
            device, n_streams, graph_max_bytes=graph_max_bytes)
        self.pool = self.executor.make_pool(pool_bytes)
        self.max_bypass = 2 * n_streams if max_bypass is None else max_bypass
        self.rejected = 0           # job registrations refused
        self.counters = collections.Counter()

    def slice_for(self, T, N, C, S, dtype, max_k=32):
        """Largest pattern slice whose max-K gradient workspace fits ``max_request_bytes``.

        Sizes come from the library (hx_job_workspace_bytes_for), so they match what a
        registered job will ask for. The size is not monotone in the slice everywhere (the
        four-state gradient switches kernels with the workgroup count), so this bisects
        keeping a slice that fits."""
        lib = _lib.load()
        code = _lib.HX_F64 if dtype == F64 else _lib.HX_F32
        cap = self.max_request_bytes

        def need(n):
            return int(lib.hx_job_workspace_bytes_for(T, S, n, C, max_k, code, 1))

        if need(1) > cap:
            raise BudgetError("one pattern of this job does not fit a request workspace")
        if need(N) <= cap:
            return N
        lo, hi = 1, N
        while hi - lo > 1:
            mid = (lo + hi) // 2
            if need(mid) <= cap:
                lo = mid
            else:
                hi = mid
        return lo

    def register(self, name, job):
        if self.resident + job.resident_bytes > self.resident_cap:
            self.rejected += 1
            raise BudgetError(f"job {name}: resident data would exceed the HBM budget")
        self.resident += job.resident_bytes
        self.jobs[name] = job
        return job

    def unregister(self, name):
        """Remove a job (no request of it may be queued or running)."""
        job = self.jobs.pop(name)
        self.executor.drop_job(job)
        self.resident -= job.resident_bytes

    @staticmethod
    def cancel(req):
        """Withdraw a request: dropped if it has not started; a running one completes and is
        reported as cancelled."""
        req.cancelled = True

    def _dispatch(self, queue, slots, t0):
        free = [i for i, s in enumerate(slots) if s is None]
        head = queue[0] if queue else None
        for r in list(queue):
            if not free:
                break
            if r.cancelled:
                queue.remove(r)
                r.status = "cancelled"
                self.counters["cancelled"] += 1
                continue
            need = r.workspace_bytes()
            if need > min(self.max_request_bytes, self.pool.size):
                queue.remove(r)
                r.status = "rejected"
                self.counters["rejected"] += 1
                continue
            off = None
            ent = self.executor.graph_for(r)
            if ent is None:
                off = self.pool.alloc(need)
                if off is None:
                    if r is head and r.bypassed >= self.max_bypass:
                        break  # hold the line: nothing starts before the oldest request
                    continue
            queue.remove(r)
            if head is not None and r is not head and head in queue:
                head.bypassed += 1
                self.counters["bypassed"] += 1
            slot = free.pop(0)
            r.t_start = self.clock() - t0
            r.status = "running"
            self.executor.start(slot, r, None if off is None else self.pool.view(off, need), ent)
            slots[slot] = (r, off, need)
            self.counters["admitted"] += 1

    def run(self, requests, realtime=True, progress=None, every=5.0, on_tick=None):
        """Serve ``requests``; returns a stats dict. ``progress(done, total)`` is called at most
        every ``every`` seconds; ``on_tick(server, now)`` once per scheduling pass."""
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        queue = []
        slots = [None] * self.executor.n
        t0 = self.clock()
        done = []
        last = t0
        total = len(pending)
        self.counters = collections.Counter()
        while pending or queue or any(s is not None for s in slots):
            if progress is not None and self.clock() - last > every:
                last = self.clock()
                progress(len(done), total)
            now = self.clock() - t0
            if on_tick is not None:
                on_tick(self, now)
            while pending and (pending[0].arrival <= now or not realtime):
                r = pending.popleft()
                r.status = "queued"
                queue.append(r)
            self.counters["queued_peak"] = max(self.counters["queued_peak"], len(queue))
            progressed = False
            for i, s in enumerate(slots):
                if s is not None and self.executor.done(i):
                    r, off, need = s
                    r.t_done = self.clock() - t0
                    r.status = "cancelled" if r.cancelled else "done"
                    if off is not None:
                        self.pool.release(off, need)
                    done.append(r)
                    slots[i] = None
                    progressed = True
            n_before = sum(s is not None for s in slots)
            self._dispatch(queue, slots, t0)
            progressed = progressed or sum(s is not None for s in slots) != n_before
            if not progressed:
                time.sleep(2e-5)
        self.executor.wait_all()
        stats = summarize([r for r in done if r.status == "done"])
        stats.update({k: int(v) for k, v in self.counters.items()})
        stats["pool_peak_bytes"] = int(self.pool.peak)
        stats["graphed"] = sum(r.graphed for r in done)
        return stats


def summarize(done):
    if not done:
        return {"requests": 0}
    lat = np.array([r.latency for r in done]) * 1e3
    upd = sum(r.job.updates(r.K, r.grad) for r in done)
    span = max(r.t_done for r in done) - min(r.arrival for r in done)
    return {
        "requests": len(done),
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "mean_ms": float(lat.mean()),
        "updates": int(upd),
        "span_s": float(span),
        "updates_per_s": float(upd / span) if span > 0 else float("nan"),
    }


def run_sequential(requests, device=0, ws_bytes=None, time_limit=None, progress=None,
                   every=5.0, realtime=False):
    """Plain per-call evaluation: one request at a time, synchronised, first come first served.

    ``realtime=False``: back to back from t = 0, each request's latency its own service time.
    ``realtime=True``: the requests' own arrival times are replayed, a request starts when it
    has arrived and the previous one is done, and its latency runs from its arrival, so the
    queueing is counted exactly as the server's is. With ``time_limit`` (seconds) the run stops
    after the request that crosses it; the stats then cover the completed requests only
    (``partial``).
    """
    dev = torch.device("cuda", device)
    need = max(r.workspace_bytes() for r in requests)
    ws = torch.empty(int(need if ws_bytes is None else ws_bytes), dtype=torch.uint8, device=dev)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    last = t0
    done = []
    for r in sorted(requests, key=lambda q: q.arrival) if realtime else requests:
        if time_limit is not None and time.perf_counter() - t0 > time_limit:
            break
        if progress is not None and time.perf_counter() - last > every:
            last = time.perf_counter()
            progress(len(done), len(requests))
        if realtime:
            wait = r.arrival - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
        done.append(r)
        a = time.perf_counter()
        r.run_on(r.bls.to(dev), ws, torch.cuda.current_stream(dev))
        torch.cuda.synchronize(dev)
        r.t_start = a - t0
        if not realtime:
            r.arrival = a - t0
        r.t_done = time.perf_counter() - t0
    out = summarize(done)
    out["partial"] = len(done) < len(requests)
    return out
