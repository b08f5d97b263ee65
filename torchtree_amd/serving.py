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
                 dtype=torch.float32, slice_patterns=0, device=None, kind="dna"):
        from .likelihood import get_plan

        lib = _lib.load()
        if device is None:
            device = torch.cuda.current_device()
        self.name, self.kind, self.dtype, self.device = name, kind, dtype, device
        codes = np.ascontiguousarray(np.asarray(codes, dtype=np.uint8))
        self.T, self.N = codes.shape
        self.B = 2 * self.T - 2
        table = np.ascontiguousarray(np.asarray(table, dtype=np.float64))
        self.S = table.shape[1]
        rates = np.ascontiguousarray(np.asarray(rates, dtype=np.float64).reshape(-1))
        props = np.ascontiguousarray(np.asarray(props, dtype=np.float64).reshape(-1))
        self.C = rates.shape[0]
        weights = np.ascontiguousarray(np.asarray(weights, dtype=np.float64))
        lam, left, right = _host_eigen(torch.as_tensor(q_norm), torch.as_tensor(freqs))
        fr = np.ascontiguousarray(np.asarray(freqs, dtype=np.float64))
        self.plan = get_plan(self.T, post_indexing, device)
        h = ctypes.c_void_p()
        code = HX_F64 if dtype == F64 else HX_F32
        prev = torch.cuda.current_device()
        torch.cuda.set_device(device)
        try:
            rc = lib.hx_job_create(self.plan.handle, self.S, self.N, self.C, code, int(slice_patterns),
                                   codes.ctypes.data, table.ctypes.data, table.shape[0],
                                   weights.ctypes.data, lam.ctypes.data, left.ctypes.data,
                                   right.ctypes.data, rates.ctypes.data, props.ctypes.data,
                                   fr.ctypes.data, device, ctypes.byref(h))
        finally:
            torch.cuda.set_device(prev)
        _lib.check(rc, "hx_job_create")
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

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and _lib._lib is not None:
            _lib._lib.hx_job_destroy(h)
            self.handle = None


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

    @property
    def K(self):
        return self.bls.shape[0]

    @property
    def latency(self):
        return self.t_done - self.arrival


class GpuExecutor:
    """Runs requests on HIP streams of one device (default stream priority)."""

    def __init__(self, device, n_streams, ws_bytes_per_stream):
        self.device = torch.device("cuda", device)
        self.n = n_streams
        self.streams = [torch.cuda.Stream(self.device) for _ in range(n_streams)]
        self.ws = [torch.empty(int(ws_bytes_per_stream), dtype=torch.uint8, device=self.device)
                   for _ in range(n_streams)]
        self.events = [None] * n_streams

    def start(self, slot, req: Request):
        st = self.streams[slot]
        job = req.job
        need = job.workspace_bytes(req.K, req.grad)
        if need > self.ws[slot].numel():
            raise BudgetError(f"request needs {need} B of workspace, stream has {self.ws[slot].numel()}")
        with torch.cuda.stream(st):
            bls = req.bls.to(self.device, non_blocking=True)
            req.loglik = torch.empty(req.K, dtype=F64, device=self.device)
            req.dbls = torch.empty(req.K, job.B, dtype=F64, device=self.device) if req.grad else None
            job.eval(bls, req.loglik, req.dbls, self.ws[slot], st.cuda_stream)
            ev = torch.cuda.Event()
            ev.record(st)
        self.events[slot] = (ev, bls)

    def done(self, slot):
        ev = self.events[slot]
        if ev is None:
            return True
        if ev[0].query():
            self.events[slot] = None
            return True
        return False

    def wait_all(self):
        torch.cuda.synchronize(self.device)


class FakeExecutor:
    """CPU stand-in: a request finishes ``cost(req)`` seconds after it starts."""

    def __init__(self, n_streams, ws_bytes_per_stream, cost, clock):
        self.n = n_streams
        self.ws_bytes = ws_bytes_per_stream
        self.cost = cost
        self.clock = clock
        self.busy_until = [None] * n_streams
        self.log = []

    def start(self, slot, req):
        need = req.job.workspace_bytes(req.K, req.grad)
        if need > self.ws_bytes:
            raise BudgetError(f"request needs {need} B of workspace")
        self.busy_until[slot] = self.clock() + self.cost(req)
        self.log.append(("start", slot, req.rid))

    def done(self, slot):
        t = self.busy_until[slot]
        if t is None:
            return True
        if self.clock() >= t:
            self.busy_until[slot] = None
            self.log.append(("done", slot))
            return True
        return False

    def wait_all(self):
        pass


class Server:
    """Admission under an HBM budget plus FCFS continuous batching over streams.

    ``hbm_budget``: bytes this server may hold on the device: resident job
    data plus ``n_streams`` workspaces of ``ws_bytes_per_stream``.
    """

    def __init__(self, hbm_budget, n_streams=4, ws_bytes_per_stream=None, executor=None,
                 device=0, clock=time.perf_counter):
        self.budget = int(hbm_budget)
        self.n_streams = n_streams
        if ws_bytes_per_stream is None:
            ws_bytes_per_stream = int(0.8 * self.budget / n_streams)
        self.ws_bytes = int(ws_bytes_per_stream)
        self.resident_cap = self.budget - n_streams * self.ws_bytes
        if self.resident_cap <= 0:
            raise BudgetError("HBM budget leaves no room for resident jobs")
        self.resident = 0
        self.jobs = {}
        self.clock = clock
        self.executor = executor if executor is not None else GpuExecutor(device, n_streams,
                                                                         self.ws_bytes)
        self.rejected = 0

    def slice_for(self, T, N, C, S, dtype, max_k=32):
        """Largest pattern slice whose max-K gradient workspace fits one stream.

        Sizes come from the library (hx_job_workspace_bytes_for), so they match what a
        registered job will ask for. The size is not monotone in the slice everywhere (the
        four-state gradient switches kernels with the workgroup count), so this bisects
        keeping a slice that fits."""
        lib = _lib.load()
        code = _lib.HX_F64 if dtype == F64 else _lib.HX_F32

        def need(n):
            return int(lib.hx_job_workspace_bytes_for(T, S, n, C, max_k, code, 1))

        if need(1) > self.ws_bytes:
            raise BudgetError("one pattern of this job does not fit a stream workspace")
        if need(N) <= self.ws_bytes:
            return N
        lo, hi = 1, N
        while hi - lo > 1:
            mid = (lo + hi) // 2
            if need(mid) <= self.ws_bytes:
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
        job = self.jobs.pop(name)
        self.resident -= job.resident_bytes

    def run(self, requests, realtime=True, progress=None, every=5.0):
        """Serve ``requests`` (sorted by arrival). Returns a stats dict.

        ``progress(done, total)`` is called at most every ``every`` seconds.
        """
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        queue = collections.deque()
        slots = [None] * self.executor.n
        t0 = self.clock()
        done = []
        last = t0
        total = len(pending)
        while pending or queue or any(s is not None for s in slots):
            progressed = False
            if progress is not None and self.clock() - last > every:
                last = self.clock()
                progress(len(done), total)
            now = self.clock() - t0
            while pending and (pending[0].arrival <= now or not realtime):
                queue.append(pending.popleft())
            for i, r in enumerate(slots):
                if r is not None and self.executor.done(i):
                    r.t_done = self.clock() - t0
                    done.append(r)
                    slots[i] = None
                    progressed = True
            for i in range(len(slots)):
                if slots[i] is None and queue:
                    r = queue.popleft()
                    r.t_start = self.clock() - t0
                    self.executor.start(i, r)
                    slots[i] = r
                    progressed = True
            if not progressed:
                time.sleep(2e-5)
        self.executor.wait_all()
        return summarize(done)


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
                   every=5.0):
    """Plain per-call evaluation: one request at a time, synchronised, back to back.

    With ``time_limit`` (seconds) the run stops after the request that crosses
    it; the stats then cover the completed requests only (``partial``).
    """
    dev = torch.device("cuda", device)
    need = max(r.job.workspace_bytes(r.K, r.grad) for r in requests)
    ws = torch.empty(int(need if ws_bytes is None else ws_bytes), dtype=torch.uint8, device=dev)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    last = t0
    done = []
    for r in requests:
        if time_limit is not None and time.perf_counter() - t0 > time_limit:
            break
        if progress is not None and time.perf_counter() - last > every:
            last = time.perf_counter()
            progress(len(done), len(requests))
        done.append(r)
        a = time.perf_counter()
        bls = r.bls.to(dev)
        r.loglik = torch.empty(r.K, dtype=F64, device=dev)
        r.dbls = torch.empty(r.K, r.job.B, dtype=F64, device=dev) if r.grad else None
        r.job.eval(bls, r.loglik, r.dbls, ws, torch.cuda.current_stream(dev).cuda_stream)
        torch.cuda.synchronize(dev)
        r.t_start = a - t0
        r.arrival = a - t0
        r.t_done = time.perf_counter() - t0
    out = summarize(done)
    out["partial"] = len(done) < len(requests)
    return out
