"""Serving scheduler on the CPU with a fake executor (no GPU): HBM-budget
admission, first-come first-served dispatch, at most n_streams in flight,
every request completed, latency accounting."""
import time

import pytest
import torch

from torchtree_amd import serving


class FakeJob:
    def __init__(self, resident, ws, T=10, N=100, C=4):
        self.resident_bytes = resident
        self._ws = ws
        self.T, self.N, self.C, self.B = T, N, C, 2 * T - 2

    def workspace_bytes(self, K, grad=True):
        return self._ws * K

    def updates(self, K, grad=True):
        return (3 * self.T - 3) * self.N * self.C * K


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 1e-4
        return self.t


def make_server(n_streams=3, budget=10_000, ws=2_000):
    clock = Clock()
    ex = serving.FakeExecutor(n_streams, ws, cost=lambda r: 0.01 * r.K, clock=clock)
    return serving.Server(budget, n_streams=n_streams, ws_bytes_per_stream=ws, executor=ex,
                          clock=clock), ex


def test_budget_admission():
    srv, _ = make_server()
    srv.register("a", FakeJob(2_000, 10))
    srv.register("b", FakeJob(1_500, 10))
    with pytest.raises(serving.BudgetError):
        srv.register("c", FakeJob(1_000, 10))  # resident cap = 10_000 - 3 * 2_000 = 4_000
    assert srv.rejected == 1
    srv.unregister("a")
    srv.register("c", FakeJob(1_000, 10))
    with pytest.raises(serving.BudgetError):
        serving.Server(1000, n_streams=2, ws_bytes_per_stream=600, executor=object())


def test_fcfs_concurrency_and_completion():
    srv, ex = make_server()
    job = FakeJob(100, 10)
    reqs = [serving.Request(i, job, torch.zeros((1 + i % 3) * 8, 4), arrival=0.001 * i)
            for i in range(20)]
    stats = srv.run(reqs)
    assert stats["requests"] == 20
    starts = [rid for kind, *rest in ex.log if kind == "start" for rid in [rest[1]]]
    assert starts == sorted(starts)  # first come, first served
    inflight = peak = 0
    for kind, *_ in ex.log:
        inflight += 1 if kind == "start" else -1
        peak = max(peak, inflight)
    assert peak <= 3
    for r in reqs:
        assert r.t_start >= r.arrival - 1e-12 and r.t_done > r.t_start
    assert stats["p99_ms"] >= stats["p50_ms"] > 0


def test_oversized_request_refused():
    srv, _ = make_server()
    job = FakeJob(100, 1_000)
    with pytest.raises(serving.BudgetError):
        srv.run([serving.Request(0, job, torch.zeros(8, 4))])


def test_late_arrivals_do_not_block_queue():
    srv, ex = make_server(n_streams=1)
    job = FakeJob(100, 10)
    reqs = [serving.Request(0, job, torch.zeros(1, 4), arrival=0.0),
            serving.Request(1, job, torch.zeros(1, 4), arrival=0.5)]
    srv.run(reqs)
    assert reqs[0].t_done < reqs[1].t_start
