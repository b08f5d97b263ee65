"""Serving scheduler on the CPU with a fake executor (no GPU).

Covers: the resident-data budget, admission by workspace memory (a request starts only when
its workspace fits the pool), first-come first-served dispatch with backfill and its bypass
bound, rejection of oversized requests, cancellation of waiting requests, randomised
completion order, at most n_streams in flight, and the workspace pool allocator itself.
"""
import random

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


def make_server(n_streams=3, budget=None, pool=6_000, order=None, **kw):
    clock = Clock()
    budget = pool + 4_000 if budget is None else budget
    ex = serving.FakeExecutor(n_streams, cost=lambda r: 0.01 * r.K, clock=clock, order=order)
    return serving.Server(budget, n_streams=n_streams, pool_bytes=pool, executor=ex,
                          clock=clock, **kw), ex


def _inflight(log):
    live, peak = set(), 0
    slot_req = {}
    for kind, *rest in log:
        if kind == "start":
            slot_req[rest[0]] = rest[1]
            live.add(rest[1])
        else:
            live.discard(slot_req.pop(rest[0]))
        peak = max(peak, len(live))
    return peak


def test_budget_admission():
    srv, _ = make_server()
    srv.register("a", FakeJob(2_000, 10))
    srv.register("b", FakeJob(1_500, 10))
    with pytest.raises(serving.BudgetError):
        srv.register("c", FakeJob(1_000, 10))  # resident cap = 10_000 - 6_000 = 4_000
    assert srv.rejected == 1
    srv.unregister("a")
    srv.register("c", FakeJob(1_000, 10))
    with pytest.raises(serving.BudgetError):
        serving.Server(1000, n_streams=2, pool_bytes=1000, executor=serving.FakeExecutor(
            2, cost=lambda r: 0.0, clock=Clock()))


def test_fcfs_concurrency_and_completion():
    srv, ex = make_server()
    job = FakeJob(100, 10)
    reqs = [serving.Request(i, job, torch.zeros((1 + i % 3) * 8, 4), arrival=0.001 * i)
            for i in range(20)]
    stats = srv.run(reqs)
    assert stats["requests"] == 20 and stats["admitted"] == 20
    starts = [rest[1] for kind, *rest in ex.log if kind == "start"]
    assert starts == sorted(starts)  # every request fits: first come, first served
    assert _inflight(ex.log) <= 3
    for r in reqs:
        assert r.status == "done"
        assert r.t_start >= r.arrival - 1e-12 and r.t_done > r.t_start
    assert stats["p99_ms"] >= stats["p50_ms"] > 0
    assert srv.pool.in_use == 0


def test_memory_limits_concurrency():
    """Three streams, but each request takes 40 % of the pool: at most two run at once."""
    srv, ex = make_server(n_streams=3, pool=10_000, max_request_bytes=10_000)
    job = FakeJob(0, 4_000)
    reqs = [serving.Request(i, job, torch.zeros(1, 4), arrival=0.0) for i in range(7)]
    stats = srv.run(reqs)
    assert stats["requests"] == 7
    assert _inflight(ex.log) == 2
    assert srv.pool.peak <= srv.pool.size and srv.pool.in_use == 0


def test_oversized_request_rejected_not_raised():
    srv, _ = make_server()
    big, small = FakeJob(100, 1_000), FakeJob(100, 10)
    reqs = [serving.Request(0, big, torch.zeros(8, 4)), serving.Request(1, small, torch.zeros(1, 4))]
    stats = srv.run(reqs)
    assert reqs[0].status == "rejected" and reqs[1].status == "done"
    assert stats["rejected"] == 1 and stats["requests"] == 1


def test_backfill_and_bypass_bound():
    """A request that does not fit the free memory is passed by smaller ones behind it, but by
    at most max_bypass of them; then it holds the line until it fits."""
    srv, ex = make_server(n_streams=4, pool=10_000, max_request_bytes=10_000, max_bypass=3)
    mid, big, small = FakeJob(0, 1_500), FakeJob(0, 8_000), FakeJob(0, 500)
    reqs = [serving.Request(0, mid, torch.zeros(4, 4), arrival=0.0),   # K=4: long, holds 6,000
            serving.Request(1, big, torch.zeros(1, 4), arrival=0.0001)]
    reqs += [serving.Request(2 + i, small, torch.zeros(1, 4), arrival=0.0002 + 1e-5 * i)
             for i in range(8)]
    srv.run(reqs)
    starts = [rest[1] for kind, *rest in ex.log if kind == "start"]
    assert starts[0] == 0
    i_big = starts.index(1)
    passed = [rid for rid in starts[1:i_big]]
    assert len(passed) == 3 and reqs[1].bypassed == 3  # three small ones went first, no more
    assert all(r.status == "done" for r in reqs)


def test_cancel_waiting_request():
    srv, ex = make_server(n_streams=1)
    job = FakeJob(0, 10)
    reqs = [serving.Request(i, job, torch.zeros(2, 4), arrival=0.0) for i in range(4)]

    def tick(server, now):
        if now > 0.005 and reqs[2].status == "queued":
            server.cancel(reqs[2])

    stats = srv.run(reqs, on_tick=tick)
    started = {rest[1] for kind, *rest in ex.log if kind == "start"}
    assert reqs[2].status == "cancelled" and 2 not in started
    assert stats["cancelled"] == 1 and stats["requests"] == 3


def test_random_completion_order():
    """Completions reported in random order among the due requests: every request finishes,
    the pool drains, and no more than n_streams run at once."""
    rng = random.Random(5)
    srv, ex = make_server(n_streams=4, pool=20_000, max_request_bytes=10_000,
                          order=lambda due: rng.choice(due))
    jobs = [FakeJob(0, w) for w in (300, 1_200, 2_500, 4_000)]
    reqs = [serving.Request(i, jobs[i % 4], torch.zeros(1 + i % 2, 4), arrival=0.0005 * i)
            for i in range(60)]
    stats = srv.run(reqs)
    assert stats["requests"] == 60 and all(r.status == "done" for r in reqs)
    assert _inflight(ex.log) <= 4
    assert srv.pool.in_use == 0 and srv.pool.free == [(0, srv.pool.size)]


def test_late_arrivals_do_not_block_queue():
    srv, ex = make_server(n_streams=1)
    job = FakeJob(100, 10)
    reqs = [serving.Request(0, job, torch.zeros(1, 4), arrival=0.0),
            serving.Request(1, job, torch.zeros(1, 4), arrival=0.5)]
    srv.run(reqs)
    assert reqs[0].t_done < reqs[1].t_start


def test_workspace_pool_against_a_byte_map():
    """Random allocations and releases: no two live ranges overlap, every range is 256-byte
    aligned and inside the pool, freed space coalesces back to one range."""
    rng = random.Random(9)
    pool = serving.WorkspacePool(1 << 20)
    live = {}
    for step in range(3000):
        if live and (rng.random() < 0.45 or len(live) > 40):
            off = rng.choice(list(live))
            pool.release(off, live.pop(off))
        else:
            n = rng.choice((1, 255, 256, 1000, 4096, 70_000))
            off = pool.alloc(n)
            if off is None:
                assert sum(sz for _, sz in pool.free) < 70_000 + 4096 or max(
                    sz for _, sz in pool.free) < n
                continue
            assert off % 256 == 0 and off + n <= pool.size
            for o2, n2 in live.items():
                assert off + n <= o2 or o2 + n2 <= off
            live[off] = n
    for off in list(live):
        pool.release(off, live.pop(off))
    assert pool.free == [(0, pool.size)] and pool.in_use == 0
