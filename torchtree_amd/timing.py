"""Per-phase device timing with HIP events on the launching stream.

``with timing.record() as rec:`` makes the engine record a start/stop event
pair around each library call (P matrices, post-order, pre-order, P
gradient) on the stream that call is launched on (torch's current stream).
After synchronising, ``rec.summary()`` gives milliseconds per phase.
"""
from __future__ import annotations

import collections
import contextlib

import torch

_ACTIVE = None


class Recorder:
    def __init__(self):
        self.events = []  # (name, start, stop, launches)

    def summary(self):
        torch.cuda.synchronize()
        tot = collections.OrderedDict()
        cnt = collections.Counter()
        launches = collections.Counter()
        for name, a, b, nl in self.events:
            tot[name] = tot.get(name, 0.0) + a.elapsed_time(b)
            cnt[name] += 1
            launches[name] += nl
        return {k: {"ms": v, "calls": cnt[k], "launches": launches[k]} for k, v in tot.items()}


@contextlib.contextmanager
def record():
    global _ACTIVE
    prev = _ACTIVE
    rec = Recorder()
    _ACTIVE = rec
    try:
        yield rec
    finally:
        _ACTIVE = prev


@contextlib.contextmanager
def phase(name: str, launches: int = 1):
    rec = _ACTIVE
    if rec is None:
        yield
        return
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    try:
        yield
    finally:
        b.record()
        rec.events.append((name, a, b, launches))
