"""Serving loop logic on the CPU: request stream, HBM-budget admission, latency accounting."""
from collections import deque

import numpy as np

from bms2_amd import serve


def test_request_stream_is_poisson_and_seeded():
    a = serve.make_requests(2000, 400.0, 100, seed=3)
    b = serve.make_requests(2000, 400.0, 100, seed=3)
    assert [r.arrival for r in a] == [r.arrival for r in b]
    gaps = np.diff([0.0] + [r.arrival for r in a])
    assert abs(gaps.mean() - 1 / 400) < 0.1 / 400          # mean inter-arrival 1 / rate
    assert abs(gaps.std() / gaps.mean() - 1) < 0.1         # exponential: cv = 1
    sizes = [len(r.queries) for r in a]
    assert min(sizes) >= 1 and max(sizes) <= 256 and max(sizes) > 200
    assert all(r.queries.max() < 100 for r in a)


def test_admission_respects_budget():
    D = 1000
    reqs = [serve.Request(i, 0.0, np.zeros(k, np.int64)) for i, k in enumerate([100, 100, 100, 50])]
    budget = serve.batch_bytes(D, 250)
    pending = deque(reqs)
    first = serve.admit(pending, D, budget)
    assert [r.rid for r in first] == [0, 1]                  # 300 queries would not fit
    second = serve.admit(pending, D, budget)
    assert [r.rid for r in second] == [2, 3]
    # an oversized request is served alone
    pending = deque([serve.Request(9, 0.0, np.zeros(400, np.int64))])
    assert len(serve.admit(pending, D, budget)) == 1


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += dt


def test_serve_loop_latency_accounting():
    clock = FakeClock()
    reqs = [serve.Request(i, 0.01 * i, np.zeros(10, np.int64)) for i in range(5)]

    def execute(batch):  # 5 ms per query batch plus 1 ms per query
        clock.t += 0.005 + 0.001 * sum(len(r.queries) for r in batch)
        for r in batch:
            r.matched = np.zeros(len(r.queries), np.uint64)

    out, loop = serve.serve(reqs, execute, 100, serve.batch_bytes(100, 1000), clock=clock,
                            sleep=clock.sleep)
    # each request arrives after the previous one is served (15 ms service < 10 ms gap + wait)
    assert all(r.done > r.arrival for r in out)
    assert abs(out[0].done - out[0].arrival - 0.015) < 1e-9
    assert loop["queries"] == 50 and sum(loop["batches"]) == 50
    rep = serve.report(out, loop, 100.0, "identity", 100)
    assert rep.requests == 5 and rep.p50_ms >= 15.0 and rep.batches == len(loop["batches"])


def test_overload_batches_waiting_requests():
    clock = FakeClock()
    reqs = [serve.Request(i, 0.0, np.zeros(4, np.int64)) for i in range(8)]  # all at t = 0

    def execute(batch):
        clock.t += 0.1

    out, loop = serve.serve(reqs, execute, 10, serve.batch_bytes(10, 16), clock=clock,
                            sleep=clock.sleep)
    assert loop["batches"] == [16, 16]                        # four requests per batch
    assert sorted({round(r.done, 6) for r in out}) == [0.1, 0.2]
