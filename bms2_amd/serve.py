"""Serving loop over a resident data batch.

A resident set of data graphs stays on the device behind one pipeline (its packs, adjacency rows
and staged objects for large graphs are built once).  Requests -- batches of 1-256 query graphs
-- arrive as a Poisson stream; the server admits the waiting requests into one query batch while
the batch's dense outputs (one u64 count per (data graph, query) pair plus the GMCR bit rows) fit
an HBM budget, binds the batch to the pipeline (bms2_pipeline_set_queries), optionally plans the
learned matching order, runs the pipeline, and answers every request with its queries' matched
data graphs and embedding counts (bms2_pipeline_query_totals).  Latency of a request = time its
answer is ready - its arrival time, on the wall clock of the serving process.

The reference has no serving mode: its batch driver streams data graphs past a fixed query set
(BMS2_batch.cpp:104-283); here the data side is resident and the queries stream.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Request:
    rid: int
    arrival: float              # seconds from the start of the stream
    queries: np.ndarray         # indices into the query pool
    done: float = -1.0
    matched: np.ndarray | None = None
    embeddings: np.ndarray | None = None


def make_requests(n: int, rate: float, pool_size: int, seed: int, min_queries: int = 1,
                  max_queries: int = 256) -> list[Request]:
    """n requests with exponential inter-arrival times (rate per second) and uniform batch
    sizes in [min_queries, max_queries] drawn from a pool of pool_size query graphs."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += float(rng.exponential(1.0 / rate))
        k = int(rng.integers(min_queries, max_queries + 1))
        out.append(Request(i, t, rng.integers(0, pool_size, size=k)))
    return out


def batch_bytes(num_data_graphs: int, num_queries: int) -> int:
    """Dense outputs of one batch: u64 counts plus u64 GMCR bit rows."""
    return 8 * num_data_graphs * num_queries + 8 * num_data_graphs * ((num_queries + 63) // 64)


def admit(pending: deque, num_data_graphs: int, budget_bytes: int) -> list[Request]:
    """Takes waiting requests in arrival order while the batch's outputs fit the budget (the
    first request is always taken, so an oversized request is served alone)."""
    taken, nq = [], 0
    while pending:
        k = len(pending[0].queries)
        if taken and batch_bytes(num_data_graphs, nq + k) > budget_bytes:
            break
        taken.append(pending.popleft())
        nq += k
    return taken


@dataclass
class ServeReport:
    requests: int
    rate: float
    order: str
    p50_ms: float
    p99_ms: float
    mean_ms: float
    wall_s: float
    batches: int
    mean_batch_queries: float
    queries_per_s: float
    pairs_per_s: float
    busy_s: float
    extra: dict = field(default_factory=dict)


def serve(requests: list[Request], execute, num_data_graphs: int, budget_bytes: int,
          clock=time.perf_counter, sleep=time.sleep) -> tuple[list[Request], dict]:
    """Runs the admission loop on the wall clock.  execute(list[Request]) answers a batch (fills
    matched / embeddings); returns the requests and loop statistics."""
    pending: deque = deque()
    i, n = 0, len(requests)
    t0 = clock()
    batches, busy, served_queries = [], 0.0, 0
    while i < n or pending:
        now = clock() - t0
        while i < n and requests[i].arrival <= now:
            pending.append(requests[i])
            i += 1
        if not pending:
            sleep(max(0.0, requests[i].arrival - now))
            continue
        batch = admit(pending, num_data_graphs, budget_bytes)
        b0 = clock()
        execute(batch)
        b1 = clock()
        busy += b1 - b0
        for r in batch:
            r.done = b1 - t0
        nq = sum(len(r.queries) for r in batch)
        served_queries += nq
        batches.append(nq)
    wall = clock() - t0
    return requests, dict(wall_s=wall, busy_s=busy, batches=batches, queries=served_queries)


def report(requests: list[Request], loop: dict, rate: float, order: str,
           num_data_graphs: int) -> ServeReport:
    lat = np.array([1000.0 * (r.done - r.arrival) for r in requests])
    wall = loop["wall_s"]
    return ServeReport(
        requests=len(requests), rate=rate, order=order,
        p50_ms=round(float(np.percentile(lat, 50)), 2),
        p99_ms=round(float(np.percentile(lat, 99)), 2),
        mean_ms=round(float(lat.mean()), 2), wall_s=round(wall, 3),
        batches=len(loop["batches"]),
        mean_batch_queries=round(float(np.mean(loop["batches"])), 1),
        queries_per_s=round(loop["queries"] / wall, 1),
        pairs_per_s=round(loop["queries"] * num_data_graphs / wall, 1),
        busy_s=round(loop["busy_s"], 3))


class PipelineExecutor:
    """Answers batches on the GPU: binds the batch's queries to the resident pipeline, runs it,
    and reads per-query totals.  With a network, a batch holding queries never planned before is
    planned (trace of plan_graphs resident graphs, bf16 inference) and the orders are cached per
    pool query: a query's plan depends on the query and the resident data only, so later batches
    reuse it without tracing."""

    def __init__(self, q, pipe, pool, num_data_graphs: int, max_queries: int, net=None,
                 plan_graphs: int = 4096):
        from .api import DeviceBuffer
        self.q, self.pipe, self.pool, self.net = q, pipe, pool, net
        self.D = num_data_graphs
        self.plan_graphs = plan_graphs
        self.counts = DeviceBuffer(q, 8 * self.D * max_queries, zero=False)
        self.valid = DeviceBuffer(q, 8 * self.D * ((max_queries + 63) // 64), zero=False)
        self.stats = DeviceBuffer(q, 64, zero=True)
        self.static = None  # structure-only feature rows of the pool's query nodes
        self.orders: dict[int, np.ndarray] = {}  # pool index -> planned order row
        self.planned_batches = 0
        self.pool_go = np.asarray(pool.graph_offsets, np.int64)
        if net is not None:
            from . import learned
            self.static = learned.static_features(pool)
        self.keep = []  # device query batches stay alive while bound

    def __call__(self, batch: list[Request]) -> None:
        from . import learned
        from .api import DeviceBatchedCSRGraph
        from .graphs import GraphBatch
        idx = np.concatenate([r.queries for r in batch])
        qb = GraphBatch.from_graphs([self.pool.graph(int(i)) for i in idx])
        dq = DeviceBatchedCSRGraph(self.q, qb)
        self.pipe.set_queries(dq)
        self.keep = [dq]
        if self.net is not None:
            if all(int(i) in self.orders for i in idx):
                self.pipe.set_order(np.stack([self.orders[int(i)] for i in idx]))
            else:
                rows = np.concatenate([np.arange(self.pool_go[i], self.pool_go[i + 1])
                                       for i in idx])
                pl = learned.plan(self.pipe, self.net, qb, sample_graphs=self.plan_graphs,
                                  static=self.static[rows])
                self.planned_batches += 1
                for k, i in enumerate(idx):
                    self.orders.setdefault(int(i), pl.order[k].copy())
        self.pipe.run(self.counts, self.valid, self.stats)
        matched, emb = self.pipe.query_totals(self.counts)
        o = 0
        for r in batch:
            k = len(r.queries)
            r.matched, r.embeddings = matched[o:o + k], emb[o:o + k]
            o += k


def _main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m bms2_amd.serve")
    ap.add_argument("--graphs", type=int, default=2_000_000, help="resident data graphs")
    ap.add_argument("--pool", type=int, default=4096, help="query pool size")
    ap.add_argument("--requests", type=int, default=60)
    ap.add_argument("--rates", default="50,2000", help="request rates per second")
    ap.add_argument("--budget-gb", type=float, default=8.0, help="HBM budget of a batch's outputs")
    ap.add_argument("--refinement-steps", type=int, default=6)
    ap.add_argument("--orders", default="identity,learned")
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args(argv)
    from . import api, graphs, learned
    log = lambda s: print(s, file=sys.stderr, flush=True)  # noqa: E731
    t = time.perf_counter()
    data, pool = graphs.generate("molecules-2m", seed=args.seed, num_data_graphs=args.graphs,
                                 num_query_graphs=args.pool)
    log(f"[serve] {data.num_graphs} resident graphs, pool of {pool.num_graphs} queries "
        f"({time.perf_counter() - t:.1f}s)")
    q = api.Queue(0)
    dd = api.DeviceBatchedCSRGraph(q, data)
    budget = int(args.budget_gb * (1 << 30))
    max_q = 1
    while batch_bytes(data.num_graphs, max_q + 1) <= budget:
        max_q += 1
    max_q = max(max_q, 256)
    # the pipeline starts on the first 256 pool queries; batches rebind it
    first = graphs.GraphBatch.from_graphs([pool.graph(i) for i in range(min(256, pool.num_graphs))])
    dq0 = api.DeviceBatchedCSRGraph(q, first)
    pipe = api.Pipeline(q, dq0, dd, refinement_steps=args.refinement_steps)
    out = []
    for order in args.orders.split(","):
        net = learned.OrderNet.load(q) if order == "learned" else None
        ex = PipelineExecutor(q, pipe, pool, data.num_graphs, max_q, net=net)
        for rate in (float(x) for x in args.rates.split(",")):
            reqs = make_requests(args.requests, rate, pool.num_graphs, seed=args.seed + 1)
            ex(reqs[:1])  # warm the bound pipeline (graph capture, caches)
            for r in reqs:
                r.done = -1.0
            reqs, loop = serve(reqs, ex, data.num_graphs, budget)
            rep = report(reqs, loop, rate, order, data.num_graphs)
            rep.extra = dict(planned_batches=ex.planned_batches, cached_orders=len(ex.orders))
            ex.planned_batches = 0
            log(f"[serve] {order} @ {rate:g} req/s: p50 {rep.p50_ms} ms, p99 {rep.p99_ms} ms, "
                f"{rep.queries_per_s} queries/s, {rep.batches} batches")
            out.append(rep.__dict__)
    print(json.dumps(dict(graphs=data.num_graphs, pool=pool.num_graphs, budget_gb=args.budget_gb,
                          max_batch_queries=max_q, runs=out)))
    return 0


if __name__ == "__main__":
    sys.exit(_main())
