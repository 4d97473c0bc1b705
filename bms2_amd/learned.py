"""Learned matching order for the fused pipeline's join.

The reference's join matches query nodes in identity order (isomorphism.hpp:371-447; its
alternative defineMatchingOrder, isomorphism.hpp:476-502, is unused).  With standard
backtracking the per-pair counts do not depend on the order, only the number of DFS nodes does,
so any permutation keeps results exact.  This module:

* logs join traces from plain runs (bms2_pipeline_trace): per query node the summed domain size
  over its GMCR pairs, per query the GMCR pairs and DFS nodes, per pair the DFS nodes;
* labels every query with the cheapest of a set of candidate orders (identity, the reference's
  max-domain root, and a connected greedy order from every root) measured by realised DFS nodes,
  and with its useful refinement depth (the smallest depth whose pruning is within 1% of depth 8);
* trains the matching-order network (bms2_mlp_*: hand-written HIP forward, backward and Adam
  kernels, f32) on per-query-node features;
* plans a query batch (bf16 inference): per query a connected order by descending predicted
  score; identity order takes over when the predicted cost ratio to identity is not below the
  budget.

The network's refinement-depth head is trained and reported but not applied: the pipeline's shell
histograms at distance k >= 3 can remove true embeddings, so a per-query depth other than the
run's K would change counts, which must stay equal to the plain pipeline's.  Deferring levels
d+1..K to the join keeps counts exact but saves nothing here: the fused screen tests every level
inside one node test, GMCR validity needs all K levels, and a shallower screen is slower (each
level ends graphs before their later node tests: 70.5 ms at K = 2 against 63.3 ms at K = 6 on
1M molecules x 1,024 queries).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
import time
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import _lib
from .api import DeviceBuffer, Pipeline, Queue, _Handle, _L, check

NUM_FEATURES = 32
NUM_OUTPUTS = 10
DEFAULT_WEIGHTS = Path(__file__).resolve().parent / "order_net.npz"
# element codes of the project generator (workload.cpp): N, Cl, Br, O, C, S, F
_ELEMENTS = (0, 1, 3, 7, 8, 9, 10)
_WILDCARD = 255


class OrderNet(_Handle):
    """The matching-order network (bms2_mlp_*) on one queue."""

    _destroy = "bms2_mlp_destroy"

    def __init__(self, q: Queue, seed: int = 0):
        self.q = q
        h = C.c_void_p()
        check(_L().bms2_mlp_create(q.h, int(seed), C.byref(h)))
        self._bind(h)
        self.num_params = int(_L().bms2_mlp_num_params())

    def get_params(self) -> np.ndarray:
        p = np.empty(self.num_params, np.float32)
        check(_L().bms2_mlp_get_params(self.h, p.ctypes.data_as(C.c_void_p), self.num_params))
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.num_params:
            raise ValueError(f"expected {self.num_params} parameters, got {p.size}")
        check(_L().bms2_mlp_set_params(self.h, p.ctypes.data_as(C.c_void_p), self.num_params))

    def forward(self, x: np.ndarray, bf16: bool = True) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        rows = x.shape[0]
        if rows == 0:
            return np.zeros((0, NUM_OUTPUTS), np.float32)
        dx = DeviceBuffer.from_array(self.q, x)
        dy = DeviceBuffer(self.q, 4 * rows * NUM_OUTPUTS, zero=False)
        check(_L().bms2_mlp_forward(self.h, C.c_void_p(dx.ptr), rows, C.c_void_p(dy.ptr),
                                    int(bf16)))
        return dy.to_numpy(np.float32, rows * NUM_OUTPUTS).reshape(rows, NUM_OUTPUTS)

    def loss_grad(self, x: np.ndarray, t: np.ndarray, depth_weight: float = 1.0):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        dx, dt = DeviceBuffer.from_array(self.q, x), DeviceBuffer.from_array(self.q, t)
        loss = C.c_float(0)
        g = np.empty(self.num_params, np.float32)
        check(_L().bms2_mlp_loss_grad(self.h, C.c_void_p(dx.ptr), C.c_void_p(dt.ptr), x.shape[0],
                                      float(depth_weight), C.byref(loss),
                                      g.ctypes.data_as(C.c_void_p)))
        return float(loss.value), g

    def train_step(self, dx: DeviceBuffer, dt: DeviceBuffer, rows: int, lr: float,
                   depth_weight: float = 1.0, beta1=0.9, beta2=0.999, eps=1e-8) -> float:
        opt = _lib.Adam(lr, beta1, beta2, eps)
        loss = C.c_float(0)
        check(_L().bms2_mlp_train_step(self.h, C.c_void_p(dx.ptr), C.c_void_p(dt.ptr), int(rows),
                                       float(depth_weight), C.byref(opt), C.byref(loss)))
        return float(loss.value)

    def save(self, path: str | Path, meta: dict | None = None) -> None:
        np.savez(path, params=self.get_params(),
                 meta=np.frombuffer(json.dumps(meta or {}).encode(), np.uint8))

    @staticmethod
    def load(q: Queue, path: str | Path = DEFAULT_WEIGHTS) -> "OrderNet":
        with np.load(path, allow_pickle=False) as z:
            p = z["params"]
        net = OrderNet(q)
        net.set_params(p)
        return net


# ---------------------------------------------------------------- query structure and features

@dataclass
class QueryInfo:
    n: int
    adj: list            # neighbour sets (local ids)
    deg: np.ndarray
    dist: np.ndarray     # all-pairs BFS distances (n x n, large for unreachable)
    labels: np.ndarray
    on_cycle: np.ndarray
    static: np.ndarray | None = None   # structure-only feature columns, computed once
    nbr: np.ndarray | None = None      # neighbour lists flattened (local ids)
    nbr_ptr: np.ndarray | None = None  # [n + 1] offsets into nbr


def query_infos(queries) -> list[QueryInfo]:
    go = np.asarray(queries.graph_offsets, np.int64)
    ro = np.asarray(queries.row_offsets, np.int64)
    ci = np.asarray(queries.column_indices, np.int64)
    lab = np.asarray(queries.node_labels)
    out = []
    for q in range(queries.num_graphs):
        b, e = go[q], go[q + 1]
        n = int(e - b)
        adj = [set((ci[ro[b + u]:ro[b + u + 1]] - b).tolist()) for u in range(n)]
        deg = np.array([len(a) for a in adj], np.int64)
        dist = np.full((n, n), 99, np.int64)
        for s in range(n):
            dist[s, s] = 0
            frontier = [s]
            while frontier:
                nxt = []
                for u in frontier:
                    for v in adj[u]:
                        if dist[s, v] == 99:
                            dist[s, v] = dist[s, u] + 1
                            nxt.append(v)
                frontier = nxt
        # a node lies on a cycle when one of its edges is not a bridge: u-v is a bridge iff v is
        # unreachable from u once the edge is removed
        on_cycle = np.zeros(n, bool)
        for u in range(n):
            for v in adj[u]:
                seen, stack = {u}, [u]
                while stack:
                    x = stack.pop()
                    for y in adj[x]:
                        if (x, y) in ((u, v), (v, u)) or y in seen:
                            continue
                        seen.add(y)
                        stack.append(y)
                if v in seen:
                    on_cycle[u] = True
                    break
        out.append(QueryInfo(n, adj, deg, dist, lab[b:e].copy(), on_cycle))
    return out


def _static_features(qi: QueryInfo) -> np.ndarray:
    """Feature columns of one query that depend on its structure only (cached on the info);
    the label-rarity and domain columns are filled per batch."""
    if qi.static is not None:
        return qi.static
    n = qi.n
    X = np.zeros((n, NUM_FEATURES), np.float32)
    qi.nbr = np.array([v for u in range(n) for v in sorted(qi.adj[u])], np.int64)
    qi.nbr_ptr = np.concatenate([[0], np.cumsum([len(a) for a in qi.adj])]).astype(np.int64)
    if n:
        m = int(qi.deg.sum()) // 2
        reach = np.where(qi.dist < 99, qi.dist, 0)
        ecc = reach.max(axis=1)
        X[:, 0] = n / 16
        X[:, 1] = m / 16
        X[:, 2] = (m - n + 1) / 4
        X[:, 3] = qi.deg.max() / 4
        X[:, 4] = ecc.max() / 8
        X[:, 5] = len(set(qi.labels.tolist())) / 8
        X[:, 8] = qi.deg / 4
        X[:, 9] = ecc / 8
        X[:, 10] = reach.sum(axis=1) / max(n - 1, 1) / 4
        for u in range(n):
            nb = list(qi.adj[u])
            X[u, 13] = sum(1 for v in nb if qi.deg[v] == 1) / 4
            X[u, 14] = sum(int(qi.deg[v]) for v in nb) / 16
            lab = int(qi.labels[u])
            X[u, 16] = float(lab == _WILDCARD)
            if lab in _ELEMENTS:
                X[u, 17 + _ELEMENTS.index(lab)] = 1.0
            elif lab != _WILDCARD:
                X[u, 24] = 1.0
        X[:, 15] = qi.on_cycle.astype(np.float32)
        X[:, 25] = np.arange(n) / max(n - 1, 1)
        X[:, 28] = np.argsort(np.argsort(-qi.deg, kind="stable"), kind="stable") / max(n - 1, 1)
        X[int(np.argmax(qi.deg)), 30] = 1.0
    qi.static = X
    return X


def static_features(queries) -> np.ndarray:
    """The structure-only columns of features() for every node of a query batch, computed
    natively (bms2_query_static_features); equal to _static_features row for row."""
    X = np.zeros((queries.total_nodes, NUM_FEATURES), np.float32)
    c = queries.to_c()
    check(_L().bms2_query_static_features(C.byref(c), X.ctypes.data_as(C.c_void_p)))
    return X


def score_orders(queries, scores: np.ndarray, use: np.ndarray) -> np.ndarray:
    """[Q][32] u8 matching orders: score_order for the queries with use[q], identity for the
    others (bms2_score_orders)."""
    Q = queries.num_graphs
    order = np.zeros((Q, 32), np.uint8)
    sc = np.ascontiguousarray(scores, np.float32)
    u = np.ascontiguousarray(use, np.uint8)
    c = queries.to_c()
    check(_L().bms2_score_orders(C.byref(c), sc.ctypes.data_as(C.c_void_p),
                                 u.ctypes.data_as(C.c_void_p), order.ctypes.data_as(C.c_void_p)))
    return order


def features(infos: list[QueryInfo], domsum: np.ndarray, qpairs: np.ndarray) -> np.ndarray:
    """[total query nodes][32] f32 features, rows in global query node order: query size, edges,
    cycle rank, largest degree, diameter, distinct labels, traced GMCR pairs and mean domain;
    per node degree, eccentricity, closeness, domain size and its ratio to the query mean, leaf
    neighbours, neighbour degrees, ring membership, wildcard / element one-hot, position, smallest
    neighbour domain, label rarity in the batch, degree and domain ranks, largest-degree and
    smallest-domain flags."""
    total = sum(qi.n for qi in infos)
    if total == 0:
        return np.zeros((0, NUM_FEATURES), np.float32)
    X = np.concatenate([_static_features(qi) for qi in infos])
    labels = np.concatenate([qi.labels for qi in infos]).astype(np.int64)
    sizes = np.array([qi.n for qi in infos], np.int64)
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    nbr = np.concatenate([qi.nbr + st for qi, st in zip(infos, starts)])
    deg = np.concatenate([np.diff(qi.nbr_ptr) for qi in infos])
    _batch_columns(X, labels, sizes, nbr, deg, domsum, qpairs)
    return X


def batch_features(queries, domsum: np.ndarray, qpairs: np.ndarray,
                   static: np.ndarray | None = None) -> np.ndarray:
    """features() of a whole query batch from its CSR: the structure-only columns natively (or
    `static`, a copy of them made earlier), the per-batch columns as in features()."""
    total = queries.total_nodes
    if total == 0:
        return np.zeros((0, NUM_FEATURES), np.float32)
    X = static_features(queries) if static is None else np.array(static, np.float32)
    labels = np.asarray(queries.node_labels).astype(np.int64)
    sizes = np.diff(np.asarray(queries.graph_offsets, np.int64))
    nbr = np.asarray(queries.column_indices, np.int64)
    deg = np.diff(np.asarray(queries.row_offsets, np.int64))
    _batch_columns(X, labels, sizes, nbr, deg, domsum, qpairs)
    return X


def _batch_columns(X: np.ndarray, labels: np.ndarray, sizes: np.ndarray, nbr: np.ndarray,
                   deg: np.ndarray, domsum: np.ndarray, qpairs: np.ndarray) -> None:
    """Columns of features() that depend on the batch and the trace: label rarity, traced pairs,
    domain sizes, smallest neighbour domain, domain rank.  nbr / deg: every node's neighbours
    (global node ids, any order, repeats allowed) and their count."""
    total = X.shape[0]
    freq = np.bincount(labels, minlength=256).astype(np.float64) / total
    X[:, 27] = -np.log2(np.maximum(freq[labels], 1e-9)) / 8
    Qn = len(sizes)
    qid = np.repeat(np.arange(Qn), sizes)
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    pairs = np.asarray(qpairs, np.float64)[:Qn]
    dom = np.where(pairs[qid] > 0,
                   np.asarray(domsum[:total], np.float64) / np.maximum(pairs[qid], 1), 0.0)
    nz = sizes > 0
    meandom = np.zeros(Qn)
    meandom[nz] = np.add.reduceat(dom, starts[nz]) / sizes[nz]
    X[:, 6] = (np.log2(1 + pairs) / 16)[qid]
    X[:, 7] = (np.log2(1 + meandom) / 6)[qid]
    X[:, 11] = np.log2(1 + dom) / 6
    X[:, 12] = np.log2((1 + dom) / (1 + meandom[qid])) / 4
    # smallest neighbour domain
    has = deg > 0
    if nbr.size:
        ptr = np.concatenate([[0], np.cumsum(deg)[:-1]])
        mins = np.minimum.reduceat(dom[nbr], ptr[has])
        X[has, 26] = np.log2(1 + mins) / 6
    # domain rank inside each query (ties by local id) and the smallest-domain flag
    local = np.arange(total) - starts[qid]
    order = np.lexsort((local, dom, qid))
    rank = np.empty(total, np.int64)
    rank[order] = np.arange(total) - starts[qid[order]]
    X[:, 29] = rank / np.maximum(sizes[qid] - 1, 1)
    X[:, 31] = (rank == 0).astype(np.float32)


# ---------------------------------------------------------------- orders

def greedy_order(qi: QueryInfo, root: int, dom: np.ndarray) -> list[int]:
    """Connected greedy order from root: next = most already-matched neighbours, then smaller
    mean domain, then higher degree, then lower id."""
    order, S = [root], {root}
    while len(order) < qi.n:
        best, key = -1, None
        for u in range(qi.n):
            if u in S:
                continue
            k = (-len(qi.adj[u] & S), dom[u], -int(qi.deg[u]), u)
            if key is None or k < key:
                best, key = u, k
        order.append(best)
        S.add(best)
    return order


def score_order(qi: QueryInfo, scores: np.ndarray) -> list[int]:
    """Connected order by descending score: start at the best-scored node, then always the
    best-scored node adjacent to the matched set (any remaining node when none is)."""
    n = qi.n
    order = [int(np.argmax(scores))]
    S = set(order)
    while len(order) < n:
        front = [u for u in range(n) if u not in S and qi.adj[u] & S]
        pool = front or [u for u in range(n) if u not in S]
        u = max(pool, key=lambda v: (scores[v], -v))
        order.append(u)
        S.add(u)
    return order


def candidate_orders(infos: list[QueryInfo], domsum: np.ndarray, qpairs: np.ndarray):
    """List of (name, order [Q][32] u8): identity, reference max-domain root, greedy from root r
    (queries with fewer nodes repeat their identity order)."""
    Q = len(infos)
    maxn = max((qi.n for qi in infos), default=0)
    offs = np.cumsum([0] + [qi.n for qi in infos])
    doms = [domsum[offs[q]:offs[q + 1]].astype(np.float64) / max(float(qpairs[q]), 1.0)
            for q in range(Q)]
    ident = np.zeros((Q, 32), np.uint8)
    for q, qi in enumerate(infos):
        ident[q, :qi.n] = np.arange(qi.n)
    out = [("identity", ident)]
    ref = ident.copy()
    for q, qi in enumerate(infos):
        if qi.n:
            root = int(np.argmax(doms[q]))  # defineMatchingOrder: max-domain root, then identity
            ref[q, :qi.n] = [root] + [u for u in range(qi.n) if u != root]
    out.append(("ref_maxdom", ref))
    for r in range(maxn):
        o = ident.copy()
        for q, qi in enumerate(infos):
            if r < qi.n:
                o[q, :qi.n] = greedy_order(qi, r, doms[q])
        out.append((f"greedy_root{r}", o))
    return out


# ---------------------------------------------------------------- traces and labels

@dataclass
class SeedTraces:
    seed: int
    X: np.ndarray
    T: np.ndarray
    pairs_traced: int
    identity_nodes: int
    best_nodes: int
    info: dict = field(default_factory=dict)


def _trace(pipe: Pipeline, graphs: int, qn: int, order=None, pair_nodes=False) -> dict:
    pipe.set_order(order)
    return pipe.trace(graphs, qn, pair_nodes=pair_nodes)


def collect_seed(q: Queue, seed: int, num_data_graphs: int = 10000, K: int = 6,
                 workload: str = "molecules-1m", depth_labels: bool = True,
                 num_query_graphs: int = 0) -> SeedTraces:
    """Plain runs on one seed's workload: candidate orders and depths, as training rows.
    num_query_graphs overrides the workload's query count (0 keeps it)."""
    from . import graphs as G
    from .api import DeviceBatchedCSRGraph
    over = {"num_query_graphs": num_query_graphs} if num_query_graphs else {}
    data, queries = G.generate(workload, seed=seed, num_data_graphs=num_data_graphs, **over)
    dq, dd = DeviceBatchedCSRGraph(q, queries), DeviceBatchedCSRGraph(q, data)
    qn = queries.total_nodes
    infos = query_infos(queries)
    pipe = Pipeline(q, dq, dd, refinement_steps=K)
    base = _trace(pipe, data.num_graphs, qn, None, pair_nodes=True)
    domsum, qpairs = base["domsum"], base["qpairs"]
    cands = candidate_orders(infos, domsum, qpairs)
    nodes = np.zeros((len(cands), queries.num_graphs), np.float64)
    for i, (name, o) in enumerate(cands):
        tr = base if name == "identity" else _trace(pipe, data.num_graphs, qn, o)
        if name != "identity" and not np.array_equal(tr["counts"], base["counts"]):
            raise RuntimeError(f"seed {seed}: counts changed under order {name}")
        nodes[i] = tr["qnodes"]
    pipe.set_order(None)
    del pipe
    best = np.argmin(nodes, axis=0)  # ties keep the earliest candidate (identity first)
    # useful refinement depth: smallest k whose GMCR pairs and DFS nodes are within 1% of k = 8
    depth_idx = np.zeros(queries.num_graphs, np.int64)
    if depth_labels:
        per_k = []
        for k in range(1, 9):
            pk = Pipeline(q, dq, dd, refinement_steps=k)
            tr = pk.trace(data.num_graphs, qn)
            per_k.append((tr["qpairs"].astype(np.float64), tr["qnodes"].astype(np.float64)))
            del pk
        p8, n8 = per_k[-1]
        for qq in range(queries.num_graphs):
            for k in range(8):
                pk_, nk_ = per_k[k][0][qq], per_k[k][1][qq]
                if pk_ <= 1.01 * p8[qq] + 1 and nk_ <= 1.01 * n8[qq] + 1:
                    depth_idx[qq] = k
                    break
    X = features(infos, domsum, qpairs)
    T = np.zeros((X.shape[0], 4), np.float32)
    r = 0
    keep = np.zeros(X.shape[0], bool)
    for qq, qi in enumerate(infos):
        n = qi.n
        o = cands[best[qq]][1][qq, :n]
        pos = np.empty(n, np.int64)
        pos[o] = np.arange(n)
        ident_nodes, best_nodes = nodes[0, qq], nodes[best[qq], qq]
        T[r:r + n, 0] = 1.0 - pos / max(n - 1, 1)
        T[r:r + n, 1] = np.log2((best_nodes + 1) / (ident_nodes + 1))
        T[r:r + n, 2] = depth_idx[qq]
        T[r:r + n, 3] = np.log2(2 + ident_nodes)
        keep[r:r + n] = qpairs[qq] > 0 and n > 1
        r += n
    return SeedTraces(seed, X[keep], T[keep], int((base["pair_nodes"] > 0).sum()),
                      int(nodes[0].sum()), int(nodes[best, np.arange(queries.num_graphs)].sum()),
                      dict(candidates=len(cands), queries=queries.num_graphs,
                           best_counts={cands[i][0]: int((best == i).sum())
                                        for i in range(len(cands)) if (best == i).any()}))


def train(net: OrderNet, X: np.ndarray, T: np.ndarray, epochs: int = 40, batch: int = 4096,
          lr: float = 2e-3, depth_weight: float = 0.25, seed: int = 0, log=None) -> list[float]:
    rng = np.random.default_rng(seed)
    rows = X.shape[0]
    losses = []
    for ep in range(epochs):
        perm = rng.permutation(rows)
        tot, nb = 0.0, 0
        for s in range(0, rows, batch):
            idx = perm[s:s + batch]
            dx = DeviceBuffer.from_array(net.q, np.ascontiguousarray(X[idx], np.float32))
            dt = DeviceBuffer.from_array(net.q, np.ascontiguousarray(T[idx], np.float32))
            tot += net.train_step(dx, dt, len(idx), lr * (0.5 ** (ep // 15)), depth_weight)
            nb += 1
        losses.append(tot / max(nb, 1))
        if log:
            log(f"epoch {ep}: loss {losses[-1]:.4f}")
    return losses


# ---------------------------------------------------------------- planning

@dataclass
class Plan:
    order: np.ndarray          # [Q][32] u8
    learned: np.ndarray        # [Q] bool: learned order used (False: identity fallback)
    depth: np.ndarray          # [Q] predicted useful refinement depth 1..8 (reported only)
    predicted_log_ratio: np.ndarray
    sample_graphs: int
    seconds: float


def plan(pipe: Pipeline, net: OrderNet, queries, sample_graphs: int = 16384,
         budget_log_ratio: float = -0.02, static: np.ndarray | None = None) -> Plan:
    """Traces a sample of the resident data graphs (identity order), predicts per query node
    with the network in bf16, sets the pipeline's matching order and returns the plan.  The
    query structure (features, orders) is handled natively; `static` (static_features of these
    queries, nullable) skips recomputing it."""
    t0 = time.perf_counter()
    pipe.set_order(None)
    tr = pipe.trace(sample_graphs, queries.total_nodes, counts=False, joins=False)
    X = batch_features(queries, tr["domsum"], tr["qpairs"], static)
    Y = net.forward(X, bf16=True)
    Q = queries.num_graphs
    sizes = np.diff(np.asarray(queries.graph_offsets, np.int64))
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    ratio = np.zeros(Q, np.float32)
    depth = np.zeros(Q, np.int64)
    nz = sizes > 0
    if nz.any():
        # per-query means of the predicted log cost ratio and of the depth scores
        sums = np.add.reduceat(Y.astype(np.float64), starts[nz], axis=0)
        means = sums / sizes[nz][:, None]
        ratio[nz] = means[:, 1]
        depth[nz] = 1 + np.argmax(means[:, 2:], axis=1)
    learned = (sizes > 1) & (ratio < budget_log_ratio)
    order = score_orders(queries, Y[:, 0], learned)
    pipe.set_order(order)
    return Plan(order, learned, depth, ratio, tr["graphs"], time.perf_counter() - t0)


# ---------------------------------------------------------------- command line

def _main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m bms2_amd.learned")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train", help="collect traces on seeds and train the order network")
    t.add_argument("--seeds", default="1-20")
    t.add_argument("--graphs-per-seed", type=int, default=10000)
    t.add_argument("--queries-per-seed", type=int, default=1400,
                   help="queries generated per seed (1,400 x 20 seeds: about 200,000 query-node rows)")
    t.add_argument("--epochs", type=int, default=40)
    t.add_argument("--out", default=str(DEFAULT_WEIGHTS))
    t.add_argument("--no-depth", action="store_true")
    e = sub.add_parser("evaluate", help="DFS nodes and counts, learned vs identity order")
    e.add_argument("--seed", type=int, default=100)
    e.add_argument("--graphs", type=int, default=100000)
    e.add_argument("--weights", default=str(DEFAULT_WEIGHTS))
    args = ap.parse_args(argv)
    log = lambda s: print(s, file=sys.stderr, flush=True)  # noqa: E731
    q = Queue(0)
    if args.cmd == "train":
        a, b = (int(x) for x in args.seeds.split("-"))
        Xs, Ts, meta = [], [], dict(seeds=args.seeds, graphs_per_seed=args.graphs_per_seed,
                                    queries_per_seed=args.queries_per_seed, per_seed=[])
        traced = 0
        for s in range(a, b + 1):
            t0 = time.perf_counter()
            st = collect_seed(q, s, args.graphs_per_seed, depth_labels=not args.no_depth,
                              num_query_graphs=args.queries_per_seed)
            Xs.append(st.X)
            Ts.append(st.T)
            traced += st.pairs_traced
            meta["per_seed"].append(dict(seed=s, rows=int(st.X.shape[0]), pairs=st.pairs_traced,
                                         identity_nodes=st.identity_nodes,
                                         best_nodes=st.best_nodes, **st.info))
            log(f"seed {s}: {st.X.shape[0]} rows, {st.pairs_traced} traced pairs, DFS nodes "
                f"identity {st.identity_nodes} best {st.best_nodes} "
                f"({time.perf_counter() - t0:.1f}s)")
        X, T = np.concatenate(Xs), np.concatenate(Ts)
        net = OrderNet(q, seed=1)
        losses = train(net, X, T, epochs=args.epochs, log=log)
        meta.update(rows=int(X.shape[0]), traced_pairs=traced, final_loss=losses[-1],
                    first_loss=losses[0])
        net.save(args.out, meta)
        print(json.dumps(dict(rows=int(X.shape[0]), traced_pairs=traced, first_loss=losses[0],
                              final_loss=losses[-1], out=args.out)))
        return 0
    if args.cmd == "evaluate":
        from . import graphs as G
        from .api import DeviceBatchedCSRGraph
        data, queries = G.generate("molecules-1m", seed=args.seed, num_data_graphs=args.graphs)
        dq, dd = DeviceBatchedCSRGraph(q, queries), DeviceBatchedCSRGraph(q, data)
        pipe = Pipeline(q, dq, dd, refinement_steps=6)
        net = OrderNet.load(q, args.weights)
        ident = pipe.trace(data.num_graphs, queries.total_nodes)
        pl = plan(pipe, net, queries)
        lrn = pipe.trace(data.num_graphs, queries.total_nodes)
        same = bool(np.array_equal(ident["counts"], lrn["counts"]))
        out = dict(seed=args.seed, graphs=data.num_graphs,
                   identity_dfs_nodes=int(ident["qnodes"].sum()),
                   learned_dfs_nodes=int(lrn["qnodes"].sum()), counts_equal=same,
                   learned_queries=int(pl.learned.sum()), plan_seconds=round(pl.seconds, 3))
        print(json.dumps(out))
        return 0 if same else 1
    return 2


if __name__ == "__main__":
    sys.exit(_main())
