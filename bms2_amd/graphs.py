"""Host-side graph batches (numpy) and the synthetic workloads.

A GraphBatch is the packed batched-CSR layout of sigmo::DeviceBatchedCSRGraph
(graph.hpp:58-66) held in numpy arrays on the host: graph_offsets (node prefix over graphs),
row_offsets, column_indices (global node ids), node_labels, edge_labels.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib

# Element and bond codes of the reference's converter (generate_data.py:8-17).
LABELS = {"N": 0, "Cl": 1, "*": 2, "Br": 3, "I": 4, "P": 5, "H": 6, "O": 7, "C": 8, "S": 9, "F": 10}
BONDS = {"unspecified": 0, "single": 1, "double": 2, "triple": 3, "aromatic": 4}


@dataclass
class GraphBatch:
    graph_offsets: np.ndarray  # u32 [G+1]
    row_offsets: np.ndarray  # u32 [N+1]
    column_indices: np.ndarray  # u32 [E]
    node_labels: np.ndarray  # u8 [N]
    edge_labels: np.ndarray  # u8 [E]

    @property
    def num_graphs(self) -> int:
        return int(self.graph_offsets.shape[0] - 1)

    @property
    def total_nodes(self) -> int:
        return int(self.node_labels.shape[0])

    @property
    def total_edges(self) -> int:
        return int(self.column_indices.shape[0])

    def graph_sizes(self) -> np.ndarray:
        return np.diff(self.graph_offsets.astype(np.int64))

    def degrees(self) -> np.ndarray:
        return np.diff(self.row_offsets.astype(np.int64))

    def to_c(self) -> _lib.HostCsr:
        """A HostCsr view of these arrays (valid while this batch is alive)."""
        for a, dt in ((self.graph_offsets, np.uint32), (self.row_offsets, np.uint32),
                      (self.column_indices, np.uint32), (self.node_labels, np.uint8),
                      (self.edge_labels, np.uint8)):
            assert a.dtype == dt and a.flags["C_CONTIGUOUS"]
        h = _lib.HostCsr()
        h.num_graphs = self.num_graphs
        h.total_nodes = self.total_nodes
        h.total_edges = self.total_edges
        h.graph_offsets = self.graph_offsets.ctypes.data_as(_lib.u32p)
        h.row_offsets = self.row_offsets.ctypes.data_as(_lib.u32p)
        h.column_indices = self.column_indices.ctypes.data_as(_lib.u32p)
        h.node_labels = self.node_labels.ctypes.data_as(_lib.u8p)
        h.edge_labels = self.edge_labels.ctypes.data_as(_lib.u8p)
        return h

    @staticmethod
    def from_c(h: _lib.HostCsr) -> "GraphBatch":
        """Copies a library-allocated HostCsr into numpy and frees it."""
        G, N, E = int(h.num_graphs), int(h.total_nodes), int(h.total_edges)

        def arr(ptr, n, dt):
            if n == 0:
                return np.zeros(0, dtype=dt)
            return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True)

        try:
            return GraphBatch(
                graph_offsets=arr(h.graph_offsets, G + 1, np.uint32) if G or h.graph_offsets
                else np.zeros(1, np.uint32),
                row_offsets=arr(h.row_offsets, N + 1, np.uint32),
                column_indices=arr(h.column_indices, E, np.uint32),
                node_labels=arr(h.node_labels, N, np.uint8),
                edge_labels=arr(h.edge_labels, E, np.uint8),
            )
        finally:
            _lib.lib().bms2_host_csr_free(C.byref(h))

    def subset(self, begin: int, end: int) -> "GraphBatch":
        """Graphs [begin, end) as a new batch with ids rebased to 0."""
        n0, n1 = int(self.graph_offsets[begin]), int(self.graph_offsets[end])
        e0, e1 = int(self.row_offsets[n0]), int(self.row_offsets[n1])
        return GraphBatch(
            graph_offsets=(self.graph_offsets[begin:end + 1] - n0).astype(np.uint32),
            row_offsets=(self.row_offsets[n0:n1 + 1] - e0).astype(np.uint32),
            column_indices=(self.column_indices[e0:e1] - n0).astype(np.uint32),
            node_labels=self.node_labels[n0:n1].copy(),
            edge_labels=self.edge_labels[e0:e1].copy(),
        )

    def local_arrays(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(node count per graph, row end of every node counted from its graph's first CSR
        entry, neighbour ids local to their graph): the input of bms2_device_csr_pack."""
        go = np.asarray(self.graph_offsets, np.int64)
        ro = np.asarray(self.row_offsets, np.int64)
        sizes = np.diff(go)
        node_graph = np.repeat(np.arange(self.num_graphs), sizes)
        ro_local = ro[1:] - ro[go[:-1]][node_graph]
        ent_graph = np.repeat(node_graph, np.diff(ro))
        ci_local = np.asarray(self.column_indices, np.int64) - go[:-1][ent_graph]
        return (sizes.astype(np.uint32), ro_local.astype(np.uint32), ci_local.astype(np.uint32))

    def graph(self, g: int) -> tuple[np.ndarray, list[tuple[int, int]]]:
        """(labels, undirected edge list with local ids, u < v) of graph g."""
        n0, n1 = int(self.graph_offsets[g]), int(self.graph_offsets[g + 1])
        edges = []
        for u in range(n0, n1):
            for k in range(int(self.row_offsets[u]), int(self.row_offsets[u + 1])):
                v = int(self.column_indices[k])
                if u < v:
                    edges.append((u - n0, v - n0))
        return self.node_labels[n0:n1].copy(), edges

    @staticmethod
    def from_graphs(graphs: list[tuple[list[int], list[tuple]]]) -> "GraphBatch":
        """Packs (labels, edges) graphs; edges are (u, v) or (u, v, label), each undirected edge
        appended under both endpoints in list order (toCSRGraph, graph.hpp:212-243)."""
        go = [0]
        ro = [0]
        ci: list[int] = []
        el: list[int] = []
        nl: list[int] = []
        for labels, edges in graphs:
            n = len(labels)
            base = go[-1]
            adj: list[list[tuple[int, int]]] = [[] for _ in range(n)]
            for e in edges:
                u, v = int(e[0]), int(e[1])
                lab = int(e[2]) if len(e) > 2 else 0
                if u == v or not (0 <= u < n and 0 <= v < n):
                    raise ValueError(f"bad edge {e} in graph of {n} nodes")
                adj[u].append((v, lab))
                adj[v].append((u, lab))
            for u in range(n):
                for v, lab in adj[u]:
                    ci.append(base + v)
                    el.append(lab)
                ro.append(len(ci))
            nl.extend(int(x) for x in labels)
            go.append(base + n)
        return GraphBatch(np.array(go, np.uint32), np.array(ro, np.uint32),
                          np.array(ci, np.uint32), np.array(nl, np.uint8),
                          np.array(el, np.uint8))

    def digest(self) -> str:
        import hashlib
        h = hashlib.sha256()
        for a in (self.graph_offsets, self.row_offsets, self.column_indices, self.node_labels,
                  self.edge_labels):
            h.update(np.ascontiguousarray(a).tobytes())
        return h.hexdigest()


# ---------------------------------------------------------------------------- workloads

def _spec(**kw) -> _lib.WorkloadSpec:
    s = _lib.WorkloadSpec()
    s.data_size_dist = 0
    s.data_min_nodes, s.data_max_nodes = 10, 40
    s.data_lognormal_mu, s.data_lognormal_sigma = 3.2, 0.5
    s.poly_min_rings, s.poly_max_rings = 2, 8
    s.max_ring_closures = 3
    s.query_shapes = 0
    for k, v in kw.items():
        setattr(s, k, v)
    return s


WORKLOADS = {
    # 2,000 molecules of 10-40 atoms x 64 queries of 4-10 nodes, half cut from the data graphs.
    "molecules-2k": dict(num_data_graphs=2000, num_query_graphs=64, query_min_nodes=4,
                         query_max_nodes=10, extracted_permille=500),
    # 1,000,000 molecules x 1,024 queries of 3-16 nodes, 70% cut from the data graphs.
    "molecules-1m": dict(num_data_graphs=1_000_000, num_query_graphs=1024, query_min_nodes=3,
                         query_max_nodes=16, extracted_permille=700),
    # Uneven sizes: lognormal(3.2, 0.5) clipped to 5-250 atoms, 10% fused ring systems of 2-8
    # rings; queries of 3-32 nodes with ring and branched shapes.
    "uneven-1m": dict(num_data_graphs=1_000_000, num_query_graphs=1024, data_size_dist=1,
                      data_min_nodes=5, data_max_nodes=250, polycyclic_permille=100,
                      query_min_nodes=3, query_max_nodes=32, extracted_permille=700,
                      query_shapes=1),
    # Resident set for the serving loop.
    "molecules-2m": dict(num_data_graphs=2_000_000, num_query_graphs=256, query_min_nodes=3,
                         query_max_nodes=24, extracted_permille=700),
}


def generate(workload: str | dict = "molecules-2k", seed: int = 0, threads: int = 0,
             **overrides) -> tuple[GraphBatch, GraphBatch]:
    """(data, queries) of a named workload (or an explicit spec dict), deterministic in seed."""
    kw = dict(WORKLOADS[workload]) if isinstance(workload, str) else dict(workload)
    kw.update(overrides)
    spec = _spec(seed=seed, num_threads=threads, **kw)
    d, q = _lib.HostCsr(), _lib.HostCsr()
    _lib.check(_lib.lib().bms2_generate_workload(C.byref(spec), C.byref(d), C.byref(q)))
    data = GraphBatch.from_c(d)
    queries = GraphBatch.from_c(q)
    return data, queries


def _graph_adj(b: GraphBatch, g: int) -> tuple[list[int], list[dict[int, int]]]:
    """Labels and adjacency (local neighbour -> edge label) of graph g of a batch."""
    n0, n1 = int(b.graph_offsets[g]), int(b.graph_offsets[g + 1])
    adj: list[dict[int, int]] = [dict() for _ in range(n1 - n0)]
    for u in range(n0, n1):
        for k in range(int(b.row_offsets[u]), int(b.row_offsets[u + 1])):
            adj[u - n0][int(b.column_indices[k]) - n0] = int(b.edge_labels[k])
    return [int(x) for x in b.node_labels[n0:n1]], adj


def _induced(labels, adj, nodes) -> tuple[list[int], list[tuple[int, int, int]]]:
    """Induced subgraph on `nodes`, numbered in the given order."""
    pos = {v: i for i, v in enumerate(nodes)}
    edges = [(pos[u], pos[v], lab) for u in nodes for v, lab in sorted(adj[u].items())
             if v in pos and pos[u] < pos[v]]
    return [labels[v] for v in nodes], edges


def _grow(adj, start: list[int], size: int, rng) -> list[int]:
    """Connected growth of a node set: random frontier nodes appended until `size` nodes."""
    chosen, inset = list(start), set(start)
    while len(chosen) < size:
        frontier = sorted({v for u in chosen for v in adj[u] if v not in inset})
        if not frontier:
            break
        v = frontier[int(rng.integers(len(frontier)))]
        chosen.append(v)
        inset.add(v)
    return chosen


def _induced_embedding(plabels, pedges, labels, adj) -> list[int] | None:
    """One label-preserving induced embedding of a small pattern (backtracking), or None."""
    n = len(plabels)
    padj = [set() for _ in range(n)]
    for u, v, _ in pedges:
        padj[u].add(v)
        padj[v].add(u)
    img: list[int] = []

    def extend(i: int) -> bool:
        if i == n:
            return True
        for c in range(len(labels)):
            if labels[c] != plabels[i] or c in img:
                continue
            if all((j in padj[i]) == (img[j] in adj[c]) for j in range(i)):
                img.append(c)
                if extend(i + 1):
                    return True
                img.pop()
        return False

    return list(img) if extend(0) else None


def cluster_queries(data: GraphBatch, num_members: int, seed: int = 0,
                    proto_nodes: tuple[int, int] = (4, 7),
                    member_nodes: tuple[int, int] = (6, 16)) -> GraphBatch:
    """A query batch in the cluster convention of the reference's CS mode (generate_cs.py:338-354,
    556-569): query 0 is the prototype, every member starts with an induced copy of it whose nodes
    come first in the prototype's order, followed by the member's own nodes.  The prototype is a
    connected induced subgraph of a data graph; members are connected induced subgraphs grown
    around an induced image of the prototype in data graphs.  Deterministic in seed."""
    rng = np.random.default_rng(seed)
    G = data.num_graphs
    sizes = np.diff(np.asarray(data.graph_offsets, np.int64))
    p = int(rng.integers(proto_nodes[0], proto_nodes[1] + 1))
    big = np.flatnonzero(sizes >= p)
    if len(big) == 0:
        raise ValueError("no data graph holds a prototype of that size")
    while True:
        g0 = int(big[rng.integers(len(big))])
        labels0, adj0 = _graph_adj(data, g0)
        S = _grow(adj0, [int(rng.integers(len(labels0)))], p, rng)
        if len(S) == p:
            break
    proto = _induced(labels0, adj0, S)
    out = [proto]
    for _ in range(num_members):
        img, labels, adj = None, None, None
        for _ in range(32):  # another data graph holding the prototype, else its source
            g = int(rng.integers(G))
            labels, adj = _graph_adj(data, g)
            img = _induced_embedding(proto[0], proto[1], labels, adj)
            if img is not None:
                break
        if img is None:
            labels, adj, img = labels0, adj0, list(S)
        size = int(rng.integers(member_nodes[0], member_nodes[1] + 1))
        out.append(_induced(labels, adj, _grow(adj, img, max(size, p), rng)))
    return GraphBatch.from_graphs(out)


def load_dat(path: str, max_graphs: int = 0, skip_malformed: bool = False) -> GraphBatch:
    """io::loadCSRGraphsFromFile (io.hpp:43-52)."""
    h = _lib.HostCsr()
    _lib.check(_lib.lib().bms2_load_dat(str(path).encode(), max_graphs, int(skip_malformed),
                                        C.byref(h)))
    return GraphBatch.from_c(h)


def parse_dat_lines(lines: list[str], skip_malformed: bool = False) -> GraphBatch:
    """io::loadCSRGraphsFromLines (io.hpp:34-41): one graph per line, in memory."""
    arr = (C.c_char_p * max(len(lines), 1))(*[ln.encode() for ln in lines])
    h = _lib.HostCsr()
    _lib.check(_lib.lib().bms2_parse_dat_lines(arr, len(lines), int(skip_malformed), C.byref(h)))
    return GraphBatch.from_c(h)


def write_dat(path: str, g: GraphBatch) -> None:
    h = g.to_c()
    _lib.check(_lib.lib().bms2_write_dat(str(path).encode(), C.byref(h)))
