"""Isomorphism classes of a query batch (bms2_query_classes, host): the pipeline screens one query
per class and copies its results to the rest, so a class must never join two non-isomorphic
queries.  Pinned against networkx's VF2 isomorphism test with label matching."""
import collections

import networkx as nx
import numpy as np
import pytest
from networkx.algorithms import isomorphism as iso

from bms2_amd import api, graphs
from bms2_amd.graphs import GraphBatch


def _nx(labels, edges):
    G = nx.Graph()
    for i, l in enumerate(labels):
        G.add_node(i, l=int(l))
    G.add_edges_from((int(e[0]), int(e[1])) for e in edges)
    return G


def _networkx_classes(batch: GraphBatch) -> np.ndarray:
    """rep[g] = lowest-indexed query isomorphic to g (labels preserved), by VF2 within buckets of
    the Weisfeiler-Lehman hash."""
    Gs = [_nx(*batch.graph(g)) for g in range(batch.num_graphs)]
    rep = np.arange(batch.num_graphs, dtype=np.uint32)
    buckets = collections.defaultdict(list)
    for g, G in enumerate(Gs):
        key = (G.number_of_nodes(), G.number_of_edges(),
               nx.weisfeiler_lehman_graph_hash(G, node_attr="l", iterations=4))
        for r in buckets[key]:
            if iso.GraphMatcher(Gs[r], G, node_match=lambda a, b: a["l"] == b["l"]).is_isomorphic():
                rep[g] = r
                break
        else:
            buckets[key].append(g)
    return rep


@pytest.mark.parametrize("workload", ["molecules-1m", "uneven-1m"])
def test_classes_equal_networkx(workload):
    _, queries = graphs.generate(workload, seed=0, num_data_graphs=300)
    rep = api.query_classes(queries)
    assert np.array_equal(rep, _networkx_classes(queries))
    assert (rep != np.arange(len(rep))).sum() > 20  # the workloads repeat small queries


def test_permuted_copies_join_their_original():
    """Copies of queries with their nodes renumbered and edges listed in another order are found
    isomorphic to the original (the first occurrence stands for the class)."""
    _, queries = graphs.generate("uneven-1m", seed=2, num_data_graphs=50)
    rng = np.random.default_rng(7)
    originals = [queries.graph(g) for g in range(0, queries.num_graphs, 8)]
    copies = []
    for labels, edges in originals:
        n = len(labels)
        perm = rng.permutation(n)  # old id -> new id
        new_labels = np.empty(n, np.uint8)
        new_labels[perm] = labels
        new_edges = [(int(perm[u]), int(perm[v])) for u, v in edges]
        rng.shuffle(new_edges)
        copies.append((new_labels, new_edges))
    batch = GraphBatch.from_graphs(originals + copies)
    rep = api.query_classes(batch)
    first = _networkx_classes(GraphBatch.from_graphs(originals))
    k = len(originals)
    assert np.array_equal(rep[k:], first)  # each copy joins its original's class
    assert np.array_equal(rep, _networkx_classes(batch))


def test_colour_refinement_ties_are_searched():
    """Pairs that colour refinement cannot tell apart stay apart unless an isomorphism exists:
    a 6-ring against two triangles, and the 3-prism against K3,3 (all carbon, all nodes of equal
    degree); a relabelled node splits a class."""
    C = 8
    hexagon = ([C] * 6, [(i, (i + 1) % 6) for i in range(6)])
    triangles = ([C] * 6, [(0, 1), (1, 2), (2, 0), (3, 4), (4, 5), (5, 3)])
    prism = ([C] * 6, [(0, 1), (1, 2), (2, 0), (3, 4), (4, 5), (5, 3), (0, 3), (1, 4), (2, 5)])
    k33 = ([C] * 6, [(a, b) for a in range(3) for b in range(3, 6)])
    path_co = ([C, C, 7], [(0, 1), (1, 2)])
    path_oc = ([7, C, C], [(0, 1), (1, 2)])
    path_cn = ([C, C, 0], [(0, 1), (1, 2)])
    ring_rot = ([C] * 6, [((i + 2) % 6, (i + 3) % 6) for i in range(6)])
    batch = GraphBatch.from_graphs([hexagon, triangles, prism, k33, path_co, path_oc, path_cn,
                                    ring_rot, triangles])
    rep = api.query_classes(batch)
    assert rep.tolist() == [0, 1, 2, 3, 4, 4, 6, 0, 1]
    assert np.array_equal(rep, _networkx_classes(batch))


def test_single_nodes_and_empty_batch_stay_alone():
    batch = GraphBatch.from_graphs([([8], []), ([8], []), ([8, 8], [(0, 1)]), ([8, 8], [(0, 1)])])
    assert api.query_classes(batch).tolist() == [0, 1, 2, 2]


def _networkx_orbits(batch: GraphBatch) -> np.ndarray:
    out = []
    for g in range(batch.num_graphs):
        G = _nx(*batch.graph(g))
        n = G.number_of_nodes()
        orb = list(range(n))
        for s in iso.GraphMatcher(G, G, node_match=lambda a, b: a["l"] == b["l"]).isomorphisms_iter():
            for u, v in s.items():
                orb[u] = min(orb[u], v)
        out.extend(orb)
    return np.asarray(out, np.uint8)


@pytest.mark.parametrize("workload", ["molecules-1m", "uneven-1m"])
def test_orbits_equal_networkx(workload):
    """Automorphism orbits (lowest local node of each orbit) equal networkx's, over every
    automorphism VF2 enumerates."""
    _, queries = graphs.generate(workload, seed=1, num_data_graphs=200)
    small = [g for g in range(queries.num_graphs)
             if queries.graph_offsets[g + 1] - queries.graph_offsets[g] <= 12]
    batch = GraphBatch.from_graphs([queries.graph(g) for g in small])
    orb = api.query_orbits(batch)
    local = np.concatenate([np.arange(n) for n in np.diff(batch.graph_offsets.astype(np.int64))])
    assert np.array_equal(orb, _networkx_orbits(batch))
    assert (orb != local).sum() > 50


def test_orbits_of_symmetric_shapes():
    C, O = 8, 7
    star = ([C, C, C, C, O], [(0, 1), (0, 2), (0, 3), (0, 4)])  # three equal arms, one oxygen
    ring = ([C] * 6, [(i, (i + 1) % 6) for i in range(6)])
    chain = ([O, C, C, O], [(0, 1), (1, 2), (2, 3)])
    batch = GraphBatch.from_graphs([star, ring, chain])
    assert api.query_orbits(batch).tolist() == [0, 1, 1, 1, 4] + [0] * 6 + [0, 1, 1, 0]


def _orders(batch: GraphBatch, rng):
    """A random connected matching order per query (u8 [Q, 32]): random scores through the
    planner's connected ordering (an order that matches a node before any of its neighbours
    leaves that depth unconstrained and can grow the search combinatorially)."""
    from bms2_amd import learned
    o = np.zeros((batch.num_graphs, 32), np.uint8)
    for g, qi in enumerate(learned.query_infos(batch)):
        o[g, :qi.n] = learned.score_order(qi, rng.normal(size=qi.n))
    return o


@pytest.mark.parametrize("workload", ["molecules-1m", "uneven-1m"])
@pytest.mark.parametrize("ordered", [False, True])
def test_symmetry_constraint_is_an_orbit(workload, ordered):
    """du and vmask are the depths of one automorphism orbit of at least two nodes (du the first
    matched), the one whose second-matched member comes earliest; queries without a non-trivial
    orbit get none."""
    _, queries = graphs.generate(workload, seed=4, num_data_graphs=100)
    order = _orders(queries, np.random.default_rng(3)) if ordered else None
    du, vm = api.query_symmetry(queries, order)
    orb = api.query_orbits(queries)
    go = queries.graph_offsets.astype(np.int64)
    have = 0
    for g in range(queries.num_graphs):
        n = int(go[g + 1] - go[g])
        o = order[g, :n] if ordered else np.arange(n)
        depth = np.empty(n, np.int64)
        depth[o] = np.arange(n)
        local = orb[go[g]:go[g + 1]]
        orbits = [np.flatnonzero(local == r) for r in np.unique(local)]
        big = [sorted(depth[m]) for m in orbits if len(m) > 1]
        if not big:
            assert du[g] == 0xff and vm[g] == 0
            continue
        have += 1
        best = min(big, key=lambda ds: ds[1])
        assert du[g] == best[0]
        assert int(vm[g]) == sum(1 << int(d) for d in best[1:])
    assert have > 100


@pytest.mark.parametrize("workload", ["molecules-1m", "uneven-1m"])
@pytest.mark.parametrize("find_first", [False, True])
def test_symmetry_breaking_counts(workload, find_first):
    """Under the constraint the oracle's join finds exactly count / (popcount(vmask) + 1) of every
    pair's embeddings and never more DFS nodes (existence unchanged in first-match mode); in
    identity order and in random connected orders."""
    import oracle
    data, queries = graphs.generate(workload, seed=2, num_data_graphs=150)
    r = oracle.staged(queries, data, 6, find_first, keep_domains=False)
    go = queries.graph_offsets.astype(np.int64)
    for order in (None, _orders(queries, np.random.default_rng(9))):
        flat = None
        if order is not None:
            flat = np.concatenate([order[g, :int(go[g + 1] - go[g])]
                                   for g in range(queries.num_graphs)]).astype(np.uint32)
        du, vm = api.query_symmetry(queries, order)
        _, c0, n0 = oracle.join(queries, data, r.final_domains, r.gmcr_offsets, r.gmcr_indices,
                                find_first, flat)
        _, c1, n1 = oracle.join(queries, data, r.final_domains, r.gmcr_offsets, r.gmcr_indices,
                                find_first, flat, sym=(du, vm))
        idx = r.gmcr_indices.astype(np.int64)
        mult = np.array([bin(int(v)).count("1") + 1 for v in vm], np.uint64)[idx]
        if find_first:  # existence is unchanged (the first match may come later or earlier)
            assert np.array_equal(c1, c0)
        else:  # every DFS node of the constrained search is one of the plain search's
            assert np.array_equal(c1 * mult, c0)
            assert np.all(n1 <= n0) and n1.sum() < n0.sum()


def test_regular_graphs_are_told_apart_quickly():
    """Colour refinement cannot tell regular graphs of one label apart; the BFS label profile in
    the initial colours and the search budgets keep such batches fast, and the classes still
    equal networkx's (renumbered copies of some graphs join their originals)."""
    import time
    rng = np.random.default_rng(11)
    gs = []
    for i in range(60):
        G = nx.random_regular_graph(3, 12, seed=i)
        gs.append(([8] * 12, [tuple(e) for e in G.edges()]))
    for i in range(0, 60, 6):
        labels, edges = gs[i]
        perm = rng.permutation(12)
        gs.append(([8] * 12, [(int(perm[u]), int(perm[v])) for u, v in edges]))
    for i in range(100):
        G = nx.random_regular_graph(4, 32, seed=1000 + i)
        gs.append(([8] * 32, [tuple(e) for e in G.edges()]))
    batch = GraphBatch.from_graphs(gs)
    t0 = time.perf_counter()
    rep = api.query_classes(batch)
    api.query_orbits(batch)
    assert time.perf_counter() - t0 < 2.0
    assert np.array_equal(rep[:70], _networkx_classes(GraphBatch.from_graphs(gs[:70])))
    assert all(rep[60 + k] <= 6 * k for k in range(10))
