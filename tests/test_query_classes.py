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
