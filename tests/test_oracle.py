"""Pins the CPU oracle: known answers, brute-force fixtures (networkx), closed forms, invariants.

Each check is chosen so that a plausible restatement mistake fails it: non-induced matching
(fixtures, star in star), shells as balls instead of exact-distance rings (P4 in C5 at K=3),
a dropped wildcard, a transposed hash combine (Bloom bit positions), visited-set leaks
(order invariance, star in star).  The filter's four terms (label or wildcard, degree, path
Bloom containment, ring counts) are pinned at the domain level by hand-derived iteration-0 domains
(tests/filter_cases.py), one case per term: every term is a sound prune, so the count fixtures
cannot see a term that is missing.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import filter_cases
import oracle
from bms2_amd.graphs import GraphBatch

GOLDEN = Path(__file__).resolve().parent / "golden" / "induced_counts.json"
C = 8  # carbon (generate_data.py label map)


def batch(graphs):
    return GraphBatch.from_graphs(graphs)


def ring(n, label=C):
    return ([label] * n, [(i, (i + 1) % n) for i in range(n)])


def path(n, label=C):
    return ([label] * n, [(i, i + 1) for i in range(n - 1)])


def counts_dense(q, d, K, find_first=False):
    r = oracle.staged(q, d, K, find_first=find_first, keep_domains=False)
    dense = np.zeros((d.num_graphs, q.num_graphs), np.uint64)
    for g in range(d.num_graphs):
        s, e = int(r.gmcr_offsets[g]), int(r.gmcr_offsets[g + 1])
        dense[g, r.gmcr_indices[s:e]] = r.counts[s:e]
    return dense


# ---------------------------------------------------------------- hash known answers

def test_fmix32_published_value():
    # MurmurHash3 fmix32 finaliser: fmix32(1) = 0x514e28b7
    assert oracle.hash32(1) == 0x514E28B7
    assert oracle.hash32(0) == 0


def test_combine_hash_closed_form():
    # boost::hash_combine(h1, h2) = h1 ^ (h2 + 0x9e3779b9 + (h1 << 6) + (h1 >> 2))
    assert oracle.combine_hash(0, 0) == 0x9E3779B9
    assert oracle.combine_hash(1, 0) == 0x9E3779F8
    assert oracle.combine_hash(0, 1) == 0x9E3779BA


# ---------------------------------------------------------------- benzene

def test_benzene_signatures():
    g = batch([ring(6)])
    lab = oracle.label_signatures(g)
    assert np.all(lab == np.uint64(2 << (4 * C)))
    # Bloom bits: C-C-C walks set {54, 66, 125}, C-C-C-C walks set {15, 16, 147}
    bits = oracle.path_signatures(g)
    expect = set([54, 66, 125, 15, 16, 147])
    for row in bits:
        got = {w * 64 + b for w in range(4) for b in range(64) if (int(row[w]) >> b) & 1}
        assert got == expect
    cyc = oracle.cycle_signatures(g)
    assert np.all(cyc == np.uint64(1 << 32))  # one same-label 6-ring per atom
    st = oracle.RefineState(g)
    shells = [st.refine(k).copy() for k in range(1, 5)]
    sizes = [2, 2, 1, 0]
    for k, n in enumerate(sizes):
        assert np.all(shells[k] == np.uint64(n << (4 * C))), k


def test_benzene_in_benzene_is_automorphism_count():
    g = batch([ring(6)])
    for K in (0, 2, 6):
        assert counts_dense(g, g, K)[0, 0] == 12


# ---------------------------------------------------------------- filter domains

@pytest.mark.parametrize("case", filter_cases.CASES, ids=[c[0] for c in filter_cases.CASES])
def test_filter_iteration0_domains_hand_derived(case):
    q, d = filter_cases.batches(case)
    r = oracle.staged(q, d, 0, keep_domains=True)
    got = filter_cases.domain_sets(r.domains[0], q.total_nodes, d.total_nodes)
    assert got == case[3]


# ---------------------------------------------------------------- closed forms

def test_star_in_star_counts_ordered_leaf_pairs():
    star3 = ([C] * 4, [(0, 1), (0, 2), (0, 3)])
    star2 = ([C] * 3, [(0, 1), (0, 2)])
    # centre -> centre, two distinct leaves in order: 3 * 2.  A visited set that leaks on
    # backtracking (isomorphism.hpp:418-447 as written) finds only 2.
    for K in (0, 1, 2):
        assert counts_dense(batch([star2]), batch([star3]), K)[0, 0] == 6


def test_p4_in_c5_exact_distance_shells():
    q, d = batch([path(4)]), batch([ring(5)])
    for K in (0, 1, 2):
        assert counts_dense(q, d, K)[0, 0] == 10
    # the end node of P4 has a node at distance 3; in C5 nothing is farther than 2
    for K in (3, 6):
        assert counts_dense(q, d, K)[0, 0] == 0


@pytest.mark.parametrize("k,n", [(2, 6), (3, 6), (5, 6), (3, 7), (4, 9)])
def test_induced_paths_in_cycles(k, n):
    assert counts_dense(batch([path(k)]), batch([ring(n)]), 2)[0, 0] == 2 * n


def test_induced_not_monomorphism():
    # a triangle contains P3 as a subgraph but not as an induced subgraph
    tri = ([C] * 3, [(0, 1), (1, 2), (0, 2)])
    assert counts_dense(batch([path(3)]), batch([tri]), 0)[0, 0] == 0


def test_wildcard_query_label():
    data = batch([([7, C], [(0, 1)])])
    q = batch([([255, C], [(0, 1)])])
    assert counts_dense(q, data, 0)[0, 0] == 1
    q_plain = batch([([7, C], [(0, 1)])])
    assert counts_dense(q_plain, data, 0)[0, 0] == 1


def test_single_node_query_never_valid():
    q = batch([([C], [])])
    d = batch([ring(6)])
    r = oracle.staged(q, d, 2)
    assert len(r.gmcr_indices) == 0


def test_first_match_is_count_positive():
    q = batch([path(3), ring(5), path(2), ([7, C], [(0, 1)])])
    d = batch([ring(6), ring(5), path(7), ([C, 7, C], [(0, 1), (1, 2)])])
    ca = counts_dense(q, d, 2)
    ff = counts_dense(q, d, 2, find_first=True)
    assert np.array_equal(ff, (ca > 0).astype(np.uint64))


# ---------------------------------------------------------------- brute force fixtures

def load_cases():
    meta = json.loads(GOLDEN.read_text())
    for c in meta["cases"]:
        data = batch([(g["labels"], [tuple(e) for e in g["edges"]]) for g in c["data"]])
        qs = batch([(g["labels"], [tuple(e) for e in g["edges"]]) for g in c["queries"]])
        yield c["name"], data, qs, np.array(c["counts"], dtype=np.uint64)


@pytest.mark.parametrize("K", [0, 1, 2])
def test_counts_equal_networkx_for_sound_depths(K):
    for name, data, qs, expect in load_cases():
        got = counts_dense(qs, data, K).T  # [query][data]
        nq = np.diff(qs.graph_offsets.astype(np.int64))
        ok = nq > 1  # single-node queries are never in the GMCR
        assert np.array_equal(got[ok], expect[ok]), name
        assert np.all(got[~ok] == 0), name


def test_counts_bounded_by_networkx_for_deep_refinement():
    for name, data, qs, expect in load_cases():
        got = counts_dense(qs, data, 6).T
        assert np.all(got <= expect), name


def test_fused_oracle_matches_fixtures():
    for name, data, qs, expect in load_cases():
        counts, valid, stats = oracle.pipeline(qs, data, 2)
        nq = np.diff(qs.graph_offsets.astype(np.int64))
        ok = nq > 1
        assert np.array_equal(counts.T[ok], expect[ok]), name


# ---------------------------------------------------------------- invariants

def test_domains_shrink_monotonically(molecules_2k):
    data, queries = molecules_2k
    r = oracle.staged(queries, data, 6)
    for a, b in zip(r.domains, r.domains[1:]):
        assert np.all((b & ~a) == 0)


def test_gmcr_is_nonempty_domain_set(molecules_2k):
    data, queries = molecules_2k
    r = oracle.staged(queries, data, 2)
    dom = r.final_domains
    bits = np.unpackbits(dom.view(np.uint8), axis=1, bitorder="little")[:, :data.total_nodes]
    go = data.graph_offsets.astype(np.int64)
    nonempty = np.add.reduceat(bits, go[:-1], axis=1) > 0
    qgo = queries.graph_offsets.astype(np.int64)
    expect = []
    for g in range(data.num_graphs):
        for q in range(queries.num_graphs):
            nodes = range(qgo[q], qgo[q + 1])
            if len(nodes) > 1 and all(nonempty[i, g] for i in nodes):
                expect.append((g, q))
    got = [(g, int(q)) for g in range(data.num_graphs)
           for q in r.gmcr_indices[r.gmcr_offsets[g]:r.gmcr_offsets[g + 1]]]
    assert got == expect


def test_count_invariant_under_matching_order(molecules_2k):
    data, queries = molecules_2k
    r = oracle.staged(queries, data, 2)
    rng = np.random.default_rng(7)
    order = np.zeros(queries.total_nodes, np.uint32)
    qgo = queries.graph_offsets
    for q in range(queries.num_graphs):
        n = int(qgo[q + 1] - qgo[q])
        order[qgo[q]:qgo[q + 1]] = rng.permutation(n)
    total, counts, _ = oracle.join(queries, data, r.final_domains, r.gmcr_offsets,
                                   r.gmcr_indices, False, order)
    assert np.array_equal(counts, r.counts)
    assert total == r.total


@pytest.mark.parametrize("K", [0, 2, 6])
def test_fused_oracle_equals_staged(molecules_2k, K):
    data, queries = molecules_2k
    sub = data.subset(0, 500)
    dense = counts_dense(queries, sub, K)
    counts, valid, stats = oracle.pipeline(queries, sub, K)
    assert np.array_equal(counts, dense)
    r = oracle.staged(queries, sub, K, keep_domains=False)
    assert int(stats[0]) == len(r.gmcr_indices)
    assert int(stats[3]) == int(r.dfs_nodes.sum())


def exact_ring_counts(labels, edges):
    """Per node (5 same, 5 diff, 6 same, 6 diff) simple-cycle counts from networkx."""
    nx = pytest.importorskip("networkx")
    g = nx.Graph()
    g.add_nodes_from(range(len(labels)))
    g.add_edges_from(edges)
    out = np.zeros((len(labels), 4), np.int64)
    for cyc in nx.simple_cycles(g, length_bound=6):
        if len(cyc) not in (5, 6):
            continue
        for s in cyc:
            same = all(labels[v] == labels[s] for v in cyc)
            out[s, (0 if len(cyc) == 5 else 2) + (0 if same else 1)] += 1
    return out


def unpack_cycles(v):
    v = int(v)
    return [(v >> (16 * f)) & 0xFFFF for f in range(4)]


def test_ring_counts_equal_simple_cycles():
    rng = np.random.default_rng(3)
    for _, data, qs, _ in load_cases():
        for b in (data, qs):
            cyc = oracle.cycle_signatures(b)
            for gi in range(b.num_graphs):
                labels, edges = b.graph(gi)
                exact = exact_ring_counts(list(labels), edges)
                n0 = int(b.graph_offsets[gi])
                for i in range(len(labels)):
                    assert unpack_cycles(cyc[n0 + i]) == list(exact[i]), (gi, i)
    del rng


def hub_bipartite(na, nb):
    """Node 0 joined to every node of A; A and B completely joined (all carbon)."""
    A = list(range(1, 1 + na))
    B = list(range(1 + na, 1 + na + nb))
    edges = [(0, a) for a in A] + [(a, b) for a in A for b in B]
    return [C] * (1 + na + nb), edges


@pytest.mark.parametrize("na,nb,overflow", [(10, 10, False), (30, 30, True)])
def test_ring_count_stack_cap_literal(na, nb, overflow):
    # Rings through node 0 are 0-a-b-a'-b'-a''-0 only (no 5-rings): na*nb*(na-1)*(nb-1)*(na-2)
    # directed closed paths, kept in a u16 counter and halved (signature.hpp:618-714).  With
    # na = nb = 10 the 128-entry stack never fills and the closed form holds exactly (64,800 is
    # just under 2^16); with 30 + 30 the stack exceeds 128 entries at depth 5
    # (signature.hpp:638,697), pushes are dropped and the literal count departs from it.
    labels, edges = hub_bipartite(na, nb)
    cyc = unpack_cycles(oracle.cycle_signatures(batch([(labels, edges)]))[0])
    directed = na * nb * (na - 1) * (nb - 1) * (na - 2)
    closed_form = (directed % 65536) // 2
    assert cyc[0] == 0 and cyc[1] == 0 and cyc[3] == 0
    if overflow:
        assert cyc[2] != closed_form
    else:
        assert cyc[2] == closed_form == 32400
