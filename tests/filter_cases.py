"""Hand-derived iteration-0 domains of the filter (isomorphism.hpp:71-97).

The filter keeps data node d for query node q when four terms hold: the labels agree (or q's
label is the wildcard 255), deg(q) <= deg(d), path(d) contains path(q), and every ring count of d
is >= q's.  Each case below is built so that exactly one term removes a candidate that the other
three keep; the expected domains are worked out by hand from the graphs (comments give the
reasoning), not computed by any code under test.  Dropping or weakening any single term in an
implementation of the filter changes at least one expected domain.  Counts alone cannot show
this: every term is a sound prune, so a looser domain never adds an induced embedding.

Path signatures: a node u's Bloom filter holds the label sequences of the simple walks u-v-w
(w != u) and u-v-w-x (x not in {u, v}) starting at u (signature.hpp:472-519).  A node whose
neighbours have no other neighbours therefore has an empty path signature.
"""
from __future__ import annotations

C, N, O = 8, 9, 10  # carbon, nitrogen, oxygen (the label codes of generate_data.py)
WILDCARD = 255


def _ring(n, label=C):
    return ([label] * n, [(i, (i + 1) % n) for i in range(n)])


# (name, query graph, data graph, expected domain per query node as a set of local data nodes,
#  the term that does the excluding)
CASES = [
    (
        # Query: path C0-C1-C2 (centre 1 has degree 2 and an empty path signature: its
        # neighbours are leaves).  Data: star, centre 0 joined to leaves 1, 2, 3.
        # Query centre vs data leaves: label C, path empty, no rings -> only deg 2 > 1 removes
        # them.  Query ends (degree 1, path {C-C-C}) vs data centre: degree 3 >= 1 and the data
        # centre's path signature is empty (its neighbours are leaves), so path removes it.
        "degree",
        ([C, C, C], [(0, 1), (1, 2)]),
        ([C, C, C, C], [(0, 1), (0, 2), (0, 3)]),
        [{1, 2, 3}, {0}, {1, 2, 3}],
    ),
    (
        # Query: all-carbon 5-ring.  Data: all-carbon 6-ring.  Every node: label C, degree 2, path
        # signature {C-C-C, C-C-C-C} on both sides.  Query nodes lie on one same-label 5-ring,
        # data nodes on none (they lie on one 6-ring): the ring term alone empties every domain.
        "ring5",
        _ring(5),
        _ring(6),
        [set()] * 5,
    ),
    (
        # The converse: a 6-ring query against a 5-ring data graph; only the 6-ring field of the
        # ring signature differs, so comparing only the 5-ring fields would keep every node.
        "ring6",
        _ring(6),
        _ring(5),
        [set()] * 6,
    ),
    (
        # Query: path C0-C1-O2.  Data: path C0-C1-N2.  Query end C0 (degree 1, path {C-C-O})
        # vs data end C0 (degree 1, path {C-C-N}) and data middle C1 (degree 2, empty path:
        # its neighbours are leaves): labels and degrees pass, the path term removes both.
        # Query middle C1 (degree 2, empty path) keeps data C1 and loses data C0 by degree.
        # Query O2 matches no label.
        "path",
        ([C, C, O], [(0, 1), (1, 2)]),
        ([C, C, N], [(0, 1), (1, 2)]),
        [set(), {1}, set()],
    ),
    (
        # Query: edge between a wildcard node and C.  Data: edge O0-C1.  Both sides have
        # degree 1 and empty path signatures.  The wildcard keeps both data nodes; the query C
        # keeps only data C1 (the label term removes O0).
        "wildcard-label",
        ([WILDCARD, C], [(0, 1)]),
        ([O, C], [(0, 1)]),
        [{0, 1}, {1}],
    ),
]


def batches(case):
    """(query batch, data batch) of one case."""
    from bms2_amd.graphs import GraphBatch
    _, qg, dg, _ = case
    return GraphBatch.from_graphs([qg]), GraphBatch.from_graphs([dg])


def domain_sets(cand, num_query_nodes, num_data_nodes):
    """Rows of a [query node][word] u64 bit matrix as sets of data node ids."""
    out = []
    for qn in range(num_query_nodes):
        out.append({d for d in range(num_data_nodes)
                    if (int(cand[qn, d // 64]) >> (d % 64)) & 1})
    return out
