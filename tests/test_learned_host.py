"""Host side of the learned matching order (no GPU): query structure, features, orders."""
import numpy as np

from bms2_amd import learned
from bms2_amd.graphs import GraphBatch


def _queries():
    path4 = ([8, 8, 7, 0], [(0, 1), (1, 2), (2, 3)])
    ring6 = ([8] * 6, [(i, (i + 1) % 6) for i in range(6)])
    # triangle 0-1-2 with a tail 2-3-4
    tri_tail = ([8, 8, 8, 7, 255], [(0, 1), (1, 2), (0, 2), (2, 3), (3, 4)])
    return GraphBatch.from_graphs([path4, ring6, tri_tail])


def test_query_infos_structure():
    infos = learned.query_infos(_queries())
    p, r, t = infos
    assert p.n == 4 and list(p.deg) == [1, 2, 2, 1]
    assert p.dist[0, 3] == 3 and not p.on_cycle.any()
    assert r.on_cycle.all() and r.dist[0, 3] == 3 and r.dist.max() == 3
    assert list(t.on_cycle) == [True, True, True, False, False]
    assert t.dist[0, 4] == 3


def _is_connected_order(qi, order):
    seen = {order[0]}
    for u in order[1:]:
        if not (qi.adj[u] & seen):
            return False
        seen.add(u)
    return True


def test_orders_are_connected_permutations():
    infos = learned.query_infos(_queries())
    rng = np.random.default_rng(0)
    for qi in infos:
        dom = rng.uniform(1, 10, size=qi.n)
        for root in range(qi.n):
            o = learned.greedy_order(qi, root, dom)
            assert sorted(o) == list(range(qi.n)) and o[0] == root
            assert _is_connected_order(qi, o)
        s = rng.normal(size=qi.n)
        o = learned.score_order(qi, s)
        assert sorted(o) == list(range(qi.n)) and o[0] == int(np.argmax(s))
        assert _is_connected_order(qi, o)


def test_score_order_follows_scores_along_frontier():
    qi = learned.query_infos(_queries())[0]  # path 0-1-2-3
    # highest score at node 3, then node 0: node 0 is not adjacent, so 2 and 1 come first
    assert learned.score_order(qi, np.array([0.9, 0.1, 0.2, 1.0])) == [3, 2, 1, 0]


def test_candidate_orders_and_features():
    qs = _queries()
    infos = learned.query_infos(qs)
    domsum = np.arange(qs.total_nodes, dtype=np.uint64) * 10
    qpairs = np.array([10, 0, 5], np.uint64)
    cands = learned.candidate_orders(infos, domsum, qpairs)
    names = [c[0] for c in cands]
    assert names[:2] == ["identity", "ref_maxdom"] and len(cands) == 2 + 6
    for _, o in cands:
        for q, qi in enumerate(infos):
            assert sorted(o[q, :qi.n].tolist()) == list(range(qi.n))
    # reference heuristic: the max-domain node first, the rest in identity order
    assert list(cands[1][1][0, :4]) == [3, 0, 1, 2]
    X = learned.features(infos, domsum, qpairs)
    assert X.shape == (qs.total_nodes, learned.NUM_FEATURES) and X.dtype == np.float32
    # path4 node 1: degree 2, eccentricity 2, mean domain 10 / 10 pairs = 1
    assert X[1, 8] == 0.5 and X[1, 9] == 0.25 and abs(X[1, 11] - 1 / 6) < 1e-7
    # ring6 has no GMCR pairs in the sample: domain features are zero
    assert np.all(X[4:10, 11] == 0) and np.all(X[4:10, 15] == 1)
    # wildcard and element one-hot columns
    assert X[14, 16] == 1 and X[0, 17 + learned._ELEMENTS.index(8)] == 1
