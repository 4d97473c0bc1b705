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


def _features_by_node(infos, domsum, qpairs):
    """Straightforward per-node statement of the 32 features (the vectorised version must match)."""
    total = sum(qi.n for qi in infos)
    X = np.zeros((total, learned.NUM_FEATURES), np.float32)
    labels = np.concatenate([qi.labels for qi in infos])
    freq = np.bincount(labels, minlength=256) / max(len(labels), 1)
    r = 0
    for q, qi in enumerate(infos):
        n = qi.n
        pairs = float(qpairs[q])
        dom = domsum[r:r + n].astype(np.float64) / pairs if pairs > 0 else np.zeros(n)
        m = int(qi.deg.sum()) // 2
        reach = np.where(qi.dist < 99, qi.dist, 0)
        ecc = reach.max(axis=1)
        meandom = dom.mean()
        for u in range(n):
            f = X[r + u]
            nb = sorted(qi.adj[u])
            lab = int(qi.labels[u])
            f[:6] = [n / 16, m / 16, (m - n + 1) / 4, qi.deg.max() / 4, ecc.max() / 8,
                     len(set(qi.labels.tolist())) / 8]
            f[6] = np.log2(1 + pairs) / 16
            f[7] = np.log2(1 + meandom) / 6
            f[8], f[9] = qi.deg[u] / 4, ecc[u] / 8
            f[10] = reach[u].sum() / max(n - 1, 1) / 4
            f[11] = np.log2(1 + dom[u]) / 6
            f[12] = np.log2((1 + dom[u]) / (1 + meandom)) / 4
            f[13] = sum(qi.deg[v] == 1 for v in nb) / 4
            f[14] = sum(qi.deg[v] for v in nb) / 16
            f[15] = float(qi.on_cycle[u])
            f[16] = float(lab == 255)
            if lab in learned._ELEMENTS:
                f[17 + learned._ELEMENTS.index(lab)] = 1
            elif lab != 255:
                f[24] = 1
            f[25] = u / max(n - 1, 1)
            f[26] = np.log2(1 + min(dom[v] for v in nb)) / 6 if nb else 0
            f[27] = -np.log2(max(freq[lab], 1e-9)) / 8
            f[28] = sorted(range(n), key=lambda v: (-qi.deg[v], v)).index(u) / max(n - 1, 1)
            f[29] = sorted(range(n), key=lambda v: (dom[v], v)).index(u) / max(n - 1, 1)
            f[30] = float(u == int(np.argmax(qi.deg)))
            f[31] = float(u == sorted(range(n), key=lambda v: (dom[v], v))[0])
        r += n
    return X


def test_vectorised_features_match_per_node_statement():
    from bms2_amd import graphs
    _, queries = graphs.generate("uneven-1m", seed=2, num_data_graphs=50, num_query_graphs=60)
    infos = learned.query_infos(queries)
    rng = np.random.default_rng(0)
    domsum = rng.integers(0, 500, size=queries.total_nodes).astype(np.uint64)
    domsum[:20] = 7  # ties
    qpairs = rng.integers(0, 30, size=queries.num_graphs).astype(np.uint64)
    qpairs[::7] = 0
    got = learned.features(infos, domsum, qpairs)
    want = _features_by_node(infos, domsum, qpairs)
    assert np.allclose(got, want, rtol=1e-6, atol=1e-6), np.argwhere(~np.isclose(got, want))[:5]
    # cached static parts give the same result on a second call
    assert np.array_equal(learned.features(infos, domsum, qpairs), got)


def _batches_for_native():
    from bms2_amd import graphs
    out = [_queries()]
    for wl, seed in (("uneven-1m", 2), ("molecules-1m", 5)):
        _, q = graphs.generate(wl, seed=seed, num_data_graphs=50, num_query_graphs=200)
        out.append(q)
    # a repeated edge collapses as in the neighbour sets
    out.append(GraphBatch.from_graphs([([8, 8, 8], [(0, 1), (0, 1), (1, 2)])]))
    return out


def test_native_static_features_equal_python():
    """bms2_query_static_features gives _static_features' rows bit for bit (ring membership,
    distances, ranks and ties included) on hand-made, uneven and molecule queries and on a query
    with a repeated edge."""
    for q in _batches_for_native():
        infos = learned.query_infos(q)
        want = np.concatenate([learned._static_features(qi) for qi in infos])
        assert np.array_equal(learned.static_features(q), want)


def test_native_score_orders_equal_python():
    rng = np.random.default_rng(3)
    for q in _batches_for_native():
        infos = learned.query_infos(q)
        go = np.asarray(q.graph_offsets)
        for quant in (None, 0.5):  # continuous scores, then many ties
            sc = rng.normal(size=q.total_nodes).astype(np.float32)
            if quant:
                sc = (np.round(sc / quant) * quant).astype(np.float32)
            use = rng.random(q.num_graphs) < 0.7
            got = learned.score_orders(q, sc, use)
            for g, qi in enumerate(infos):
                want = learned.score_order(qi, sc[go[g]:go[g + 1]]) if use[g] else list(range(qi.n))
                assert list(got[g, :qi.n]) == want, (g, quant)


def test_batch_features_equal_features():
    from bms2_amd import graphs
    _, queries = graphs.generate("uneven-1m", seed=4, num_data_graphs=50, num_query_graphs=80)
    infos = learned.query_infos(queries)
    rng = np.random.default_rng(1)
    domsum = rng.integers(0, 500, size=queries.total_nodes).astype(np.uint64)
    qpairs = rng.integers(0, 30, size=queries.num_graphs).astype(np.uint64)
    qpairs[::5] = 0
    want = learned.features(infos, domsum, qpairs)
    assert np.array_equal(learned.batch_features(queries, domsum, qpairs), want)
    static = learned.static_features(queries)
    assert np.array_equal(learned.batch_features(queries, domsum, qpairs, static), want)
