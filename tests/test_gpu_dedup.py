"""One query per isomorphism class: bms2_pipeline_run screens and joins each class's first query
and copies its count and GMCR bit to the others.  Every output must equal the oracle's, which
runs every query on its own, and the run with dedup off."""
import numpy as np
import pytest
from conftest import expected_stats

import oracle
from bms2_amd import api, graphs
from bms2_amd.graphs import GraphBatch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def queue():
    q = api.Queue(0)
    yield q
    q.close()


def _permuted_copies(queries: GraphBatch, step: int, seed: int) -> GraphBatch:
    """The batch followed by renumbered copies of every step-th query (edges shuffled)."""
    rng = np.random.default_rng(seed)
    out = [queries.graph(g) for g in range(queries.num_graphs)]
    for g in range(0, queries.num_graphs, step):
        labels, edges = queries.graph(g)
        perm = rng.permutation(len(labels))
        nl = np.empty(len(labels), np.uint8)
        nl[perm] = labels
        ne = [(int(perm[u]), int(perm[v])) for u, v in edges]
        rng.shuffle(ne)
        out.append((nl, ne))
    return GraphBatch.from_graphs(out)


def _run(queue, queries, data, K, find_first, dedup):
    pipe = api.Pipeline(queue, api.DeviceBatchedCSRGraph(queue, queries),
                        api.DeviceBatchedCSRGraph(queue, data), refinement_steps=K,
                        find_first=find_first)
    ndup = pipe.set_dedup(dedup)
    counts, valid, stats = pipe.alloc_outputs()
    pipe.run(counts, valid, stats)
    queue.wait()
    c, v = pipe.results(counts, valid)
    return c, v, stats.to_numpy(np.uint64, 8), ndup


@pytest.mark.parametrize("find_first", [False, True])
def test_dedup_equals_oracle_and_full_run(queue, molecules_1m_shard, find_first):
    data, queries = molecules_1m_shard
    c1, v1, s1, ndup = _run(queue, queries, data, 6, find_first, True)
    c0, v0, s0, _ = _run(queue, queries, data, 6, find_first, False)
    rc, rv, rs = oracle.pipeline(queries, data, 6, find_first)
    assert ndup > 100  # the 1,024-query batch repeats many small queries
    assert np.array_equal(c1, rc) and np.array_equal(v1, rv)
    assert np.array_equal(c0, rc) and np.array_equal(v0, rv)
    assert np.array_equal(s0[:4], rs)
    assert np.array_equal(s1[:4], expected_stats(queries, data, 6, find_first, rs))
    assert s1[3] < s0[3]


@pytest.mark.parametrize("K", [2, 6])
def test_dedup_renumbered_copies_uneven(queue, K):
    """Renumbered copies of the uneven workload's queries, against graphs of up to 250 nodes
    (the big-graph screen and joins copy classes too), both modes."""
    data, queries = graphs.generate("uneven-1m", seed=0, num_data_graphs=1500)
    batch = _permuted_copies(queries, 5, seed=K)
    rep = api.query_classes(batch)
    assert np.all(rep[queries.num_graphs:] < queries.num_graphs)
    for find_first in (False, True):
        c, v, s, ndup = _run(queue, batch, data, K, find_first, True)
        rc, rv, rs = oracle.pipeline(batch, data, K, find_first)
        assert ndup >= (queries.num_graphs + 4) // 5
        assert np.array_equal(c, rc) and np.array_equal(v, rv)
        assert np.array_equal(s[:4], expected_stats(batch, data, K, find_first, rs))


def test_dedup_switch_on_one_pipeline(queue, molecules_1m_shard):
    """Turning the class copies and the symmetry breaking off and on again on one pipeline (whose
    runs replay captured graphs) gives the oracle's results every time and the matching DFS-node
    statistics: every query searched in full when off, the constrained searches of the classes'
    first queries when on."""
    data, queries = molecules_1m_shard
    rc, rv, rs = oracle.pipeline(queries, data, 6, False)
    on = expected_stats(queries, data, 6, False, rs)
    assert on[3] < rs[3]
    pipe = api.Pipeline(queue, api.DeviceBatchedCSRGraph(queue, queries),
                        api.DeviceBatchedCSRGraph(queue, data), refinement_steps=6)
    counts, valid, stats = pipe.alloc_outputs()
    for dedup in (True, False, True, False):
        pipe.set_dedup(dedup)
        for _ in range(2):  # the second run replays the captured chunk
            pipe.run(counts, valid, stats)
            queue.wait()
            c, v = pipe.results(counts, valid)
            assert np.array_equal(c, rc) and np.array_equal(v, rv)
            assert np.array_equal(stats.to_numpy(np.uint64, 8)[:4], on if dedup else rs), dedup
