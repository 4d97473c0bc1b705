"""Resident data batch with rebound query batches (bms2_pipeline_set_queries), per-query totals,
and the serving executor, against the CPU oracle."""
import numpy as np
import pytest
from conftest import expected_stats

import oracle
from bms2_amd import api, graphs, serve

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def queue():
    q = api.Queue(0)
    yield q
    q.close()


def _subset(queries, idx):
    return graphs.GraphBatch.from_graphs([queries.graph(int(i)) for i in idx])


@pytest.mark.parametrize("workload", ["molecules-1m", "uneven-1m"])
def test_set_queries_and_totals(queue, workload):
    data, queries = graphs.generate(workload, seed=0, num_data_graphs=2500)
    dd = api.DeviceBatchedCSRGraph(queue, data)
    b0 = _subset(queries, range(64))
    pipe = api.Pipeline(queue, api.DeviceBatchedCSRGraph(queue, b0), dd, refinement_steps=6)
    rng = np.random.default_rng(1)
    for size in (100, 7, 256):
        qb = _subset(queries, rng.integers(0, queries.num_graphs, size=size))
        pipe.set_queries(api.DeviceBatchedCSRGraph(queue, qb))
        counts, valid, stats = pipe.alloc_outputs()
        for _ in range(2):  # the second run replays the captured graph
            pipe.run(counts, valid, stats)
            queue.wait()
            c, v = pipe.results(counts, valid)
            rc, rv, rs = oracle.pipeline(qb, data, 6, False)
            assert np.array_equal(c, rc) and np.array_equal(v, rv)
            assert np.array_equal(stats.to_numpy(np.uint64, 8)[:4],
                                  expected_stats(qb, data, 6, False, rs))
        m, e = pipe.query_totals(counts)
        assert np.array_equal(m, (rc > 0).sum(axis=0).astype(np.uint64))
        assert np.array_equal(e, rc.sum(axis=0).astype(np.uint64))


def test_serving_executor_answers(queue):
    data, pool = graphs.generate("molecules-2m", seed=0, num_data_graphs=2000, num_query_graphs=300)
    dd = api.DeviceBatchedCSRGraph(queue, data)
    pipe = api.Pipeline(queue, api.DeviceBatchedCSRGraph(queue, _subset(pool, range(16))), dd,
                        refinement_steps=6)
    reqs = serve.make_requests(12, 500.0, pool.num_graphs, seed=4, max_queries=40)
    ex = serve.PipelineExecutor(queue, pipe, pool, data.num_graphs, max_queries=512)
    out, loop = serve.serve(reqs, ex, data.num_graphs, serve.batch_bytes(data.num_graphs, 80))
    rc, _, _ = oracle.pipeline(pool, data, 6, False)
    for r in out:
        assert r.done >= r.arrival
        assert np.array_equal(r.matched, (rc[:, r.queries] > 0).sum(axis=0).astype(np.uint64))
        assert np.array_equal(r.embeddings, rc[:, r.queries].sum(axis=0).astype(np.uint64))
    assert len(loop["batches"]) >= 2


def test_serving_executor_learned_orders_keep_answers(queue):
    from bms2_amd import learned
    data, pool = graphs.generate("molecules-2m", seed=0, num_data_graphs=2000, num_query_graphs=120)
    dd = api.DeviceBatchedCSRGraph(queue, data)
    pipe = api.Pipeline(queue, api.DeviceBatchedCSRGraph(queue, _subset(pool, range(16))), dd,
                        refinement_steps=6)
    net = learned.OrderNet.load(queue)
    ex = serve.PipelineExecutor(queue, pipe, pool, data.num_graphs, max_queries=512, net=net,
                                plan_graphs=1000)
    reqs = serve.make_requests(10, 500.0, pool.num_graphs, seed=5, max_queries=60)
    out, loop = serve.serve(reqs, ex, data.num_graphs, serve.batch_bytes(data.num_graphs, 100))
    rc, _, _ = oracle.pipeline(pool, data, 6, False)
    for r in out:
        assert np.array_equal(r.embeddings, rc[:, r.queries].sum(axis=0).astype(np.uint64))
    assert ex.planned_batches >= 1 and len(ex.orders) > 0
