"""bms2_join_candidates_noalloc (the join without any allocation, its scratch in a static device
array) against bms2_join_candidates, through the C ABI: per-pair counts and totals bit-equal on
two calls in a row, the second on a smaller batch, in both modes; a batch whose scratch would
exceed the array is refused before any device work."""
import ctypes as C

import numpy as np
import pytest

from bms2_amd import _lib, api, graphs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def queue():
    q = api.Queue(0)
    yield q
    q.close()


def _to_gmcr(q, queries, data, K):
    """Device batches, candidates and GMCR of the staged sequence with K refinement steps."""
    dq = api.DeviceBatchedCSRGraph(q, queries)
    dd = api.DeviceBatchedCSRGraph(q, data)
    sig = api.Signature(q, dd.total_nodes, dq.total_nodes)
    path = api.PathSignature(q, dd.total_nodes, dq.total_nodes)
    cyc = api.CycleSignature(q, dd.total_nodes, dq.total_nodes)
    cand = api.Candidates(q, dq.total_nodes, dd.total_nodes)
    for g, fn in ((dd, "data"), (dq, "query")):
        getattr(sig, f"generate_{fn}_signatures")(g)
        getattr(path, f"generate_{fn}_path_signatures")(g)
        getattr(cyc, f"generate_{fn}_cycle_signatures")(g)
    api.filter_candidates(q, dq, dd, sig, path, cyc, cand)
    for k in range(1, K + 1):
        sig.refine_data_signatures(dd, k)
        sig.refine_query_signatures(dq, k)
        api.refine_candidates(q, dq, dd, sig, cand)
    gm = api.GMCR(q)
    gm.generate_gmcr(dq, dd, cand)
    q.wait()
    return dq, dd, cand, gm


def _noalloc(q, dq, dd, cand, gm, total, find_first, pairs):
    return _lib.lib().bms2_join_candidates_noalloc(
        q.h, C.byref(dq.c), C.byref(dd.c), cand.h, gm.h, C.c_void_p(total.ptr), int(find_first),
        C.c_void_p(pairs.ptr), pairs.nbytes // 8)


@pytest.mark.parametrize("find_first", [False, True])
def test_noalloc_join_equals_join(queue, molecules_2k, find_first):
    data, queries = molecules_2k
    for size in (2000, 700):  # the second call on a smaller batch reuses the array
        print(f"noalloc join: {size} graphs, find_first={find_first}", flush=True)
        dq, dd, cand, gm = _to_gmcr(queue, queries, data.subset(0, size), 6)
        T = gm.total
        assert T > 0
        t1 = api.DeviceBuffer(queue, 8, zero=True)
        p1 = api.DeviceBuffer(queue, 8 * T, zero=True)
        api.join_candidates(queue, dq, dd, cand, gm, t1, find_first, p1)
        t2 = api.DeviceBuffer(queue, 8, zero=True)
        p2 = api.DeviceBuffer(queue, 8 * T, zero=True)
        _lib.check(_noalloc(queue, dq, dd, cand, gm, t2, find_first, p2))
        queue.wait()
        a, b = p1.to_numpy(np.uint64, T), p2.to_numpy(np.uint64, T)
        assert np.array_equal(a, b)
        assert int(t1.to_numpy(np.uint64)[0]) == int(t2.to_numpy(np.uint64)[0]) == int(a.sum())


def test_noalloc_join_large_graphs_and_queries(queue):
    """Data graphs above 64 nodes take the size-classed join (multi-word rows and sorted entries in
    the array); queries of 17-32 nodes the 32-deep instances."""
    data, queries = graphs.generate("uneven-1m", seed=0, num_data_graphs=600)
    print("noalloc join: uneven 600 graphs", flush=True)
    dq, dd, cand, gm = _to_gmcr(queue, queries, data, 2)
    T = gm.total
    t1 = api.DeviceBuffer(queue, 8, zero=True)
    p1 = api.DeviceBuffer(queue, 8 * T, zero=True)
    api.join_candidates(queue, dq, dd, cand, gm, t1, False, p1)
    t2 = api.DeviceBuffer(queue, 8, zero=True)
    p2 = api.DeviceBuffer(queue, 8 * T, zero=True)
    _lib.check(_noalloc(queue, dq, dd, cand, gm, t2, False, p2))
    queue.wait()
    assert np.array_equal(p1.to_numpy(np.uint64, T), p2.to_numpy(np.uint64, T))
    assert int(t1.to_numpy(np.uint64)[0]) == int(t2.to_numpy(np.uint64)[0])


def test_noalloc_join_refuses_oversized_batch(queue):
    """8.4 M data nodes of 2-node graphs need more than the 64 MiB array: refused with
    BMS2_ERR_CAPACITY before any kernel, nothing written."""
    n_graphs = 4_200_000
    data = graphs.GraphBatch(
        graph_offsets=np.arange(0, 2 * n_graphs + 1, 2, dtype=np.uint32),
        row_offsets=np.arange(0, 2 * n_graphs + 1, dtype=np.uint32),
        column_indices=(np.arange(2 * n_graphs, dtype=np.uint32) ^ 1),
        node_labels=np.full(2 * n_graphs, 8, np.uint8),
        edge_labels=np.zeros(2 * n_graphs, np.uint8))
    queries = graphs.GraphBatch.from_graphs([([8, 8], [(0, 1)])])
    print("noalloc join: oversized batch", flush=True)
    dq, dd, cand, gm = _to_gmcr(queue, queries, data, 0)
    assert gm.total == n_graphs
    total = api.DeviceBuffer(queue, 8, zero=True)
    pairs = api.DeviceBuffer(queue, 8 * gm.total, zero=True)
    rc = _noalloc(queue, dq, dd, cand, gm, total, False, pairs)
    assert rc == _lib.ERR_CAPACITY
    assert "allocation-free join" in _lib.lib().bms2_last_error().decode()
    assert int(total.to_numpy(np.uint64)[0]) == 0
