"""CS (cluster) join on the GPU (joinPartialCandidates2 + joinWithPartialMatches,
isomorphism.hpp:756-977, 1021-1178; BMS2 --use-cs) against the oracle: partial-match records
bit-exact and in the same order, per-GMCR-entry counts and totals exact, in both modes."""
import numpy as np
import pytest

import oracle
from bms2_amd import api, graphs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def queue():
    q = api.Queue(0)
    yield q
    q.close()


def _check(queue, data, queries, K, find_first):
    print(f"[cs] D={data.num_graphs} Q={queries.num_graphs} K={K} find_first={find_first}",
          flush=True)
    res = api.run_staged(queue, queries, data, K, find_first=find_first, keep_domains=False,
                         use_cs=True)
    ref = oracle.staged(queries, data, K, find_first=find_first, keep_domains=False)
    assert np.array_equal(res["final_domains"], ref.final_domains)
    off, idx = res["gmcr"]
    assert np.array_equal(off, ref.gmcr_offsets) and np.array_equal(idx, ref.gmcr_indices)
    end = int(queries.graph_offsets[1] - queries.graph_offsets[0])
    part = oracle.join_partial(queries, data, ref.final_domains, end)
    assert np.array_equal(res["partial"], part)
    total, counts = oracle.join_with_partial(queries, data, ref.final_domains, off, idx, end, part,
                                             find_first)
    assert np.array_equal(res["counts"], counts)
    assert res["total"] == total
    return res, ref


@pytest.mark.parametrize("K", [0, 2, 6])
@pytest.mark.parametrize("find_first", [False, True])
def test_cs_join_matches_oracle(queue, molecules_2k, K, find_first):
    data, _ = molecules_2k
    queries = graphs.cluster_queries(data, 63, seed=20 + K)
    res, ref = _check(queue, data, queries, K, find_first)
    assert len(res["partial"]) > 0 and res["total"] > 0
    if K <= 2:  # sound filters and the cluster convention: the regular join's counts
        assert np.array_equal(res["counts"], ref.counts)


def test_cs_join_large_graphs(queue):
    # graphs above 64 nodes take the four-word bitsets
    data, _ = graphs.generate("uneven-1m", seed=5, num_data_graphs=300, num_query_graphs=8)
    assert int(np.diff(data.graph_offsets).max()) > 64
    queries = graphs.cluster_queries(data, 20, seed=6, member_nodes=(6, 20))
    for ff in (False, True):
        res, ref = _check(queue, data, queries, 2, ff)
        assert np.array_equal(res["counts"], ref.counts)


def _candidates(queue, dq, dd, K):
    """Domains through the stage calls (filter, K refinement steps) of BMS2.cpp:131-199."""
    sig = api.Signature(queue, dd.total_nodes, dq.total_nodes)
    path = api.PathSignature(queue, dd.total_nodes, dq.total_nodes)
    cyc = api.CycleSignature(queue, dd.total_nodes, dq.total_nodes)
    cand = api.Candidates(queue, dq.total_nodes, dd.total_nodes)
    for g, fn in ((dd, "data"), (dq, "query")):
        getattr(sig, f"generate_{fn}_signatures")(g)
        getattr(path, f"generate_{fn}_path_signatures")(g)
        getattr(cyc, f"generate_{fn}_cycle_signatures")(g)
    api.filter_candidates(queue, dq, dd, sig, path, cyc, cand)
    for k in range(1, K + 1):
        sig.refine_data_signatures(dd, k)
        sig.refine_query_signatures(dq, k)
        api.refine_candidates(queue, dq, dd, sig, cand)
    queue.wait()
    return cand


def test_cs_partial_capacity_and_depth(queue, molecules_2k):
    data, _ = molecules_2k
    queries = graphs.cluster_queries(data, 7, seed=3)
    dq = api.DeviceBatchedCSRGraph(queue, queries)
    dd = api.DeviceBatchedCSRGraph(queue, data)
    ref = oracle.staged(queries, data, 2, keep_domains=False)
    cand = _candidates(queue, dq, dd, 2)
    assert np.array_equal(cand.host(), ref.final_domains)
    end = int(queries.graph_offsets[1] - queries.graph_offsets[0])
    buf, n = api.join_partial_candidates(queue, dq, dd, cand, end)
    assert n == len(oracle.join_partial(queries, data, ref.final_domains, end)) > 1
    with pytest.raises(api.Bms2Error) as e:
        api.join_partial_candidates(queue, dq, dd, cand, end, capacity=n - 1)
    assert e.value.code == 7
    # a depth beyond query 0's nodes yields nothing (isomorphism.hpp:1087)
    _, n2 = api.join_partial_candidates(queue, dq, dd, cand, end + 1)
    assert n2 == 0
    # a shallower depth: the oracle's prefix matches
    _, n3 = api.join_partial_candidates(queue, dq, dd, cand, end - 1)
    assert n3 == len(oracle.join_partial(queries, data, ref.final_domains, end - 1))
