"""HIP path vs the CPU oracle, through the C ABI (bit-exact: everything here is integer work)."""
import numpy as np
import pytest

import oracle
from bms2_amd import _lib, api, graphs
from bms2_amd.graphs import GraphBatch
from conftest import expected_stats

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def queue():
    q = api.Queue(0)
    yield q
    q.close()


def dense_from_gmcr(offsets, indices, counts, D, Q):
    dense = np.zeros((D, Q), np.uint64)
    valid = np.zeros((D, Q), bool)
    for g in range(D):
        s, e = int(offsets[g]), int(offsets[g + 1])
        dense[g, indices[s:e]] = counts[s:e]
        valid[g, indices[s:e]] = True
    return dense, valid


def unpack_valid(v, Q):
    bits = np.unpackbits(v.view(np.uint8), axis=1, bitorder="little")
    return bits[:, :Q].astype(bool)


def test_signatures_match_oracle(queue, molecules_2k):
    data, queries = molecules_2k
    out = api.run_staged(queue, queries, data, K=6, keep_domains=True)
    ref = oracle.staged(queries, data, 6, keep_domains=True)
    for key in ("dlabel", "qlabel", "dpath", "qpath", "dcycle", "qcycle"):
        assert np.array_equal(out["signatures"][key], ref.signatures[key]), key
    for k in range(1, 7):
        assert np.array_equal(out["signatures"][f"dsig{k}"], ref.signatures[f"dsig{k}"]), k
        assert np.array_equal(out["signatures"][f"qsig{k}"], ref.signatures[f"qsig{k}"]), k


@pytest.mark.parametrize("K", [0, 2, 6])
@pytest.mark.parametrize("find_first", [False, True])
def test_staged_matches_oracle_every_iteration(queue, molecules_2k, K, find_first):
    data, queries = molecules_2k
    out = api.run_staged(queue, queries, data, K=K, find_first=find_first)
    ref = oracle.staged(queries, data, K, find_first=find_first)
    assert len(out["domains"]) == K + 1
    for it, (a, b) in enumerate(zip(out["domains"], ref.domains)):
        assert np.array_equal(a, b), f"domains differ after iteration {it}"
    off, idx = out["gmcr"]
    assert np.array_equal(off, ref.gmcr_offsets)
    assert np.array_equal(idx, ref.gmcr_indices)
    assert np.array_equal(out["counts"], ref.counts)
    assert out["total"] == ref.total


@pytest.mark.parametrize("K", [0, 2, 6])
@pytest.mark.parametrize("find_first", [False, True])
def test_fused_pipeline_matches_oracle(queue, molecules_2k, K, find_first):
    data, queries = molecules_2k
    counts, valid, stats = api.run_pipeline(queue, queries, data, K, find_first)
    rc, rv, rs = oracle.pipeline(queries, data, K, find_first)
    assert np.array_equal(valid, rv)
    assert np.array_equal(counts, rc)
    assert int(stats[0]) == int(rs[0])  # GMCR pairs
    assert int(stats[1]) == int(rs[1])  # matched pairs
    assert int(stats[2]) == int(rs[2])  # embeddings
    if not find_first:  # DFS nodes (identity order) of the queries screened
        assert int(stats[3]) == int(expected_stats(queries, data, K, find_first, rs)[3])


def test_fused_equals_staged(queue, molecules_2k):
    data, queries = molecules_2k
    out = api.run_staged(queue, queries, data, K=6, keep_domains=False)
    counts, valid, _ = api.run_pipeline(queue, queries, data, 6, False)
    off, idx = out["gmcr"]
    dense, dvalid = dense_from_gmcr(off, idx, out["counts"], data.num_graphs, queries.num_graphs)
    assert np.array_equal(unpack_valid(valid, queries.num_graphs), dvalid)
    assert np.array_equal(counts, dense)


def test_fused_1m_shard(queue, molecules_1m_shard):
    data, queries = molecules_1m_shard
    counts, valid, stats = api.run_pipeline(queue, queries, data, 6, False)
    rc, rv, rs = oracle.pipeline(queries, data, 6, False)
    assert np.array_equal(valid, rv)
    assert np.array_equal(counts, rc)
    assert np.array_equal(stats[:4], expected_stats(queries, data, 6, False, rs))


@pytest.mark.parametrize("which", ["small", "large"])
@pytest.mark.parametrize("find_first", [False, True])
def test_fused_query_size_classes(queue, molecules_1m_shard, which, find_first):
    """Queries of up to 8 nodes go to the small-record join, larger ones to the large-record
    join; a batch of only one class exercises each join alone."""
    data, queries = molecules_1m_shard
    n = np.diff(np.asarray(queries.graph_offsets))
    sel = np.flatnonzero(n <= 8 if which == "small" else n > 8)[:200]
    qs = graphs.GraphBatch.from_graphs([queries.graph(int(i)) for i in sel])
    sub = data.subset(0, 1500)
    counts, valid, stats = api.run_pipeline(queue, qs, sub, 6, find_first)
    rc, rv, rs = oracle.pipeline(qs, sub, 6, find_first)
    assert np.array_equal(valid, rv)
    assert np.array_equal(counts, rc)
    assert int(stats[0]) == int(rs[0]) and int(stats[1]) == int(rs[1])
    if not find_first:
        assert np.array_equal(stats[:4], expected_stats(qs, sub, 6, False, rs))


@pytest.fixture(scope="module")
def uneven_2k():
    """2,000 graphs of the uneven generator (lognormal sizes up to 250 nodes, 10% fused ring
    systems; 55 graphs above 64 nodes) with its 1,024 queries of 3-32 nodes (seed 0)."""
    return graphs.generate("uneven-1m", seed=0, num_data_graphs=2000)


@pytest.mark.parametrize("K", [0, 2, 6])
@pytest.mark.parametrize("find_first", [False, True])
def test_fused_uneven_sizes(queue, uneven_2k, K, find_first):
    """Graphs above 64 nodes take the big-graph screen and size-class joins inside
    bms2_pipeline_run, queries above 16 nodes the 32-deep join; all counts, GMCR rows and
    statistics equal the oracle's."""
    data, queries = uneven_2k
    assert np.diff(np.asarray(data.graph_offsets)).max() > 64
    assert np.diff(np.asarray(queries.graph_offsets)).max() > 16
    counts, valid, stats = api.run_pipeline(queue, queries, data, K, find_first)
    rc, rv, rs = oracle.pipeline(queries, data, K, find_first)
    assert np.array_equal(valid, rv)
    assert np.array_equal(counts, rc)
    assert np.array_equal(stats[:4], expected_stats(queries, data, K, find_first, rs))


@pytest.fixture(scope="module")
def big_graphs():
    """400 graphs of 65-250 nodes (lognormal around 120 nodes, 10% fused ring systems; about 45%
    above 128 nodes) with the uneven workload's 1,024 queries: every graph takes the big-graph
    screen, in both size classes."""
    return graphs.generate("uneven-1m", seed=3, num_data_graphs=400, data_min_nodes=65,
                           data_lognormal_mu=4.8, data_lognormal_sigma=0.4)


@pytest.mark.parametrize("K", [0, 1, 6])
@pytest.mark.parametrize("find_first", [False, True])
def test_fused_big_graphs(queue, big_graphs, K, find_first):
    data, queries = big_graphs
    sizes = np.diff(data.graph_offsets.astype(np.int64))
    assert sizes.min() > 64 and (sizes > 128).sum() >= 100
    counts, valid, stats = api.run_pipeline(queue, queries, data, K, find_first)
    rc, rv, rs = oracle.pipeline(queries, data, K, find_first)
    assert rs[0] > 10000  # GMCR pairs: the screen's entries span several query blocks per graph
    assert np.array_equal(valid, rv)
    assert np.array_equal(counts, rc)
    assert np.array_equal(stats[:4], expected_stats(queries, data, K, find_first, rs))


def test_fused_big_graphs_autotuned(queue, big_graphs):
    """Every launch shape of the autotuner (both row-word splits of the size classes) leaves the
    results of the big-graph path unchanged; the best shape's run is compared again."""
    data, queries = big_graphs
    rc, rv, rs = oracle.pipeline(queries, data, 6, False)
    pipe = api.Pipeline(queue, api.DeviceBatchedCSRGraph(queue, queries),
                        api.DeviceBatchedCSRGraph(queue, data), refinement_steps=6)
    counts, valid, stats = pipe.alloc_outputs()
    t = pipe.autotune(counts, valid, runs=1)
    assert t["big_row_words"] in (2, 4)
    c, v = pipe.results(counts, valid)
    assert np.array_equal(c, rc) and np.array_equal(v, rv)
    counts.zero()
    pipe.run(counts, valid, stats)
    queue.wait()
    c, v = pipe.results(counts, valid)
    assert np.array_equal(c, rc) and np.array_equal(v, rv)
    assert np.array_equal(stats.to_numpy(np.uint64, 8)[:4], expected_stats(queries, data, 6, False, rs))


def test_fused_big_graphs_buffers_grow(queue):
    """600 queries of 2-3 carbons against 40 graphs of 65-250 nodes: every graph holds far more
    GMCR pairs than the big-graph screen's first buffers (128 entries per graph), so the run
    sizes its buffers from the counters and redoes the screen; results equal the oracle's.  The
    queries are copies of two graphs, so every query is screened (dedup off)."""
    data, _ = graphs.generate("uneven-1m", seed=5, num_data_graphs=40, data_min_nodes=65,
                              data_lognormal_mu=4.8, data_lognormal_sigma=0.4)
    queries = GraphBatch.from_graphs([([8, 8], [(0, 1)])] * 300 +
                                     [([8, 8, 8], [(0, 1), (1, 2)])] * 300)
    for find_first in (False, True):
        counts, valid, stats = api.run_pipeline(queue, queries, data, 2, find_first, dedup=False)
        rc, rv, rs = oracle.pipeline(queries, data, 2, find_first)
        assert rs[0] > 128 * 40  # GMCR pairs beyond the first sizing
        assert np.array_equal(valid, rv)
        assert np.array_equal(counts, rc)
        assert np.array_equal(stats[:4], rs)


def test_fused_big_graphs_new_query_batch(queue, big_graphs):
    """A pipeline over graphs above 64 nodes bound to a second query batch
    (bms2_pipeline_set_queries) serves it exactly, twice in a row."""
    data, queries = big_graphs
    second = graphs.generate("uneven-1m", seed=9, num_data_graphs=10, num_query_graphs=300)[1]
    dd = api.DeviceBatchedCSRGraph(queue, data)
    pipe = api.Pipeline(queue, api.DeviceBatchedCSRGraph(queue, queries), dd, refinement_steps=6)
    counts, valid, stats = pipe.alloc_outputs()
    pipe.run(counts, valid, stats)
    queue.wait()
    rc, rv, _ = oracle.pipeline(queries, data, 6, False)
    c, v = pipe.results(counts, valid)
    assert np.array_equal(c, rc) and np.array_equal(v, rv)
    dq2 = api.DeviceBatchedCSRGraph(queue, second)
    pipe.set_queries(dq2)
    counts, valid, stats = pipe.alloc_outputs()
    rc, rv, rs = oracle.pipeline(second, data, 6, False)
    rs = expected_stats(second, data, 6, False, rs)
    for _ in range(2):
        stats.zero()
        pipe.run(counts, valid, stats)
        queue.wait()
        c, v = pipe.results(counts, valid)
        assert np.array_equal(c, rc) and np.array_equal(v, rv)
        assert np.array_equal(stats.to_numpy(np.uint64, 8)[:4], rs)


def test_fused_uneven_order_and_trace(queue, uneven_2k):
    data, queries = uneven_2k
    dq = api.DeviceBatchedCSRGraph(queue, queries)
    dd = api.DeviceBatchedCSRGraph(queue, data)
    pipe = api.Pipeline(queue, dq, dd, refinement_steps=6)
    rc, rv, _ = oracle.pipeline(queries, data, 6, False)
    # connected random orders (random scores through the planner's connected ordering): a
    # disconnected order leaves early depths unconstrained and the search of a 32-node query on
    # a 250-node graph grows combinatorially
    from bms2_amd import learned
    rng = np.random.default_rng(3)
    order = np.zeros((queries.num_graphs, 32), np.uint8)
    for q, qi in enumerate(learned.query_infos(queries)):
        order[q, :qi.n] = learned.score_order(qi, rng.normal(size=qi.n))
    pipe.set_order(order)
    counts, valid, stats = pipe.alloc_outputs()
    for _ in range(2):  # the staged objects are reused across runs
        pipe.run(counts, valid, stats)
        queue.wait()
        c, v = pipe.results(counts, valid)
        assert np.array_equal(c, rc) and np.array_equal(v, rv)
    # a trace over a prefix holding big graphs writes exactly the prefix's counts
    tr = pipe.trace(700, queries.total_nodes)
    g = tr["graphs"]
    assert np.array_equal(tr["counts"], rc[:g])


@pytest.mark.parametrize("workload", ["molecules", "uneven"])
def test_graph_replays_match_eager(queue, molecules_1m_shard, uneven_2k, workload):
    """From the second run on, the chunk kernels replay from a hipGraph; every replay, with
    eager work (traces, order changes, another pipeline) between replays, equals the oracle and
    the pipeline that never captures."""
    data, queries = molecules_1m_shard if workload == "molecules" else uneven_2k
    rc, rv, rs = oracle.pipeline(queries, data, 6, False)
    rs = expected_stats(queries, data, 6, False, rs)
    dq = api.DeviceBatchedCSRGraph(queue, queries)
    dd = api.DeviceBatchedCSRGraph(queue, data)
    eager = api.Pipeline(queue, dq, dd, refinement_steps=6, use_graph=False)
    ec, ev, es = eager.alloc_outputs()
    eager.run(ec, ev, es)
    queue.wait()
    assert np.array_equal(eager.results(ec, ev)[0], rc)
    pipe = api.Pipeline(queue, dq, dd, refinement_steps=6)
    counts, valid, stats = pipe.alloc_outputs()
    for i in range(4):
        counts.zero()
        pipe.run(counts, valid, stats)
        queue.wait()
        c, v = pipe.results(counts, valid)
        assert np.array_equal(c, rc), f"run {i}"
        assert np.array_equal(v, rv), f"run {i}"
        assert np.array_equal(stats.to_numpy(np.uint64, 8)[:4], rs), f"run {i}"
        # eager work between replays
        pipe.trace(300, queries.total_nodes)
        eager.run(ec, ev, es)
        queue.wait()


def test_autotune_keeps_results(queue, uneven_2k):
    data, queries = uneven_2k
    rc, rv, rs = oracle.pipeline(queries, data, 2, False)
    rs = expected_stats(queries, data, 2, False, rs)
    pipe = api.Pipeline(queue, api.DeviceBatchedCSRGraph(queue, queries),
                        api.DeviceBatchedCSRGraph(queue, data), refinement_steps=2)
    counts, valid, stats = pipe.alloc_outputs()
    t = pipe.autotune(counts, valid, runs=1)
    assert t["candidates"] >= 6 and t["slots"] in (8, 16) and t["best_ms"] > 0
    c, v = pipe.results(counts, valid)
    assert np.array_equal(c, rc) and np.array_equal(v, rv)
    counts.zero()
    pipe.run(counts, valid, stats)
    queue.wait()
    c, v = pipe.results(counts, valid)
    assert np.array_equal(c, rc) and np.array_equal(v, rv)
    assert np.array_equal(stats.to_numpy(np.uint64, 8)[:4], rs)


def _random_labelled(rng, n_graphs, nmin, nmax, labels):
    """Random connected graphs (spanning tree plus a few extra edges), labels uniform in labels."""
    out = []
    for _ in range(n_graphs):
        n = int(rng.integers(nmin, nmax + 1))
        edges = {(int(rng.integers(0, v)), v) for v in range(1, n)}
        for _ in range(int(rng.integers(0, 4))):
            a, b = sorted(int(x) for x in rng.choice(n, 2, replace=False))
            edges.add((a, b))
        out.append(([int(x) for x in rng.choice(labels, n)], sorted(edges)))
    return out


@pytest.mark.parametrize("qlabels", [8, 13])
def test_shell_histogram_label_sets(queue, qlabels):
    """Queries over at most 8 labels use 2-word shell levels restricted to those labels, more
    labels the full 4-word levels; data graphs carry 13 labels either way."""
    rng = np.random.default_rng(qlabels)
    dgs = _random_labelled(rng, 300, 8, 40, np.arange(13))
    data = graphs.GraphBatch.from_graphs(dgs)
    # queries: connected pieces of data graphs relabelled into the first qlabels labels
    qgs = []
    for labs, edges in _random_labelled(rng, 60, 3, 7, np.arange(qlabels)):
        qgs.append((labs, edges))
    for g in range(0, 300, 10):  # induced pieces of data graphs keep matches
        labs, edges = dgs[g]
        keep = {0}
        for a, b in edges:
            if len(keep) < 5 and (a in keep) != (b in keep):
                keep |= {a, b}
        ids = sorted(keep)
        pos = {v: i for i, v in enumerate(ids)}
        ql = [min(labs[v], qlabels - 1) for v in ids]
        qgs.append((ql, [(pos[a], pos[b]) for a, b in edges if a in pos and b in pos]))
    queries = graphs.GraphBatch.from_graphs(qgs)
    assert len(set(np.asarray(queries.node_labels).tolist())) == qlabels
    for K in (2, 6):
        counts, valid, stats = api.run_pipeline(queue, queries, data, K, False)
        rc, rv, rs = oracle.pipeline(queries, data, K, False)
        assert np.array_equal(valid, rv) and np.array_equal(counts, rc)
        assert np.array_equal(stats[:4], expected_stats(queries, data, K, False, rs))


@pytest.fixture(scope="module")
def large_queries_small_graphs(uneven_2k):
    """Queries of 17-32 nodes of the uneven workload against its first 1,000 data graphs of at
    most 64 nodes: the staged join's 32-deep variant over precomputed adjacency rows."""
    data, queries = uneven_2k
    n = np.diff(np.asarray(data.graph_offsets))
    small = [int(i) for i in np.flatnonzero(n <= 64)[:1000]]
    qn = np.diff(np.asarray(queries.graph_offsets))
    big = [int(i) for i in np.flatnonzero(qn > 16)]
    return (graphs.GraphBatch.from_graphs([data.graph(i) for i in small]),
            graphs.GraphBatch.from_graphs([queries.graph(i) for i in big]))


@pytest.mark.parametrize("K,find_first", [(2, False), (6, False), (6, True)])
def test_staged_large_queries_small_graphs(queue, large_queries_small_graphs, K, find_first):
    data, queries = large_queries_small_graphs
    assert np.diff(np.asarray(data.graph_offsets)).max() <= 64
    assert np.diff(np.asarray(queries.graph_offsets)).min() > 16
    out = api.run_staged(queue, queries, data, K=K, find_first=find_first, keep_domains=False)
    ref = oracle.staged(queries, data, K, find_first=find_first, keep_domains=False)
    assert len(ref.gmcr_indices) > 100
    off, idx = out["gmcr"]
    assert np.array_equal(off, ref.gmcr_offsets) and np.array_equal(idx, ref.gmcr_indices)
    assert np.array_equal(out["counts"], ref.counts)
    assert out["total"] == ref.total


def test_join_rejects_small_pair_buffer(queue, molecules_2k):
    """A per-pair count buffer smaller than the GMCR is refused on the host before any kernel
    runs (both joins), and the same call with the right size succeeds."""
    data, queries = molecules_2k
    sub = data.subset(0, 200)
    dq = api.DeviceBatchedCSRGraph(queue, queries)
    dd = api.DeviceBatchedCSRGraph(queue, sub)
    sig = api.Signature(queue, dd.total_nodes, dq.total_nodes)
    path = api.PathSignature(queue, dd.total_nodes, dq.total_nodes)
    cyc = api.CycleSignature(queue, dd.total_nodes, dq.total_nodes)
    cand = api.Candidates(queue, dq.total_nodes, dd.total_nodes)
    for g, fn in ((dd, "data"), (dq, "query")):
        getattr(sig, f"generate_{fn}_signatures")(g)
        getattr(path, f"generate_{fn}_path_signatures")(g)
        getattr(cyc, f"generate_{fn}_cycle_signatures")(g)
    api.filter_candidates(queue, dq, dd, sig, path, cyc, cand)
    gm = api.GMCR(queue)
    gm.generate_gmcr(dq, dd, cand)
    queue.wait()
    T = gm.total
    assert T > 1
    total = api.DeviceBuffer(queue, 8, zero=True)
    short = api.DeviceBuffer(queue, 8 * (T - 1), zero=True)
    with pytest.raises(api.Bms2Error) as e:
        api.join_candidates(queue, dq, dd, cand, gm, total, False, short)
    assert e.value.code == _lib.ERR_CAPACITY
    end = int(queries.graph_offsets[1] - queries.graph_offsets[0])
    part, n = api.join_partial_candidates(queue, dq, dd, cand, end)
    with pytest.raises(api.Bms2Error) as e:
        api.join_with_partial_matches(queue, dq, dd, cand, gm, end, part, n, total, short, False)
    assert e.value.code == _lib.ERR_CAPACITY
    assert int(total.to_numpy(np.uint64)[0]) == 0  # nothing ran
    full = api.DeviceBuffer(queue, 8 * T, zero=True)
    api.join_candidates(queue, dq, dd, cand, gm, total, False, full)
    queue.wait()
    ref = oracle.staged(queries, sub, 0, keep_domains=False)
    assert np.array_equal(full.to_numpy(np.uint64, T), ref.counts)
