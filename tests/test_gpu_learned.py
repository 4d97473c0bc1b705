"""Learned matching order on the GPU: join traces and orders against the CPU oracle, the HIP
network kernels against the float64 restatement (oracle/mlp.py), and the gain of a network
trained on the GPU over identity order with bit-exact counts."""
import numpy as np
import pytest

import oracle
from oracle import mlp as omlp
from bms2_amd import api, graphs, learned

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def queue():
    q = api.Queue(0)
    yield q
    q.close()


def _random_orders(queries, rng):
    go = np.asarray(queries.graph_offsets)
    o = np.zeros((queries.num_graphs, 32), np.uint8)
    flat = np.zeros(queries.total_nodes, np.uint32)
    for q in range(queries.num_graphs):
        n = int(go[q + 1] - go[q])
        p = rng.permutation(n)
        o[q, :n] = p
        flat[go[q]:go[q] + n] = p
    return o, flat


def test_trace_and_order_match_oracle(queue, molecules_2k):
    data, queries = molecules_2k
    dq = api.DeviceBatchedCSRGraph(queue, queries)
    dd = api.DeviceBatchedCSRGraph(queue, data)
    pipe = api.Pipeline(queue, dq, dd, refinement_steps=6)
    ref = oracle.staged(queries, data, 6, keep_domains=False)
    off, idx = ref.gmcr_offsets, ref.gmcr_indices
    D, Q = data.num_graphs, queries.num_graphs
    go_d = np.asarray(data.graph_offsets, np.int64)
    go_q = np.asarray(queries.graph_offsets, np.int64)
    # oracle per-query-node domain sums, GMCR pairs per query
    bits = np.unpackbits(ref.final_domains.view(np.uint8), axis=1, bitorder="little")
    domsum = np.zeros(queries.total_nodes, np.uint64)
    qpairs = np.zeros(Q, np.uint64)
    pair_g = np.repeat(np.arange(D), np.diff(off.astype(np.int64)))
    for t, (g, q) in enumerate(zip(pair_g, idx)):
        b, e = go_q[q], go_q[q + 1]
        domsum[b:e] += bits[b:e, go_d[g]:go_d[g + 1]].sum(axis=1).astype(np.uint64)
        qpairs[q] += 1

    tr = pipe.trace(D, queries.total_nodes, pair_nodes=True)
    assert tr["graphs"] == D
    assert np.array_equal(tr["domsum"], domsum)
    assert np.array_equal(tr["qpairs"], qpairs)
    dense_nodes = np.zeros((D, Q), np.uint64)
    dense_nodes[pair_g, idx] = ref.dfs_nodes
    assert np.array_equal(tr["pair_nodes"].astype(np.uint64), dense_nodes)
    assert np.array_equal(tr["qnodes"], dense_nodes.sum(axis=0))

    # a random matching order: same counts, DFS nodes equal to the oracle's under that order
    rng = np.random.default_rng(5)
    o, flat = _random_orders(queries, rng)
    pipe.set_order(o)
    tr2 = pipe.trace(D, queries.total_nodes, pair_nodes=True)
    _, rcounts, rnodes = oracle.join(queries, data, ref.final_domains, off, idx, False, flat)
    dense_counts = np.zeros((D, Q), np.uint64)
    dense_counts[pair_g, idx] = rcounts
    assert np.array_equal(tr2["counts"], dense_counts)
    assert np.array_equal(tr2["counts"], tr["counts"])
    dn = np.zeros((D, Q), np.uint64)
    dn[pair_g, idx] = rnodes
    assert np.array_equal(tr2["pair_nodes"].astype(np.uint64), dn)
    assert not np.array_equal(tr2["qnodes"], tr["qnodes"])

    # full runs under the order equal the oracle's fused counts
    counts, valid, stats = pipe.alloc_outputs()
    pipe.run(counts, valid, stats)
    queue.wait()
    c, v = pipe.results(counts, valid)
    rc, rv, _ = oracle.pipeline(queries, data, 6, False)
    assert np.array_equal(c, rc) and np.array_equal(v, rv)

    bad = o.copy()
    bad[0, 1] = bad[0, 0]
    with pytest.raises(api.Bms2Error):
        pipe.set_order(bad)
    pipe.set_order(None)
    tr3 = pipe.trace(D, queries.total_nodes)
    assert np.array_equal(tr3["qnodes"], tr["qnodes"])


def test_trace_prefix(queue, molecules_2k):
    data, queries = molecules_2k
    pipe = api.Pipeline(queue, api.DeviceBatchedCSRGraph(queue, queries),
                        api.DeviceBatchedCSRGraph(queue, data), refinement_steps=2)
    tr = pipe.trace(500, queries.total_nodes)
    assert tr["graphs"] == api.PACK_WINDOW  # whole pack windows
    rc, _, rs = oracle.pipeline(queries, data, 2, False, g_end=tr["graphs"])
    assert np.array_equal(tr["counts"], rc)
    assert int(tr["qpairs"].sum()) == int(rs[0])
    assert int(tr["qnodes"].sum()) == int(rs[3])
    # past the first window: the rest of the 2,000-graph batch
    tr = pipe.trace(api.PACK_WINDOW + 1, queries.total_nodes)
    assert tr["graphs"] == data.num_graphs


def test_screen_only_trace_equals_full_trace(queue, molecules_2k):
    """The planner's trace runs the screens only (null DFS-node buffers): its domain sums and
    GMCR pairs equal the full trace's, and it reports no DFS nodes or counts."""
    data, queries = molecules_2k
    pipe = api.Pipeline(queue, api.DeviceBatchedCSRGraph(queue, queries),
                        api.DeviceBatchedCSRGraph(queue, data), refinement_steps=6)
    full = pipe.trace(data.num_graphs, queries.total_nodes)
    screen = pipe.trace(data.num_graphs, queries.total_nodes, counts=False, joins=False)
    assert screen["graphs"] == full["graphs"]
    assert np.array_equal(screen["domsum"], full["domsum"])
    assert np.array_equal(screen["qpairs"], full["qpairs"])
    assert "qnodes" not in screen and "counts" not in screen
    # a full run afterwards is unaffected
    rc, _, _ = oracle.pipeline(queries, data, 6, False)
    counts, valid, stats = pipe.alloc_outputs()
    pipe.run(counts, valid, stats)
    queue.wait()
    assert np.array_equal(pipe.results(counts, valid)[0], rc)


def _rand_batch(rows, rng):
    x = rng.normal(size=(rows, omlp.IN)).astype(np.float32)
    t = np.stack([rng.normal(size=rows), rng.normal(size=rows),
                  rng.integers(0, 8, size=rows).astype(np.float64),
                  rng.uniform(0.5, 2.0, size=rows)], axis=1).astype(np.float32)
    return x, t


@pytest.mark.parametrize("rows", [1, 63, 64, 1000])
def test_mlp_forward_and_gradient_match_oracle(queue, rows):
    rng = np.random.default_rng(rows)
    net = learned.OrderNet(queue, seed=7)
    p = (rng.normal(scale=0.2, size=net.num_params)).astype(np.float32)
    net.set_params(p)
    assert np.array_equal(net.get_params(), p)
    x, t = _rand_batch(rows, rng)
    # rows whose pre-activations sit within f32 rounding of a ReLU kink are redrawn: there the
    # f32 and f64 masks may differ and the subgradient is a matter of rounding
    while True:
        bad = omlp.min_abs_preactivation(p.astype(np.float64), x) < 1e-5
        if not bad.any():
            break
        x[bad] = rng.normal(size=(int(bad.sum()), omlp.IN)).astype(np.float32)
    # f32 forward: float64 restatement, tolerance 1e-5 of the output scale (f32 accumulation)
    y = net.forward(x, bf16=False)
    yr, _ = omlp.forward(p.astype(np.float64), x.astype(np.float64))
    assert np.abs(y - yr).max() <= 1e-5 * (1 + np.abs(yr).max())
    # bf16 forward: same rounding points as the restatement; accumulation order may move an
    # activation across a bf16 rounding boundary, so allow 2% of the output scale
    yb = net.forward(x, bf16=True)
    ybr, _ = omlp.forward(p.astype(np.float64), x, bf16=True)
    assert np.abs(yb - ybr).max() <= 2e-2 * (1 + np.abs(ybr).max())
    # gradient: tolerance 1e-4 of the largest component (f32 activations and deltas)
    loss, g = net.loss_grad(x, t, depth_weight=0.5)
    lr, gr = omlp.loss_grad(p.astype(np.float64), x, t, depth_weight=0.5)
    assert abs(loss - lr) <= 1e-5 * (1 + abs(lr))
    assert np.abs(g - gr).max() <= 1e-4 * (np.abs(gr).max() + 1e-6)


def test_mlp_adam_matches_oracle(queue):
    rng = np.random.default_rng(11)
    net = learned.OrderNet(queue, seed=3)
    p0 = net.get_params().astype(np.float64)
    x, t = _rand_batch(300, rng)
    while True:
        bad = omlp.min_abs_preactivation(p0, x) < 1e-5
        if not bad.any():
            break
        x[bad] = rng.normal(size=(int(bad.sum()), omlp.IN)).astype(np.float32)
    dx, dt = api.DeviceBuffer.from_array(queue, x), api.DeviceBuffer.from_array(queue, t)
    opt = omlp.Adam(omlp.NUM_PARAMS, lr=1e-3)
    p = p0.copy()
    firm = np.ones(omlp.NUM_PARAMS, bool)
    for _ in range(4):
        net.train_step(dx, dt, 300, lr=1e-3, depth_weight=1.0)
        _, g = omlp.loss_grad(p, x, t, 1.0)
        # Adam normalises each component: one whose gradient is within f32 rounding of zero may
        # move by up to 2 lr either way, so only components with a clear gradient are compared
        firm &= np.abs(g) > 1e-4 * np.abs(g).max()
        p = opt.step(p, g)
    got = net.get_params().astype(np.float64)
    assert firm.mean() > 0.5
    # after 4 steps each firm parameter moved by about 4e-3; f32 state agrees to a few 1e-6
    assert np.abs(got - p)[firm].max() <= 2e-5
    assert np.abs(got - p).max() <= 8e-3 + 1e-6
    assert np.abs(got - p0).max() > 1e-3


def test_mlp_init_deterministic(queue):
    a = learned.OrderNet(queue, seed=42).get_params()
    b = learned.OrderNet(queue, seed=42).get_params()
    c = learned.OrderNet(queue, seed=43).get_params()
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    lim = np.sqrt(6.0 / 32)
    assert np.abs(a[:32 * 64]).max() <= lim and np.all(a[32 * 64:32 * 64 + 64] == 0)


def test_learned_order_gains_over_identity(queue):
    """A network trained on the GPU (traces of seeds 1-3) expands fewer DFS nodes than identity
    order on a held-out seed, with bit-exact counts."""
    rows = [learned.collect_seed(queue, s, num_data_graphs=3000, depth_labels=False)
            for s in (1, 2, 3)]
    X = np.concatenate([r.X for r in rows])
    T = np.concatenate([r.T for r in rows])
    assert all(r.best_nodes < r.identity_nodes for r in rows)
    net = learned.OrderNet(queue, seed=1)
    losses = learned.train(net, X, T, epochs=30, batch=2048)
    assert losses[-1] < 0.7 * losses[0]

    data, queries = graphs.generate("molecules-1m", seed=100, num_data_graphs=6000)
    pipe = api.Pipeline(queue, api.DeviceBatchedCSRGraph(queue, queries),
                        api.DeviceBatchedCSRGraph(queue, data), refinement_steps=6)
    ident = pipe.trace(data.num_graphs, queries.total_nodes)
    pl = learned.plan(pipe, net, queries, sample_graphs=2000)
    assert pl.learned.any()
    lrn = pipe.trace(data.num_graphs, queries.total_nodes)
    assert np.array_equal(lrn["counts"], ident["counts"])
    n_id, n_l = int(ident["qnodes"].sum()), int(lrn["qnodes"].sum())
    print(f"held-out seed 100: DFS nodes identity {n_id}, learned {n_l} ({n_l / n_id:.3f}); "
          f"{int(pl.learned.sum())} of {queries.num_graphs} queries reordered")
    assert n_l < 0.97 * n_id
