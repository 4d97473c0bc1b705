"""GPU-side CSR packing (bms2_device_csr_pack, createDeviceCSRGraph graph.hpp:249-310 with the
offsets added on the device) against the host-packed upload, and its validation."""
import ctypes as C

import numpy as np
import pytest

from bms2_amd import _lib, api, graphs
from bms2_amd.graphs import GraphBatch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def queue():
    q = api.Queue(0)
    yield q
    q.close()


def _same(a: GraphBatch, b: GraphBatch):
    for f in ("graph_offsets", "row_offsets", "column_indices", "node_labels", "edge_labels"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f


@pytest.mark.parametrize("which", ["molecules", "queries", "uneven", "empty-graphs", "many"])
def test_pack_equals_host_packing(queue, which):
    if which in ("molecules", "queries"):
        d, q = graphs.generate("molecules-2k", seed=4)
        host = d if which == "molecules" else q
    elif which == "many":
        # more graphs than one scan block covers (the offsets go through the multi-block scan)
        host, _ = graphs.generate("molecules-1m", seed=3, num_data_graphs=40_000)
    elif which == "uneven":
        host, _ = graphs.generate("uneven-1m", seed=2, num_data_graphs=400, num_query_graphs=4)
        assert int(np.diff(host.graph_offsets).max()) > 64
    else:
        host = GraphBatch.from_graphs([([], []), ([3, 4], [(0, 1, 2)]), ([], []),
                                       ([1, 1, 1], [(0, 1), (1, 2), (0, 2)]), ([7], [])])
    print(f"[pack] {which}: {host.num_graphs} graphs, {host.total_nodes} nodes", flush=True)
    dp = api.DeviceBatchedCSRGraph.pack(queue, host)
    dc = api.DeviceBatchedCSRGraph(queue, host)
    _same(dp.to_host(), dc.to_host())
    _same(dp.to_host(), host)
    assert dp.c.max_graph_nodes == dc.c.max_graph_nodes
    assert dp.c.max_degree == dc.c.max_degree


def test_pack_empty_batch(queue):
    host = GraphBatch.from_graphs([])
    dp = api.DeviceBatchedCSRGraph.pack(queue, host)
    assert dp.num_graphs == 0 and dp.total_nodes == 0


def _raw_pack(queue, nodes, ro, ci, nl=None, el=None):
    """bms2_device_csr_pack on raw arrays; totals are the arrays' own lengths."""
    nodes = np.asarray(nodes, np.uint32)
    ro = np.asarray(ro, np.uint32)
    ci = np.asarray(ci, np.uint32)
    nl = np.zeros(max(len(ro), 1), np.uint8) if nl is None else np.asarray(nl, np.uint8)
    el = np.zeros(max(len(ci), 1), np.uint8) if el is None else np.asarray(el, np.uint8)
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    out = _lib.DeviceCsr()
    rc = _lib.lib().bms2_device_csr_pack(queue.h, len(nodes), p(nodes), p(ro), p(ci), p(nl),
                                        p(el), len(ro), len(ci), C.byref(out))
    if rc == 0:
        _lib.lib().bms2_device_csr_destroy(queue.h, C.byref(out))
    return rc


def test_pack_validation(queue):
    # a valid triangle for reference
    assert _raw_pack(queue, [3], [2, 4, 6], [1, 2, 0, 2, 0, 1]) == 0
    # asymmetric: 0 lists 1, 1 does not list 0
    assert _raw_pack(queue, [3], [2, 3, 5], [1, 2, 2, 0, 1]) == _lib.ERR_INVALID
    # self-loop
    assert _raw_pack(queue, [2], [2, 3], [1, 0, 0]) == _lib.ERR_UNSUPPORTED
    # endpoint outside its graph (local id 5 in a 2-node graph)
    assert _raw_pack(queue, [2], [1, 2], [5, 0]) == _lib.ERR_INVALID
    # row ends not monotone
    assert _raw_pack(queue, [3], [2, 1, 4], [1, 2, 0, 0]) == _lib.ERR_INVALID
    # node counts disagree with the node arrays (4 entries, counts sum to 3)
    assert _raw_pack(queue, [3], [0, 0, 0, 0], []) == _lib.ERR_INVALID
    # row ends give 2 CSR entries, the entry arrays hold 3
    assert _raw_pack(queue, [2], [1, 2], [1, 0, 0]) == _lib.ERR_INVALID
    # a graph above the size limit (a 300-node path)
    n = 300
    ro = np.cumsum([1] + [2] * (n - 2) + [1])
    ci = []
    for u in range(n):
        if u > 0:
            ci.append(u - 1)
        if u < n - 1:
            ci.append(u + 1)
    assert _raw_pack(queue, [n], ro, ci) == _lib.ERR_UNSUPPORTED
