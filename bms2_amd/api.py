"""Python mirror of the reference's sigmo:: API over the C ABI (libbms2.so).

Names and argument meaning follow library/include of AnySparse/BMS2:
  Queue                   sycl::queue                          (BMS2.cpp:14)
  create_device_csr_graph createDeviceCSRGraph                 (graph.hpp:251-310)
  Signature               signature::Signature<PowerGraph, 4>  (signature.hpp:18-390)
  PathSignature           signature::PathSignature             (signature.hpp:406-532)
  CycleSignature          signature::CycleSignature            (signature.hpp:535-732)
  Candidates              candidates::Candidates               (candidates.hpp:10-200)
  filter_candidates       isomorphism::filter::filterCandidates  (isomorphism.hpp:30-104)
  refine_candidates       isomorphism::filter::refineCandidates  (isomorphism.hpp:106-203)
  GMCR                    isomorphism::mapping::GMCR           (gmcr.hpp:16-145)
  join_candidates         isomorphism::join::joinCandidates    (isomorphism.hpp:291-472)
  Pipeline                the fused stage sequence of BMS2.cpp:131-234
Every call goes to HIP kernels; errors raise Bms2Error with the library's message.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import Bms2Error, check
from .graphs import GraphBatch

SCOPE_DATA = _lib.SCOPE_DATA
SCOPE_QUERY = _lib.SCOPE_QUERY
PACK_WINDOW = 1024  # BMS2_PACK_WINDOW: data packs are planned per window of graphs


def _L():
    return _lib.lib()


class Event:
    """utils::BatchedEvent: elapsed device time of one call's kernels."""

    def __init__(self, handle):
        self._h = handle

    def elapsed_ms(self) -> float:
        ms = C.c_float()
        check(_L().bms2_event_elapsed_ms(self._h, C.byref(ms)))
        return float(ms.value)

    def __del__(self):
        if getattr(self, "_h", None):
            _L().bms2_event_destroy(self._h)
            self._h = None


def _ev_out(want: bool):
    return C.c_void_p() if want else None


def _ev(ptr):
    return Event(ptr.value) if ptr is not None and ptr.value else None


class Queue:
    """A HIP stream on one device (sycl::queue).

    Device objects created on a queue (buffers, graph batches, handles) count themselves against
    it while alive: close() on a queue that still has some only marks it, and the last of them to
    be freed destroys it, so no object is ever freed through a destroyed queue, whatever order the
    garbage collector finalises them in."""

    def __init__(self, device: int = 0, stream: int | None = None):
        self._deps = 0
        self._closing = False
        h = C.c_void_p()
        if stream is None:
            check(_L().bms2_queue_create(device, C.byref(h)))
        else:
            check(_L().bms2_queue_wrap(device, C.c_void_p(stream), C.byref(h)))
        self.h = h
        self.device = device

    def _acquire(self) -> None:
        self._deps += 1

    def _release(self) -> None:
        self._deps -= 1
        if self._closing and self._deps <= 0:
            self._destroy()

    def _destroy(self) -> None:
        if self.h:
            h, self.h = self.h, None
            check(_L().bms2_queue_destroy(h))

    def wait(self) -> None:
        check(_L().bms2_queue_wait(self.h))

    @property
    def stream(self) -> int:
        return int(_L().bms2_queue_stream(self.h) or 0)

    def close(self) -> None:
        self._closing = True
        if self._deps <= 0:
            self._destroy()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    """Plain device allocation (u8 bytes) owned by Python."""

    def __init__(self, q: Queue, nbytes: int, zero: bool = True):
        self.q = q
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(_L().bms2_device_malloc(q.h, max(self.nbytes, 8), C.byref(p)))
        self.ptr = p.value
        q._acquire()
        if zero:
            check(_L().bms2_memset_zero(q.h, C.c_void_p(self.ptr), max(self.nbytes, 8)))

    @staticmethod
    def from_array(q: Queue, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = DeviceBuffer(q, a.nbytes, zero=False)
        check(_L().bms2_memcpy_h2d(q.h, C.c_void_p(b.ptr), a.ctypes.data_as(C.c_void_p), a.nbytes))
        q.wait()
        return b

    def to_numpy(self, dtype, count: int | None = None) -> np.ndarray:
        dt = np.dtype(dtype)
        n = self.nbytes // dt.itemsize if count is None else count
        out = np.empty(n, dtype=dt)
        if n:
            check(_L().bms2_memcpy_d2h(self.q.h, out.ctypes.data_as(C.c_void_p),
                                       C.c_void_p(self.ptr), n * dt.itemsize))
        return out

    def zero(self) -> None:
        check(_L().bms2_memset_zero(self.q.h, C.c_void_p(self.ptr), max(self.nbytes, 8)))

    def free(self) -> None:
        if self.ptr:
            ptr, self.ptr = self.ptr, None
            try:
                check(_L().bms2_device_free(self.q.h, C.c_void_p(ptr)))
            finally:
                self.q._release()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceBatchedCSRGraph:
    """sigmo::DeviceBatchedCSRGraph (graph.hpp:58-108) created by createDeviceCSRGraph."""

    def __init__(self, q: Queue, host: GraphBatch):
        self.q = q
        self.c = _lib.DeviceCsr()
        h = host.to_c()
        check(_L().bms2_device_csr_create(q.h, C.byref(h), C.byref(self.c)))
        q._acquire()
        self.host = host

    @staticmethod
    def pack(q: Queue, host: GraphBatch) -> "DeviceBatchedCSRGraph":
        """The same batch packed on the GPU (bms2_device_csr_pack): the graphs' local arrays go
        up end to end and the device adds the offsets and validates."""
        self = DeviceBatchedCSRGraph.__new__(DeviceBatchedCSRGraph)
        self.q = q
        self.c = _lib.DeviceCsr()
        nodes, ro_local, ci_local = host.local_arrays()
        el = np.ascontiguousarray(host.edge_labels, np.uint8)
        nl = np.ascontiguousarray(host.node_labels, np.uint8)
        p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
        check(_L().bms2_device_csr_pack(q.h, host.num_graphs, p(nodes), p(ro_local), p(ci_local),
                                        p(nl), p(el), host.total_nodes, len(ci_local),
                                        C.byref(self.c)))
        q._acquire()
        self.host = host
        return self

    def to_host(self) -> GraphBatch:
        """The batch's device arrays copied back."""
        q = self.q

        def get(ptr, n, dt):
            out = np.empty(n, dt)
            if n:
                check(_L().bms2_memcpy_d2h(q.h, out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr),
                                           out.nbytes))
            return out

        c = self.c
        return GraphBatch(get(c.graph_offsets, c.num_graphs + 1, np.uint32),
                          get(c.row_offsets, c.total_nodes + 1, np.uint32),
                          get(c.column_indices, c.total_edges, np.uint32),
                          get(c.node_labels, c.total_nodes, np.uint8),
                          get(c.edge_labels, c.total_edges, np.uint8))

    @property
    def num_graphs(self) -> int:
        return int(self.c.num_graphs)

    @property
    def total_nodes(self) -> int:
        return int(self.c.total_nodes)

    @property
    def total_edges(self) -> int:
        return int(self.c.total_edges)

    def alloc_size(self) -> int:
        return int(_L().bms2_device_csr_alloc_size(C.byref(self.c)))

    def destroy(self) -> None:
        if self.c.graph_offsets:
            try:
                check(_L().bms2_device_csr_destroy(self.q.h, C.byref(self.c)))
            finally:
                self.c.graph_offsets = None
                self.q._release()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def create_device_csr_graph(q: Queue, host: GraphBatch) -> DeviceBatchedCSRGraph:
    return DeviceBatchedCSRGraph(q, host)


class _Handle:
    """An opaque library handle on queue self.q, destroyed with the Python object."""

    _destroy = ""

    def _bind(self, h) -> None:
        self.h = h
        self.q._acquire()

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.h = None
            try:
                getattr(_L(), self._destroy)(h)
            except Exception:
                pass
            self.q._release()


class Signature(_Handle):
    """Signature<PowerGraph, 4>(queue, data_nodes, query_nodes) (signature.hpp:156-167)."""

    _destroy = "bms2_signature_destroy"

    def __init__(self, q: Queue, data_nodes: int, query_nodes: int):
        self.q, self.n = q, (int(data_nodes), int(query_nodes))
        h = C.c_void_p()
        check(_L().bms2_signature_create(q.h, data_nodes, query_nodes, C.byref(h)))
        self._bind(h)

    def _gen(self, g, scope, event):
        e = _ev_out(event)
        check(_L().bms2_signature_generate(self.h, C.byref(g.c), scope,
                                           C.byref(e) if e is not None else None))
        return _ev(e)

    def _ref(self, g, view_size, scope, event):
        e = _ev_out(event)
        check(_L().bms2_signature_refine(self.h, C.byref(g.c), view_size, scope,
                                         C.byref(e) if e is not None else None))
        return _ev(e)

    def generate_data_signatures(self, g, event=False):
        return self._gen(g, SCOPE_DATA, event)

    def generate_query_signatures(self, g, event=False):
        return self._gen(g, SCOPE_QUERY, event)

    def refine_data_signatures(self, g, view_size=1, event=False):
        return self._ref(g, view_size, SCOPE_DATA, event)

    def refine_query_signatures(self, g, view_size=1, event=False):
        return self._ref(g, view_size, SCOPE_QUERY, event)

    def host(self, scope: int) -> np.ndarray:
        out = np.zeros(self.n[scope], np.uint64)
        check(_L().bms2_signature_copy_to_host(self.h, scope, out.ctypes.data_as(C.c_void_p)))
        return out


class PathSignature(_Handle):
    _destroy = "bms2_path_signature_destroy"

    def __init__(self, q: Queue, data_nodes: int, query_nodes: int):
        self.q, self.n = q, (int(data_nodes), int(query_nodes))
        h = C.c_void_p()
        check(_L().bms2_path_signature_create(q.h, data_nodes, query_nodes, C.byref(h)))
        self._bind(h)

    def _gen(self, g, scope, event):
        e = _ev_out(event)
        check(_L().bms2_path_signature_generate(self.h, C.byref(g.c), scope,
                                                C.byref(e) if e is not None else None))
        return _ev(e)

    def generate_data_path_signatures(self, g, event=False):
        return self._gen(g, SCOPE_DATA, event)

    def generate_query_path_signatures(self, g, event=False):
        return self._gen(g, SCOPE_QUERY, event)

    def host(self, scope: int) -> np.ndarray:
        out = np.zeros((self.n[scope], 4), np.uint64)
        check(_L().bms2_path_signature_copy_to_host(self.h, scope, out.ctypes.data_as(C.c_void_p)))
        return out


class CycleSignature(_Handle):
    _destroy = "bms2_cycle_signature_destroy"

    def __init__(self, q: Queue, data_nodes: int, query_nodes: int):
        self.q, self.n = q, (int(data_nodes), int(query_nodes))
        h = C.c_void_p()
        check(_L().bms2_cycle_signature_create(q.h, data_nodes, query_nodes, C.byref(h)))
        self._bind(h)

    def _gen(self, g, scope, event):
        e = _ev_out(event)
        check(_L().bms2_cycle_signature_generate(self.h, C.byref(g.c), scope,
                                                 C.byref(e) if e is not None else None))
        return _ev(e)

    def generate_data_cycle_signatures(self, g, event=False):
        return self._gen(g, SCOPE_DATA, event)

    def generate_query_cycle_signatures(self, g, event=False):
        return self._gen(g, SCOPE_QUERY, event)

    def host(self, scope: int) -> np.ndarray:
        out = np.zeros(self.n[scope], np.uint64)
        check(_L().bms2_cycle_signature_copy_to_host(self.h, scope, out.ctypes.data_as(C.c_void_p)))
        return out


class Candidates(_Handle):
    """candidates::Candidates(queue, source_nodes, target_nodes), zero-filled."""

    _destroy = "bms2_candidates_destroy"

    def __init__(self, q: Queue, source_nodes: int, target_nodes: int):
        self.q = q
        self.rows = int(source_nodes)
        h = C.c_void_p()
        check(_L().bms2_candidates_create(q.h, source_nodes, target_nodes, C.byref(h)))
        self._bind(h)
        self.W = int(_L().bms2_candidates_words_per_row(h))

    def host(self) -> np.ndarray:
        out = np.zeros((self.rows, self.W), np.uint64)
        check(_L().bms2_candidates_copy_to_host(self.h, out.ctypes.data_as(C.c_void_p)))
        return out

    def row_counts(self) -> np.ndarray:
        out = np.zeros(self.rows, np.uint32)
        check(_L().bms2_candidates_row_counts(self.h, out.ctypes.data_as(C.c_void_p)))
        return out


def filter_candidates(q: Queue, query, data, sig: Signature, path: PathSignature,
                      cycle: CycleSignature, cand: Candidates, event=False):
    e = _ev_out(event)
    check(_L().bms2_filter_candidates(q.h, C.byref(query.c), C.byref(data.c), sig.h, path.h,
                                      cycle.h, cand.h, C.byref(e) if e is not None else None))
    return _ev(e)


def refine_candidates(q: Queue, query, data, sig: Signature, cand: Candidates, event=False):
    e = _ev_out(event)
    check(_L().bms2_refine_candidates(q.h, C.byref(query.c), C.byref(data.c), sig.h, cand.h,
                                      C.byref(e) if e is not None else None))
    return _ev(e)


class GMCR(_Handle):
    _destroy = "bms2_gmcr_destroy"

    def __init__(self, q: Queue):
        self.q = q
        h = C.c_void_p()
        check(_L().bms2_gmcr_create(q.h, C.byref(h)))
        self._bind(h)
        self.num_data = 0

    def generate_gmcr(self, query, data, cand: Candidates, event=False):
        e = _ev_out(event)
        check(_L().bms2_gmcr_generate(self.h, C.byref(query.c), C.byref(data.c), cand.h,
                                      C.byref(e) if e is not None else None))
        self.num_data = data.num_graphs
        return _ev(e)

    @property
    def total(self) -> int:
        return int(_L().bms2_gmcr_total(self.h))

    def host(self) -> tuple[np.ndarray, np.ndarray]:
        off = np.zeros(self.num_data + 1, np.uint32)
        idx = np.zeros(max(self.total, 1), np.uint32)
        check(_L().bms2_gmcr_copy_to_host(self.h, off.ctypes.data_as(C.c_void_p),
                                          idx.ctypes.data_as(C.c_void_p)))
        return off, idx[:self.total]


def join_candidates(q: Queue, query, data, cand: Candidates, gmcr: GMCR, num_matches: DeviceBuffer,
                    find_first: bool = True, pair_counts: DeviceBuffer | None = None, event=False):
    """joinCandidates: adds the match total into the device u64 `num_matches`; `pair_counts`
    (optional) receives one u64 per GMCR entry and must hold gmcr.total of them (Bms2Error with
    code ERR_CAPACITY otherwise, before any device work)."""
    e = _ev_out(event)
    check(_L().bms2_join_candidates(q.h, C.byref(query.c), C.byref(data.c), cand.h, gmcr.h,
                                    C.c_void_p(num_matches.ptr), int(find_first),
                                    C.c_void_p(pair_counts.ptr) if pair_counts else None,
                                    pair_counts.nbytes // 8 if pair_counts else 0,
                                    C.byref(e) if e is not None else None))
    return _ev(e)


def join_partial_candidates(q: Queue, query, data, cand: Candidates, end: int,
                            capacity: int | None = None, event=False):
    """joinPartialCandidates2 (the --use-cs step 1): every embedding of query 0's first `end`
    nodes, as device records of end + 1 int32 {mapped global data node ids, data graph id}.
    Returns (records DeviceBuffer, count[, Event]).  With capacity None the count is taken first
    and the buffer sized to it."""
    n = C.c_uint64(0)
    if capacity is None:
        rc = _L().bms2_join_partial_candidates(q.h, C.byref(query.c), C.byref(data.c), cand.h,
                                               int(end), None, 0, C.byref(n), None)
        if rc not in (_lib.OK, _lib.ERR_CAPACITY):
            check(rc)
        capacity = int(n.value)
    buf = DeviceBuffer(q, 4 * (int(end) + 1) * max(int(capacity), 1), zero=False)
    e = _ev_out(event)
    check(_L().bms2_join_partial_candidates(q.h, C.byref(query.c), C.byref(data.c), cand.h,
                                            int(end), C.c_void_p(buf.ptr), int(capacity),
                                            C.byref(n), C.byref(e) if e is not None else None))
    if event:
        return buf, int(n.value), _ev(e)
    return buf, int(n.value)


def join_with_partial_matches(q: Queue, query, data, cand: Candidates, gmcr: GMCR, end: int,
                              partial: DeviceBuffer, num_partial: int, num_matches: DeviceBuffer,
                              pair_counts: DeviceBuffer, find_first: bool = True, event=False):
    """joinWithPartialMatches (the --use-cs step 2): adds the total into the device u64
    `num_matches`; `pair_counts` (u64 per GMCR entry) receives the counts (0/1 flags when
    find_first)."""
    e = _ev_out(event)
    check(_L().bms2_join_with_partial_matches(
        q.h, C.byref(query.c), C.byref(data.c), cand.h, gmcr.h, int(end), C.c_void_p(partial.ptr),
        int(num_partial), C.c_void_p(num_matches.ptr), int(find_first),
        C.c_void_p(pair_counts.ptr), pair_counts.nbytes // 8,
        C.byref(e) if e is not None else None))
    return _ev(e)


class Pipeline(_Handle):
    """The fused stage sequence over a resident data batch."""

    _destroy = "bms2_pipeline_destroy"

    def __init__(self, q: Queue, query, data, refinement_steps: int = 0, find_first: bool = False,
                 use_graph: bool = True):
        self.q = q
        self.Q = query.num_graphs
        self.D = data.num_graphs
        # the pipeline reads the batches' device memory on every run: keep them alive with it
        self.query_graphs, self.data_graphs = query, data
        o = _lib.PipelineOptions()
        o.refinement_steps = refinement_steps
        o.find_first = int(find_first)
        o.no_graph = 0 if use_graph else 1
        h = C.c_void_p()
        check(_L().bms2_pipeline_create(q.h, C.byref(query.c), C.byref(data.c), C.byref(o),
                                        C.byref(h)))
        self._bind(h)
        self.VW = (self.Q + 63) // 64

    def alloc_outputs(self) -> tuple[DeviceBuffer, DeviceBuffer, DeviceBuffer]:
        counts = DeviceBuffer(self.q, 8 * self.D * self.Q, zero=False)
        valid = DeviceBuffer(self.q, 8 * self.D * self.VW, zero=False)
        stats = DeviceBuffer(self.q, 64, zero=True)
        return counts, valid, stats

    def run(self, counts: DeviceBuffer, valid: DeviceBuffer, stats: DeviceBuffer | None = None,
            event=False, kernel_event=False):
        """Enqueues one run.  Returns the whole-call Event (event=True), or the pair
        (whole-call Event, fused-kernel Event) when kernel_event=True."""
        e = _ev_out(event or kernel_event)
        k = _ev_out(kernel_event)
        check(_L().bms2_pipeline_run(self.h, C.c_void_p(counts.ptr), C.c_void_p(valid.ptr),
                                     C.c_void_p(stats.ptr) if stats else None,
                                     C.byref(e) if e is not None else None,
                                     C.byref(k) if k is not None else None))
        if kernel_event:
            return _ev(e), _ev(k)
        return _ev(e)

    def results(self, counts: DeviceBuffer, valid: DeviceBuffer, rows: int | None = None):
        """Host copies of the first `rows` data graphs' counts and GMCR bit rows."""
        rows = self.D if rows is None else min(int(rows), self.D)
        c = counts.to_numpy(np.uint64, rows * self.Q).reshape(rows, self.Q)
        v = valid.to_numpy(np.uint64, rows * self.VW).reshape(rows, self.VW)
        return c, v

    def set_queries(self, query) -> None:
        """Binds another query batch to the resident data batch (bms2_pipeline_set_queries)."""
        check(_L().bms2_pipeline_set_queries(self.h, C.byref(query.c)))
        self.query_graphs = query
        self.Q = query.num_graphs
        self.VW = (self.Q + 63) // 64

    def query_totals(self, counts: DeviceBuffer) -> tuple[np.ndarray, np.ndarray]:
        """Per query of the bound batch: data graphs matched and summed counts of the last run."""
        m = DeviceBuffer(self.q, 8 * max(self.Q, 1), zero=False)
        e = DeviceBuffer(self.q, 8 * max(self.Q, 1), zero=False)
        check(_L().bms2_pipeline_query_totals(self.h, C.c_void_p(counts.ptr), C.c_void_p(m.ptr),
                                              C.c_void_p(e.ptr)))
        return m.to_numpy(np.uint64, self.Q), e.to_numpy(np.uint64, self.Q)

    def autotune(self, counts: DeviceBuffer, valid: DeviceBuffer, runs: int = 1) -> dict:
        """Times the join launch shapes on this batch and keeps the fastest
        (bms2_pipeline_autotune); the outputs hold the last run's (identical) results."""
        t = _lib.Tuning()
        check(_L().bms2_pipeline_autotune(self.h, C.c_void_p(counts.ptr), C.c_void_p(valid.ptr),
                                          int(runs), C.byref(t)))
        return dict(small_waves=t.small_waves, large_waves=t.large_waves, slots=t.slots,
                    big_waves=t.big_waves, big_block_waves=t.big_block_waves,
                    big_row_words=t.big_row_words, candidates=t.candidates,
                    best_ms=round(float(t.best_ms), 3))

    def set_dedup(self, on: bool) -> int:
        """Screen one query per isomorphism class and copy its outputs to the others (on, the
        default) or screen every query (off); returns the number of queries that copy."""
        n = C.c_uint32(0)
        check(_L().bms2_pipeline_set_dedup(self.h, int(bool(on)), C.byref(n)))
        return int(n.value)

    def set_order(self, order: np.ndarray | None) -> None:
        """Join matching order: order[q, d] = local query node matched at depth d (u8 [Q, 32]);
        None restores identity."""
        if order is None:
            check(_L().bms2_pipeline_set_order(self.h, None))
            return
        o = np.ascontiguousarray(order, np.uint8)
        if o.shape != (self.Q, 32):
            raise ValueError(f"order must have shape ({self.Q}, 32), got {o.shape}")
        check(_L().bms2_pipeline_set_order(self.h, o.ctypes.data_as(C.c_void_p)))

    def trace(self, max_graphs: int, query_nodes: int, pair_nodes: bool = False,
              counts: bool = True, joins: bool = True) -> dict:
        """Join traces over the leading data graphs (at least max_graphs of them): summed
        domain size per query node, GMCR pairs and DFS nodes per query, optionally DFS nodes per
        (graph, query) pair, plus (counts=True) the counts of the traced graphs.  joins=False
        runs the screens only: domain sums and GMCR pairs, no DFS nodes or counts."""
        if not joins and (pair_nodes or counts):
            raise ValueError("a trace without joins has no DFS nodes or counts")
        W = PACK_WINDOW  # whole pack windows: max_graphs rounded up to the window
        G = min((int(max_graphs) + W - 1) // W * W, self.D)
        dcounts = DeviceBuffer(self.q, 8 * max(G * self.Q, 1), zero=False)
        valid = DeviceBuffer(self.q, 8 * max(G * self.VW, 1), zero=False)
        domsum = DeviceBuffer(self.q, 8 * max(query_nodes, 1), zero=True)
        qpairs = DeviceBuffer(self.q, 8 * max(self.Q, 1), zero=True)
        qnodes = DeviceBuffer(self.q, 8 * max(self.Q, 1), zero=True) if joins else None
        pn = DeviceBuffer(self.q, 4 * max(G * self.Q, 1), zero=True) if pair_nodes else None
        got = C.c_uint32(0)
        check(_L().bms2_pipeline_trace(self.h, int(max_graphs), C.c_void_p(dcounts.ptr),
                                       C.c_void_p(valid.ptr), C.c_void_p(domsum.ptr),
                                       C.c_void_p(qpairs.ptr),
                                       C.c_void_p(qnodes.ptr) if qnodes else None,
                                       C.c_void_p(pn.ptr) if pn else None, C.byref(got)))
        g = int(got.value)
        out = dict(graphs=g,
                   domsum=domsum.to_numpy(np.uint64, query_nodes),
                   qpairs=qpairs.to_numpy(np.uint64, self.Q))
        if qnodes is not None:
            out["qnodes"] = qnodes.to_numpy(np.uint64, self.Q)
        if counts:
            out["counts"] = dcounts.to_numpy(np.uint64, g * self.Q).reshape(g, self.Q)
        if pn is not None:
            out["pair_nodes"] = pn.to_numpy(np.uint32, g * self.Q).reshape(g, self.Q)
        return out


def run_staged(q: Queue, query: GraphBatch, data: GraphBatch, K: int, find_first: bool = False,
               keep_domains: bool = True, use_cs: bool = False):
    """The stage sequence of BMS2.cpp:131-234 through the staged entry points.

    Returns dict(domains=[iteration 0..K], gmcr=(offsets, indices), counts=per GMCR entry,
    total=match total, signatures=...).  use_cs joins as BMS2 --use-cs does (BMS2.cpp:217-300):
    partial matches of query 0 at depth = its node count, then their extensions; the result also
    holds the partial-match records (`partial`)."""
    dq = DeviceBatchedCSRGraph(q, query)
    dd = DeviceBatchedCSRGraph(q, data)
    sig = Signature(q, dd.total_nodes, dq.total_nodes)
    path = PathSignature(q, dd.total_nodes, dq.total_nodes)
    cyc = CycleSignature(q, dd.total_nodes, dq.total_nodes)
    cand = Candidates(q, dq.total_nodes, dd.total_nodes)
    sig.generate_data_signatures(dd)
    path.generate_data_path_signatures(dd)
    cyc.generate_data_cycle_signatures(dd)
    sig.generate_query_signatures(dq)
    path.generate_query_path_signatures(dq)
    cyc.generate_query_cycle_signatures(dq)
    q.wait()
    out = dict(domains=[], signatures=dict(
        dlabel=sig.host(SCOPE_DATA), qlabel=sig.host(SCOPE_QUERY), dpath=path.host(SCOPE_DATA),
        qpath=path.host(SCOPE_QUERY), dcycle=cyc.host(SCOPE_DATA), qcycle=cyc.host(SCOPE_QUERY)))
    filter_candidates(q, dq, dd, sig, path, cyc, cand)
    if keep_domains:
        out["domains"].append(cand.host())
    for k in range(1, K + 1):
        sig.refine_data_signatures(dd, k)
        sig.refine_query_signatures(dq, k)
        refine_candidates(q, dq, dd, sig, cand)
        if keep_domains:
            out["domains"].append(cand.host())
            out["signatures"][f"dsig{k}"] = sig.host(SCOPE_DATA)
            out["signatures"][f"qsig{k}"] = sig.host(SCOPE_QUERY)
    q.wait()
    out["final_domains"] = cand.host()
    gm = GMCR(q)
    gm.generate_gmcr(dq, dd, cand)
    out["gmcr"] = gm.host()
    total = DeviceBuffer(q, 8, zero=True)
    per_pair = DeviceBuffer(q, 8 * max(gm.total, 1), zero=True)
    if use_cs and query.num_graphs:
        end = int(query.graph_offsets[1] - query.graph_offsets[0])
        part, n = join_partial_candidates(q, dq, dd, cand, end)
        join_with_partial_matches(q, dq, dd, cand, gm, end, part, n, total, per_pair, find_first)
        q.wait()
        out["partial"] = part.to_numpy(np.int32, n * (end + 1)).reshape(n, end + 1)
    else:
        join_candidates(q, dq, dd, cand, gm, total, find_first, per_pair)
    q.wait()
    out["total"] = int(total.to_numpy(np.uint64)[0])
    out["counts"] = per_pair.to_numpy(np.uint64, gm.total)
    return out


def query_classes(query: GraphBatch) -> np.ndarray:
    """rep[g] = lowest-indexed query of the batch isomorphic to query g (bms2_query_classes)."""
    rep = np.zeros(max(query.num_graphs, 1), np.uint32)
    n = C.c_uint32(0)
    check(_L().bms2_query_classes(C.byref(query.to_c()), rep.ctypes.data_as(C.c_void_p),
                                  C.byref(n)))
    return rep[:query.num_graphs]


def query_orbits(query: GraphBatch) -> np.ndarray:
    """orbit[v] = lowest local node of node v's automorphism orbit in its query
    (bms2_query_orbits), one u8 per global query node."""
    orb = np.zeros(max(query.total_nodes, 1), np.uint8)
    check(_L().bms2_query_orbits(C.byref(query.to_c()), orb.ctypes.data_as(C.c_void_p)))
    return orb[:query.total_nodes]


def query_symmetry(query: GraphBatch, order: np.ndarray | None = None) -> tuple[np.ndarray, np.ndarray]:
    """(du u8[Q], vmask u32[Q]): the symmetry breaking of bms2_query_symmetry under a matching
    order (u8 [Q, 32], None = identity)."""
    Q = query.num_graphs
    du = np.zeros(max(Q, 1), np.uint8)
    vm = np.zeros(max(Q, 1), np.uint32)
    o = None if order is None else np.ascontiguousarray(order, np.uint8)
    check(_L().bms2_query_symmetry(C.byref(query.to_c()),
                                   None if o is None else o.ctypes.data_as(C.c_void_p),
                                   du.ctypes.data_as(C.c_void_p), vm.ctypes.data_as(C.c_void_p)))
    return du[:Q], vm[:Q]


def run_pipeline(q: Queue, query: GraphBatch, data: GraphBatch, K: int, find_first: bool = False,
                 dedup: bool = True):
    """Fused pipeline: (counts [D][Q] u64, valid bit rows [D][ceil(Q/64)], stats[8])."""
    dq = DeviceBatchedCSRGraph(q, query)
    dd = DeviceBatchedCSRGraph(q, data)
    p = Pipeline(q, dq, dd, K, find_first)
    p.set_dedup(dedup)
    counts, valid, stats = p.alloc_outputs()
    p.run(counts, valid, stats)
    q.wait()
    c, v = p.results(counts, valid)
    return c, v, stats.to_numpy(np.uint64, 8)


__all__ = [
    "Bms2Error", "Queue", "DeviceBuffer", "DeviceBatchedCSRGraph", "create_device_csr_graph",
    "Signature", "PathSignature", "CycleSignature", "Candidates", "filter_candidates",
    "refine_candidates", "GMCR", "join_candidates", "Pipeline", "run_staged", "run_pipeline",
    "query_classes", "query_orbits", "query_symmetry",
    "SCOPE_DATA", "SCOPE_QUERY",
]
