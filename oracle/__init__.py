"""ctypes wrapper of liboracle.so, the CPU restatement of the reference path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's CPU
baseline leg, as the checker.  The product package (bms2_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liboracle.so"

u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)


class OrcCsr(C.Structure):
    _fields_ = [
        ("num_graphs", C.c_uint32),
        ("reserved", C.c_uint32),
        ("total_nodes", C.c_uint64),
        ("total_edges", C.c_uint64),
        ("graph_offsets", u32p),
        ("row_offsets", u32p),
        ("column_indices", u32p),
        ("node_labels", u8p),
        ("edge_labels", u8p),
    ]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is not built (make -C oracle)")
        h = C.CDLL(str(LIB_PATH))
        P = C.POINTER(OrcCsr)
        vp = C.c_void_p
        h.orc_hash32.restype = C.c_uint32
        h.orc_hash32.argtypes = [C.c_uint32]
        h.orc_combine_hash.restype = C.c_uint32
        h.orc_combine_hash.argtypes = [C.c_uint32, C.c_uint32]
        for name in ("orc_label_signatures", "orc_path_signatures", "orc_cycle_signatures"):
            getattr(h, name).restype = None
            getattr(h, name).argtypes = [P, vp]
        h.orc_refine_signatures.restype = None
        h.orc_refine_signatures.argtypes = [P, C.c_uint32, vp, vp]
        h.orc_filter.restype = None
        h.orc_filter.argtypes = [P, P, vp, vp, vp, vp, vp, C.c_uint64]
        h.orc_refine_candidates.restype = None
        h.orc_refine_candidates.argtypes = [P, P, vp, vp, vp, C.c_uint64]
        h.orc_gmcr.restype = C.c_uint64
        h.orc_gmcr.argtypes = [P, P, vp, C.c_uint64, vp, vp, C.c_uint64]
        h.orc_join.restype = C.c_uint64
        h.orc_join.argtypes = [P, P, vp, C.c_uint64, vp, vp, C.c_int, vp, vp, vp]
        h.orc_join_sym.restype = C.c_uint64
        h.orc_join_sym.argtypes = [P, P, vp, C.c_uint64, vp, vp, C.c_int, vp, vp, vp, vp, vp]
        h.orc_join_partial.restype = C.c_uint64
        h.orc_join_partial.argtypes = [P, P, vp, C.c_uint64, C.c_uint32, vp, C.c_uint64]
        h.orc_join_with_partial.restype = C.c_uint64
        h.orc_join_with_partial.argtypes = [P, P, vp, C.c_uint64, vp, vp, C.c_uint32, vp,
                                            C.c_uint64, C.c_int, vp]
        h.orc_pipeline.restype = C.c_int
        h.orc_pipeline.argtypes = [P, P, C.c_uint32, C.c_int, C.c_uint32, C.c_uint32, vp, vp, vp,
                                   C.c_int]
        _lib = h
    return _lib


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def csr(g) -> OrcCsr:
    """OrcCsr view of a GraphBatch-like object (keep g alive while in use)."""
    c = OrcCsr()
    c.num_graphs = g.num_graphs
    c.total_nodes = g.total_nodes
    c.total_edges = g.total_edges
    c.graph_offsets = g.graph_offsets.ctypes.data_as(u32p)
    c.row_offsets = g.row_offsets.ctypes.data_as(u32p)
    c.column_indices = g.column_indices.ctypes.data_as(u32p)
    c.node_labels = g.node_labels.ctypes.data_as(u8p)
    c.edge_labels = g.edge_labels.ctypes.data_as(u8p)
    return c


def hash32(h: int) -> int:
    return int(lib().orc_hash32(h))


def combine_hash(a: int, b: int) -> int:
    return int(lib().orc_combine_hash(a, b))


def label_signatures(g) -> np.ndarray:
    out = np.zeros(g.total_nodes, np.uint64)
    c = csr(g)
    lib().orc_label_signatures(C.byref(c), _ptr(out))
    return out


def path_signatures(g) -> np.ndarray:
    out = np.zeros((g.total_nodes, 4), np.uint64)
    c = csr(g)
    lib().orc_path_signatures(C.byref(c), _ptr(out))
    return out


def cycle_signatures(g) -> np.ndarray:
    out = np.zeros(g.total_nodes, np.uint64)
    c = csr(g)
    lib().orc_cycle_signatures(C.byref(c), _ptr(out))
    return out


class RefineState:
    """Signature<PowerGraph> state of one scope: signatures and reachability (zeroed)."""

    def __init__(self, g):
        self.g = g
        self.sig = label_signatures(g)
        self.reach = np.zeros((g.total_nodes, 4), np.uint64)

    def refine(self, view_size: int) -> np.ndarray:
        c = csr(self.g)
        lib().orc_refine_signatures(C.byref(c), view_size, _ptr(self.sig), _ptr(self.reach))
        return self.sig


def words_per_row(data) -> int:
    return (data.total_nodes + 63) // 64


def filter_candidates(q, d, qpath, dpath, qcyc, dcyc) -> np.ndarray:
    W = words_per_row(d)
    cand = np.zeros((q.total_nodes, W), np.uint64)
    cq, cd = csr(q), csr(d)
    lib().orc_filter(C.byref(cq), C.byref(cd), _ptr(qpath), _ptr(dpath), _ptr(qcyc), _ptr(dcyc),
                     _ptr(cand), W)
    return cand


def refine_candidates(q, d, qsig, dsig, cand) -> np.ndarray:
    cq, cd = csr(q), csr(d)
    lib().orc_refine_candidates(C.byref(cq), C.byref(cd), _ptr(qsig), _ptr(dsig), _ptr(cand),
                                cand.shape[1])
    return cand


def gmcr(q, d, cand) -> tuple[np.ndarray, np.ndarray]:
    cq, cd = csr(q), csr(d)
    offsets = np.zeros(d.num_graphs + 1, np.uint32)
    total = lib().orc_gmcr(C.byref(cq), C.byref(cd), _ptr(cand), cand.shape[1], _ptr(offsets),
                           None, 0)
    indices = np.zeros(max(total, 1), np.uint32)
    lib().orc_gmcr(C.byref(cq), C.byref(cd), _ptr(cand), cand.shape[1], _ptr(offsets),
                   _ptr(indices), total)
    return offsets, indices[:total]


def join(q, d, cand, offsets, indices, find_first=False, order=None, sym=None):
    """(total, per-pair counts, per-pair DFS nodes).  sym = (du u8[Q], vmask u32[Q]): the
    symmetry-breaking constraint per query (counts are then the constrained counts)."""
    cq, cd = csr(q), csr(d)
    T = len(indices)
    counts = np.zeros(max(T, 1), np.uint64)
    nodes = np.zeros(max(T, 1), np.uint64)
    idx = np.ascontiguousarray(indices, np.uint32)
    ordp = None if order is None else np.ascontiguousarray(order, np.uint32)
    if sym is None:
        total = lib().orc_join(C.byref(cq), C.byref(cd), _ptr(cand), cand.shape[1], _ptr(offsets),
                               _ptr(idx), int(find_first), _ptr(ordp), _ptr(counts), _ptr(nodes))
    else:
        du = np.ascontiguousarray(sym[0], np.uint8)
        vm = np.ascontiguousarray(sym[1], np.uint32)
        total = lib().orc_join_sym(C.byref(cq), C.byref(cd), _ptr(cand), cand.shape[1],
                                   _ptr(offsets), _ptr(idx), int(find_first), _ptr(ordp),
                                   _ptr(du), _ptr(vm), _ptr(counts), _ptr(nodes))
    return int(total), counts[:T], nodes[:T]


def join_partial(q, d, cand, end: int) -> np.ndarray:
    """joinPartialCandidates2: every embedding of query 0's first `end` nodes, as rows of
    end + 1 ints {mapped global data node ids, data graph id} (data graph order, DFS order)."""
    cq, cd = csr(q), csr(d)
    n = lib().orc_join_partial(C.byref(cq), C.byref(cd), _ptr(cand), cand.shape[1], end, None, 0)
    out = np.zeros((max(n, 1), end + 1), np.int32)
    lib().orc_join_partial(C.byref(cq), C.byref(cd), _ptr(cand), cand.shape[1], end, _ptr(out), n)
    return out[:n]


def join_with_partial(q, d, cand, offsets, indices, end: int, partial: np.ndarray,
                      find_first: bool = False) -> tuple[int, np.ndarray]:
    """joinWithPartialMatches: (total, per-GMCR-entry counts)."""
    cq, cd = csr(q), csr(d)
    T = len(indices)
    counts = np.zeros(max(T, 1), np.uint64)
    part = np.ascontiguousarray(partial, np.int32)
    idx = np.ascontiguousarray(indices, np.uint32)
    total = lib().orc_join_with_partial(C.byref(cq), C.byref(cd), _ptr(cand), cand.shape[1],
                                        _ptr(offsets), _ptr(idx), end, _ptr(part), len(part),
                                        int(find_first), _ptr(counts))
    return int(total), counts[:T]


class StagedResult:
    """Outputs of the staged restatement: domains after each refinement step, GMCR, counts."""

    def __init__(self):
        self.domains: list[np.ndarray] = []
        self.gmcr_offsets = None
        self.gmcr_indices = None
        self.counts = None
        self.dfs_nodes = None
        self.total = 0
        self.signatures = {}


def staged(q, d, K: int, find_first: bool = False, keep_domains: bool = True,
           order=None) -> StagedResult:
    """The stage sequence of BMS2.cpp:131-234 through the oracle's staged functions."""
    r = StagedResult()
    dpath, qpath = path_signatures(d), path_signatures(q)
    dcyc, qcyc = cycle_signatures(d), cycle_signatures(q)
    dstate, qstate = RefineState(d), RefineState(q)
    r.signatures = dict(dlabel=dstate.sig.copy(), qlabel=qstate.sig.copy(), dpath=dpath,
                        qpath=qpath, dcycle=dcyc, qcycle=qcyc)
    cand = filter_candidates(q, d, qpath, dpath, qcyc, dcyc)
    if keep_domains:
        r.domains.append(cand.copy())
    for k in range(1, K + 1):
        dsig = dstate.refine(k)
        qsig = qstate.refine(k)
        r.signatures[f"dsig{k}"] = dsig.copy()
        r.signatures[f"qsig{k}"] = qsig.copy()
        refine_candidates(q, d, qsig, dsig, cand)
        if keep_domains:
            r.domains.append(cand.copy())
    r.final_domains = cand
    r.gmcr_offsets, r.gmcr_indices = gmcr(q, d, cand)
    r.total, r.counts, r.dfs_nodes = join(q, d, cand, r.gmcr_offsets, r.gmcr_indices,
                                          find_first, order)
    return r


def pipeline(q, d, K: int, find_first: bool = False, g_begin: int = 0, g_end: int | None = None,
             threads: int = 0):
    """Fused CPU pipeline: (counts [G][Q] u64, valid bit rows [G][ceil(Q/64)], stats[4])."""
    g_end = d.num_graphs if g_end is None else g_end
    G, Q = g_end - g_begin, q.num_graphs
    counts = np.zeros((G, Q), np.uint64)
    valid = np.zeros((G, (Q + 63) // 64), np.uint64)
    stats = np.zeros(4, np.uint64)
    cq, cd = csr(q), csr(d)
    rc = lib().orc_pipeline(C.byref(cq), C.byref(cd), K, int(find_first), g_begin, g_end,
                            _ptr(counts), _ptr(valid), _ptr(stats), threads)
    if rc != 0:
        raise RuntimeError(f"orc_pipeline failed with {rc}")
    return counts, valid, stats
