"""ctypes binding of libbms2.so (the C ABI declared in include/bms2.h).

The library is the product: if it is missing or lacks a symbol this module raises instead of
falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libbms2.so"

OK = 0
ERR_INVALID = 1  # BMS2_ERR_INVALID
ERR_UNSUPPORTED = 2  # BMS2_ERR_UNSUPPORTED
ERR_CAPACITY = 7  # BMS2_ERR_CAPACITY: caller buffer too small
SCOPE_DATA = 0
SCOPE_QUERY = 1
WILDCARD_NODE = 255
MAX_QUERY_NODES = 32
MAX_GRAPH_NODES = 256

u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)


class HostCsr(C.Structure):
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


class DeviceCsr(C.Structure):
    _fields_ = [
        ("graph_offsets", C.c_void_p),
        ("row_offsets", C.c_void_p),
        ("column_indices", C.c_void_p),
        ("node_labels", C.c_void_p),
        ("edge_labels", C.c_void_p),
        ("num_graphs", C.c_uint32),
        ("max_graph_nodes", C.c_uint32),
        ("total_nodes", C.c_uint64),
        ("total_edges", C.c_uint64),
        ("max_degree", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


class WorkloadSpec(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("num_data_graphs", C.c_uint32),
        ("num_query_graphs", C.c_uint32),
        ("data_size_dist", C.c_uint32),
        ("data_min_nodes", C.c_uint32),
        ("data_max_nodes", C.c_uint32),
        ("polycyclic_permille", C.c_uint32),
        ("data_lognormal_mu", C.c_double),
        ("data_lognormal_sigma", C.c_double),
        ("poly_min_rings", C.c_uint32),
        ("poly_max_rings", C.c_uint32),
        ("query_min_nodes", C.c_uint32),
        ("query_max_nodes", C.c_uint32),
        ("extracted_permille", C.c_uint32),
        ("query_shapes", C.c_uint32),
        ("max_ring_closures", C.c_uint32),
        ("num_threads", C.c_uint32),
    ]


class PipelineOptions(C.Structure):
    _fields_ = [
        ("refinement_steps", C.c_uint32),
        ("find_first", C.c_uint32),
        ("block_waves", C.c_uint32),
        ("join_waves", C.c_uint32),
        ("no_graph", C.c_uint32),
        ("reserved", C.c_uint32 * 3),
    ]


class Tuning(C.Structure):
    _fields_ = [("small_waves", C.c_uint32), ("large_waves", C.c_uint32), ("slots", C.c_uint32),
                ("candidates", C.c_uint32), ("best_ms", C.c_float), ("big_waves", C.c_uint32),
                ("big_block_waves", C.c_uint32), ("big_row_words", C.c_uint32)]


class Adam(C.Structure):
    _fields_ = [("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float)]


vp = C.c_void_p
vpp = C.POINTER(C.c_void_p)
i32 = C.c_int
u32 = C.c_uint32
u64 = C.c_uint64
HCP = C.POINTER(HostCsr)
DCP = C.POINTER(DeviceCsr)

# name -> (restype, argtypes); mirrors include/bms2.h
PROTOTYPES: dict[str, tuple] = {
    "bms2_last_error": (C.c_char_p, []),
    "bms2_version": (i32, []),
    "bms2_host_csr_free": (None, [HCP]),
    "bms2_load_dat": (i32, [C.c_char_p, u64, i32, HCP]),
    "bms2_parse_dat_lines": (i32, [C.POINTER(C.c_char_p), u64, i32, HCP]),
    "bms2_write_dat": (i32, [C.c_char_p, HCP]),
    "bms2_generate_workload": (i32, [C.POINTER(WorkloadSpec), HCP, HCP]),
    "bms2_queue_create": (i32, [i32, vpp]),
    "bms2_queue_wrap": (i32, [i32, vp, vpp]),
    "bms2_queue_destroy": (i32, [vp]),
    "bms2_queue_wait": (i32, [vp]),
    "bms2_queue_stream": (vp, [vp]),
    "bms2_event_elapsed_ms": (i32, [vp, C.POINTER(C.c_float)]),
    "bms2_event_destroy": (i32, [vp]),
    "bms2_device_malloc": (i32, [vp, u64, vpp]),
    "bms2_device_free": (i32, [vp, vp]),
    "bms2_memcpy_h2d": (i32, [vp, vp, vp, u64]),
    "bms2_memcpy_d2h": (i32, [vp, vp, vp, u64]),
    "bms2_memset_zero": (i32, [vp, vp, u64]),
    "bms2_device_csr_create": (i32, [vp, HCP, DCP]),
    "bms2_device_csr_pack": (i32, [vp, u32, vp, vp, vp, vp, vp, u64, u64, DCP]),
    "bms2_device_csr_destroy": (i32, [vp, DCP]),
    "bms2_device_csr_alloc_size": (u64, [DCP]),
    "bms2_signature_create": (i32, [vp, u64, u64, vpp]),
    "bms2_signature_destroy": (i32, [vp]),
    "bms2_signature_generate": (i32, [vp, DCP, i32, vpp]),
    "bms2_signature_refine": (i32, [vp, DCP, u32, i32, vpp]),
    "bms2_signature_copy_to_host": (i32, [vp, i32, vp]),
    "bms2_signature_device_ptr": (vp, [vp, i32]),
    "bms2_path_signature_create": (i32, [vp, u64, u64, vpp]),
    "bms2_path_signature_destroy": (i32, [vp]),
    "bms2_path_signature_generate": (i32, [vp, DCP, i32, vpp]),
    "bms2_path_signature_copy_to_host": (i32, [vp, i32, vp]),
    "bms2_cycle_signature_create": (i32, [vp, u64, u64, vpp]),
    "bms2_cycle_signature_destroy": (i32, [vp]),
    "bms2_cycle_signature_generate": (i32, [vp, DCP, i32, vpp]),
    "bms2_cycle_signature_copy_to_host": (i32, [vp, i32, vp]),
    "bms2_candidates_create": (i32, [vp, u64, u64, vpp]),
    "bms2_candidates_destroy": (i32, [vp]),
    "bms2_candidates_words_per_row": (u64, [vp]),
    "bms2_candidates_copy_to_host": (i32, [vp, vp]),
    "bms2_candidates_row_counts": (i32, [vp, vp]),
    "bms2_candidates_copy_rows": (i32, [vp, u64, u64, vp]),
    "bms2_signature_alloc_size": (u64, [vp, i32]),
    "bms2_path_signature_alloc_size": (u64, [vp, i32]),
    "bms2_cycle_signature_alloc_size": (u64, [vp, i32]),
    "bms2_queue_device_info": (i32, [vp, C.POINTER(C.c_uint64), C.POINTER(u32), C.c_char_p, u64]),
    "bms2_shared_malloc": (i32, [vp, u64, vpp]),
    "bms2_shared_free": (i32, [vp, vp]),
    "bms2_filter_candidates": (i32, [vp, DCP, DCP, vp, vp, vp, vp, vpp]),
    "bms2_refine_candidates": (i32, [vp, DCP, DCP, vp, vp, vpp]),
    "bms2_gmcr_create": (i32, [vp, vpp]),
    "bms2_gmcr_destroy": (i32, [vp]),
    "bms2_gmcr_generate": (i32, [vp, DCP, DCP, vp, vpp]),
    "bms2_gmcr_total": (u64, [vp]),
    "bms2_gmcr_device_ptrs": (i32, [vp, vpp, vpp]),
    "bms2_gmcr_copy_to_host": (i32, [vp, vp, vp]),
    "bms2_join_candidates": (i32, [vp, DCP, DCP, vp, vp, vp, i32, vp, u64, vpp]),
    "bms2_join_candidates_noalloc": (i32, [vp, DCP, DCP, vp, vp, vp, i32, vp, u64]),
    "bms2_join_partial_candidates": (i32, [vp, DCP, DCP, vp, u32, vp, u64, C.POINTER(C.c_uint64),
                                           vpp]),
    "bms2_join_with_partial_matches": (i32, [vp, DCP, DCP, vp, vp, u32, vp, u64, vp, i32, vp,
                                             u64, vpp]),
    "bms2_pipeline_create": (i32, [vp, DCP, DCP, C.POINTER(PipelineOptions), vpp]),
    "bms2_pipeline_destroy": (i32, [vp]),
    "bms2_pipeline_run": (i32, [vp, vp, vp, vp, vpp, vpp]),
    "bms2_pipeline_set_order": (i32, [vp, vp]),
    "bms2_pipeline_set_dedup": (i32, [vp, i32, C.POINTER(C.c_uint32)]),
    "bms2_query_classes": (i32, [HCP, vp, C.POINTER(C.c_uint32)]),
    "bms2_query_orbits": (i32, [HCP, vp]),
    "bms2_query_symmetry": (i32, [HCP, vp, vp, vp]),
    "bms2_pipeline_autotune": (i32, [vp, vp, vp, u32, C.POINTER(Tuning)]),
    "bms2_pipeline_set_queries": (i32, [vp, DCP]),
    "bms2_pipeline_query_totals": (i32, [vp, vp, vp, vp]),
    "bms2_query_static_features": (i32, [HCP, vp]),
    "bms2_score_orders": (i32, [HCP, vp, vp, vp]),
    "bms2_mlp_num_params": (u64, []),
    "bms2_mlp_create": (i32, [vp, u64, vpp]),
    "bms2_mlp_destroy": (i32, [vp]),
    "bms2_mlp_get_params": (i32, [vp, vp, u64]),
    "bms2_mlp_set_params": (i32, [vp, vp, u64]),
    "bms2_mlp_forward": (i32, [vp, vp, u32, vp, i32]),
    "bms2_mlp_loss_grad": (i32, [vp, vp, vp, u32, C.c_float, C.POINTER(C.c_float), vp]),
    "bms2_mlp_train_step": (i32, [vp, vp, vp, u32, C.c_float, C.POINTER(Adam),
                                  C.POINTER(C.c_float)]),
    "bms2_pipeline_trace": (i32, [vp, u32, vp, vp, vp, vp, vp, vp, C.POINTER(u32)]),
}


class Bms2Error(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"bms2 error {code}: {message}")
        self.code = code


_lib: C.CDLL | None = None


def lib() -> C.CDLL:
    """Loads libbms2.so and binds every prototype; raises if the library or a symbol is missing."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(handle, name)  # AttributeError names a missing export
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(rc: int) -> None:
    if rc != OK:
        msg = lib().bms2_last_error()
        raise Bms2Error(rc, msg.decode() if msg else "")
