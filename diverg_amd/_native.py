"""ctypes declarations of the C ABI (include/diverg_amd.h, diverg_amd_gen.h,
diverg_amd_graph.h)."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdiverg_amd.so")

DV_HOST, DV_DEVICE = 0, 1

STATUS = {
    0: "ok", 1: "invalid argument", 2: "overflow", 3: "out of memory",
    4: "device error", 5: "HBM budget exceeded", 6: "index out of range",
    7: "busy", 8: "not found", 9: "I/O error",
}

DV_ORDERS = {"native": 0, "topological": 1, "min-breaks": 2, "cuthill-mckee": 3, "random": 4}


class dv_rcrs(C.Structure):
    _fields_ = [
        ("nrows", C.c_uint64),
        ("ncols", C.c_uint64),
        ("rowmap", C.c_void_p),
        ("entries", C.POINTER(C.c_uint32)),
        ("n_entries", C.c_uint64),
        ("nnz", C.c_uint64),
        ("size_bytes", C.c_uint32),
        ("location", C.c_uint32),
        ("owner", C.c_void_p),
    ]


class dv_build_stats(C.Structure):
    _fields_ = [
        ("build_ms", C.c_double),
        ("h2d_ms", C.c_double),
        ("d2h_ms", C.c_double),
        ("alg_bytes", C.c_uint64),
        ("peak_device_bytes", C.c_uint64),
        ("n_spgemm", C.c_uint32),
        ("n_spadd", C.c_uint32),
        ("overflow_rows", C.c_uint64),
        ("deferred_rows", C.c_uint64),
        ("planned", C.c_uint32),
        ("retried", C.c_uint32),
        ("layout", C.c_uint32),
        ("builds", C.c_uint32),
    ]


class dv_graph_info(C.Structure):
    _fields_ = [
        ("nodes", C.c_uint64),
        ("edges", C.c_uint64),
        ("loci", C.c_uint64),
        ("components", C.c_uint64),
        ("breaks", C.c_uint64),
        ("bandwidth", C.c_uint64),
    ]


class dv_tuning(C.Structure):
    _fields_ = [
        ("lds_words", C.c_uint32),
        ("stage_entries", C.c_uint32),
        ("rows_per_group", C.c_uint32),
        ("lanes_small", C.c_uint32),
        ("small_slice_max", C.c_uint32),
        ("lane_slice_max", C.c_uint32),
        ("autotune", C.c_uint32),
        ("plan_mode", C.c_uint32),
        ("graphs", C.c_uint32),
        ("summary_span", C.c_uint32),
        ("schedule", C.c_uint32),
        ("layout", C.c_uint32),
        ("chain_span", C.c_uint32),
    ]


class dv_kernel_time(C.Structure):
    _fields_ = [
        ("launches", C.c_uint64),
        ("total_ms", C.c_double),
        ("alg_bytes", C.c_uint64),
        ("slice_bytes", C.c_uint64),
    ]


KERNEL_KINDS = ["spgemm_lane", "spgemm_write", "spgemm_groups_scan", "spadd"]


class dv_graph_params(C.Structure):
    _fields_ = [
        ("backbone_bp", C.c_uint64),
        ("seed", C.c_uint64),
        ("node_min", C.c_uint32),
        ("node_max", C.c_uint32),
        ("snp_rate", C.c_double),
        ("indel_rate", C.c_double),
        ("indel_geometric", C.c_uint32),
        ("indel_min", C.c_uint32),
        ("indel_max", C.c_uint32),
        ("indel_mean", C.c_double),
        ("sv_rate", C.c_double),
        ("sv_min", C.c_uint32),
        ("sv_max", C.c_uint32),
        ("sv_inversion_frac", C.c_double),
        ("sv_deletion_frac", C.c_double),
    ]


class dv_server_opts(C.Structure):
    _fields_ = [
        ("hbm_budget_bytes", C.c_uint64),
        ("max_batch", C.c_uint32),
        ("max_inflight", C.c_uint32),
        ("query_cus", C.c_uint32),
        ("pad", C.c_uint32),
    ]


class dv_request_info(C.Structure):
    _fields_ = [
        ("kind", C.c_uint32),
        ("status", C.c_int32),
        ("index_id", C.c_uint32),
        ("pad", C.c_uint32),
        ("n", C.c_uint64),
        ("submit_us", C.c_double),
        ("start_us", C.c_double),
        ("done_us", C.c_double),
    ]


P = C.POINTER
VP = C.c_void_p
U32P = C.POINTER(C.c_uint32)
DP = C.POINTER(C.c_double)

# (name, restype, argtypes) of every function include/*.h declares
SIGNATURES = [
    ("dv_abi_version", C.c_int, []),
    ("dv_status_string", C.c_char_p, [C.c_int]),
    ("dv_init", C.c_int, [C.c_int, P(VP)]),
    ("dv_finalize", C.c_int, [VP]),
    ("dv_synchronize", C.c_int, [VP]),
    ("dv_range_spadd_noalloc_max_rows", C.c_uint64, []),
    ("dv_range_spadd_noalloc", C.c_int,
     [VP, P(dv_rcrs), P(dv_rcrs), VP, U32P, C.c_uint64, P(C.c_uint64), P(C.c_uint64), VP]),
    ("dv_last_error", C.c_char_p, [VP]),
    ("dv_stream", VP, [VP]),
    ("dv_get_tuning", C.c_int, [VP, P(dv_tuning)]),
    ("dv_last_tuning", C.c_int, [VP, P(dv_tuning)]),
    ("dv_set_tuning", C.c_int, [VP, P(dv_tuning)]),
    ("dv_profile", C.c_int, [VP, C.c_int]),
    ("dv_profile_read", C.c_int, [VP, P(dv_kernel_time)]),
    ("dv_copy", C.c_int, [VP, P(dv_rcrs), C.c_uint32, P(dv_rcrs)]),
    ("dv_free", C.c_int, [VP, P(dv_rcrs)]),
    ("dv_range_identity", C.c_int, [VP, C.c_uint64, C.c_uint32, C.c_uint32, P(dv_rcrs)]),
    ("dv_range_spadd", C.c_int, [VP, P(dv_rcrs), P(dv_rcrs), C.c_uint32, P(dv_rcrs)]),
    ("dv_range_spgemm", C.c_int, [VP, P(dv_rcrs), P(dv_rcrs), C.c_uint32, P(dv_rcrs)]),
    ("dv_range_power", C.c_int, [VP, P(dv_rcrs), C.c_uint32, C.c_uint32, P(dv_rcrs)]),
    ("dv_create_distance_index", C.c_int,
     [VP, P(dv_rcrs), C.c_uint32, C.c_uint32, C.c_uint32, P(dv_rcrs), P(dv_build_stats)]),
    ("dv_create_distance_index_batched", C.c_int,
     [VP, P(dv_rcrs), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, P(dv_rcrs),
      P(dv_build_stats)]),
    ("dv_query_batch", C.c_int,
     [VP, P(dv_rcrs), P(C.c_uint32), P(C.c_uint32), P(C.c_uint8), C.c_uint64, C.c_uint32]),
    ("dv_server_start", C.c_int, [C.c_int, P(dv_server_opts), P(VP)]),
    ("dv_server_stop", C.c_int, [VP]),
    ("dv_server_add_index", C.c_int, [VP, P(dv_rcrs), U32P]),
    ("dv_server_submit_queries", C.c_int,
     [VP, C.c_uint32, U32P, U32P, C.c_uint64, P(C.c_uint64)]),
    ("dv_server_submit_build", C.c_int,
     [VP, P(dv_rcrs), C.c_uint32, C.c_uint32, P(C.c_uint64)]),
    ("dv_server_poll", C.c_int, [VP, C.c_uint64, P(dv_request_info)]),
    ("dv_server_wait", C.c_int, [VP, C.c_uint64, P(dv_request_info)]),
    ("dv_server_results", C.c_int, [VP, C.c_uint64, P(C.c_uint8), C.c_uint64]),
    ("dv_server_get_index", C.c_int, [VP, C.c_uint32, P(dv_rcrs)]),
    ("dv_server_replay", C.c_int,
     [VP, C.c_uint64, U32P, U32P, U32P, DP, P(C.c_uint8), DP, C.c_uint32, P(dv_rcrs),
      U32P, U32P, DP, DP, U32P, DP]),
    ("dv_graph_preset", C.c_int, [C.c_uint32, C.c_uint32, P(dv_graph_params)]),
    ("dv_generate_graph", C.c_int,
     [P(dv_graph_params), C.c_uint32, P(dv_rcrs), P(P(C.c_uint32))]),
    ("dv_free_host", None, [VP]),
    ("dv_merge_distance_index", C.c_int, [VP, P(dv_rcrs), P(dv_rcrs), C.c_uint32, P(dv_rcrs)]),
    ("dv_save", C.c_int, [VP, P(dv_rcrs), C.c_char_p]),
    ("dv_load", C.c_int, [VP, C.c_char_p, C.c_uint32, C.c_uint32, P(dv_rcrs)]),
    ("dv_graph_load_gfa", C.c_int, [C.c_char_p, C.c_uint32, P(VP)]),
    ("dv_graph_from_arrays", C.c_int,
     [C.c_uint64, U32P, C.c_uint64, U32P, U32P, P(VP)]),
    ("dv_graph_free", None, [VP]),
    ("dv_graph_info_get", C.c_int, [VP, P(dv_graph_info)]),
    ("dv_graph_arrays", C.c_int, [VP, U32P, P(C.c_uint64), P(C.c_uint64), U32P]),
    ("dv_graph_node_name", C.c_char_p, [VP, C.c_uint64]),
    ("dv_graph_sort", C.c_int, [VP, C.c_uint32, C.c_uint64]),
    ("dv_graph_components", C.c_int, [VP, P(C.c_uint64), C.c_uint64, P(C.c_uint64)]),
    ("dv_adjacency_matrix", C.c_int,
     [VP, VP, C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, P(dv_rcrs)]),
]

_lib = None


class NativeMissing(ImportError):
    pass


def lib():
    """Load libdiverg_amd.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeMissing(
                f"{LIB_PATH} is missing: build it with `python -m diverg_amd.build` "
                "(or __graft_entry__.build()); there is no fallback path")
        l = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            try:
                f = getattr(l, name)
            except AttributeError:
                continue
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib
