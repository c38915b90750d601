"""ctypes mirror of include/gpuqo_amd.h (the C ABI of the optimizer core).

The structures below must stay byte-identical to the header; tests check the
sizes against the C library's own sizeof values.
"""
import ctypes as C

GQ_MAX_RELS = 256
GQ_MAX_WORDS = 4
GQ_MAX_DP_RELS = 32

GQ_OK = 0
GQ_ERR_INVALID = 1
GQ_ERR_UNSUPPORTED = 2
GQ_ERR_NOMEM = 3
GQ_ERR_DEVICE = 4
GQ_ERR_CAPACITY = 5
GQ_ERR_NO_PLAN = 6
GQ_ERR_INTERRUPTED = 7
GQ_ERR_BUSY = 8
GQ_ERR_INTERNAL = 9

# gpuqo_algorithm (src/include/optimizer/gpuqo_common.h:13-29)
ALGO = {
    "dpsize": 0, "dpsub": 1, "cpu_dpsize": 2, "cpu_dpsub": 3,
    "parallel_cpu_dpsub": 4, "cpu_dpsub_bicc": 5, "parallel_cpu_dpsub_bicc": 6,
    "cpu_dpccp": 7, "dpe_dpsize": 8, "dpe_dpsub": 9, "dpe_dpccp": 10,
    "cpu_goo": 11,
}
IDP = {"idp1": 0, "idp2": 1, "union": 2}
MEMO = {"auto": 0, "dense": 1, "hash": 2, "rank": 3}
COST_FUNCTION = {"postgres": 0, "simple": 1, "cout": 2}


class PathCost(C.Structure):
    _fields_ = [("startup", C.c_float), ("total", C.c_float)]


class BaseRel(C.Structure):
    _fields_ = [("rows", C.c_float), ("tuples", C.c_float), ("cost", PathCost),
                ("width", C.c_int32)]


class VarInfo(C.Structure):
    _fields_ = [("stadistinct", C.c_float), ("stanullfrac", C.c_float),
                ("mcvfreq", C.c_float), ("index_available", C.c_uint8),
                ("index_unique", C.c_uint8), ("pad_", C.c_uint8 * 2),
                ("index_rows", C.c_float), ("index_cost", PathCost)]


class Query(C.Structure):
    _fields_ = [("n_rels", C.c_int32), ("n_words", C.c_int32),
                ("base_rels", C.POINTER(BaseRel)),
                ("edge_table", C.POINTER(C.c_uint64)),
                ("n_eq_classes", C.c_int32),
                ("ec_member_offsets", C.POINTER(C.c_int32)),
                ("ec_members", C.POINTER(C.c_int32)),
                ("ec_sels", C.POINTER(C.c_float)),
                ("ec_fks", C.POINTER(C.c_uint64)),
                ("ec_vars", C.POINTER(VarInfo))]


class Config(C.Structure):
    _fields_ = [("gpuqo_algorithm", C.c_int32),
                ("gpuqo_dpsub_bicc", C.c_int32),
                ("gpuqo_dpsub_filter", C.c_int32),
                ("gpuqo_dpsub_filter_threshold", C.c_int32),
                ("gpuqo_dpsub_filter_cpu_enum_threshold", C.c_int32),
                ("gpuqo_dpsub_filter_keys_overprovisioning", C.c_int32),
                ("gpuqo_dpsub_ccc", C.c_int32),
                ("gpuqo_dpsub_csg", C.c_int32),
                ("gpuqo_dpsub_csg_threshold", C.c_int32),
                ("gpuqo_dpsub_tree", C.c_int32),
                ("gpuqo_spanning_tree", C.c_int32),
                ("gpuqo_n_parallel", C.c_int32),
                ("gpuqo_scratchpad_size_mb", C.c_int32),
                ("gpuqo_min_memo_size_mb", C.c_int32),
                ("gpuqo_max_memo_size_mb", C.c_int32),
                ("gpuqo_idp_type", C.c_int32),
                ("gpuqo_idp_n_iters", C.c_int32),
                ("gpuqo_dpe_n_threads", C.c_int32),
                ("gpuqo_cpu_dpsub_parallel_chunk_size", C.c_int32),
                ("gpuqo_dpe_pairs_per_depbuf", C.c_int32),
                ("seq_page_cost", C.c_double),
                ("random_page_cost", C.c_double),
                ("cpu_tuple_cost", C.c_double),
                ("cpu_index_tuple_cost", C.c_double),
                ("cpu_operator_cost", C.c_double),
                ("disable_cost", C.c_double),
                ("effective_cache_size", C.c_int32),
                ("work_mem", C.c_int32),
                ("enable_seqscan", C.c_uint8), ("enable_indexscan", C.c_uint8),
                ("enable_tidscan", C.c_uint8), ("enable_sort", C.c_uint8),
                ("enable_hashagg", C.c_uint8), ("enable_nestloop", C.c_uint8),
                ("enable_mergejoin", C.c_uint8), ("enable_hashjoin", C.c_uint8),
                ("memo_kind", C.c_int32),
                ("device", C.c_int32),
                ("profile_kernels", C.c_int32),
                ("cost_function", C.c_int32)]


class PlanNode(C.Structure):
    _fields_ = [("relids", C.c_uint64 * GQ_MAX_WORDS), ("rows", C.c_float),
                ("cost", PathCost), ("width", C.c_int32), ("left", C.c_int32),
                ("right", C.c_int32)]


class PlanStats(C.Structure):
    _fields_ = [("candidates", C.c_uint64), ("sets", C.c_uint64),
                ("pairs", C.c_uint64), ("memo_bytes", C.c_uint64),
                ("gpu_ms", C.c_double), ("wall_ms", C.c_double),
                ("n_dp_calls", C.c_int32), ("memo_kind", C.c_int32),
                ("filter_ms", C.c_double), ("eval_ms", C.c_double),
                ("filter_launches", C.c_int32), ("eval_launches", C.c_int32),
                ("host_prep_ms", C.c_double), ("host_enqueue_ms", C.c_double), ("host_post_ms", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("tree_two_lane_waves", C.c_int32), ("general_grid_max", C.c_int32),
                ("tree_task_iters", C.c_int32), ("tree_expand_max_rels", C.c_int32),
                ("tree_one_xcd_wgs", C.c_int32)]


class PlanOut(C.Structure):
    _fields_ = [("nodes", C.POINTER(PlanNode)), ("capacity", C.c_int32),
                ("n_nodes", C.c_int32), ("stats", PlanStats)]


def default_config(**overrides):
    """Reference GUC defaults (src/backend/utils/misc/guc.c), plus overrides.

    Mirrors gq_config_default() so the oracle can be configured without the
    product library; tests check the two agree.
    """
    c = Config()
    c.gpuqo_algorithm = ALGO["dpsize"]
    c.gpuqo_dpsub_bicc = 0
    c.gpuqo_dpsub_filter = 1
    c.gpuqo_dpsub_filter_threshold = 16
    c.gpuqo_dpsub_filter_cpu_enum_threshold = 1024
    c.gpuqo_dpsub_filter_keys_overprovisioning = 100
    c.gpuqo_dpsub_ccc = 1
    c.gpuqo_dpsub_csg = 1
    c.gpuqo_dpsub_csg_threshold = 128
    c.gpuqo_dpsub_tree = 0
    c.gpuqo_spanning_tree = 0
    c.gpuqo_n_parallel = 10240
    c.gpuqo_scratchpad_size_mb = 10
    c.gpuqo_min_memo_size_mb = 1
    c.gpuqo_max_memo_size_mb = 1024
    c.gpuqo_idp_type = IDP["idp2"]
    c.gpuqo_idp_n_iters = 0
    c.gpuqo_dpe_n_threads = 4
    c.gpuqo_cpu_dpsub_parallel_chunk_size = 64
    c.gpuqo_dpe_pairs_per_depbuf = 128
    c.seq_page_cost = 1.0
    c.random_page_cost = 4.0
    c.cpu_tuple_cost = 0.01
    c.cpu_index_tuple_cost = 0.005
    c.cpu_operator_cost = 0.0025
    c.disable_cost = 1.0e10
    c.effective_cache_size = 524288
    c.work_mem = 4096
    c.enable_seqscan = 1
    c.enable_indexscan = 1
    c.enable_tidscan = 1
    c.enable_sort = 1
    c.enable_hashagg = 1
    c.enable_nestloop = 1
    c.enable_mergejoin = 1              # guc.c:1009-1017 (not read by the join cost model)
    c.enable_hashjoin = 1
    c.memo_kind = MEMO["auto"]
    c.device = 0
    c.profile_kernels = 0
    c.cost_function = COST_FUNCTION["postgres"]
    for k, v in overrides.items():
        if k == "algorithm":
            v = ALGO[v] if isinstance(v, str) else v
            k = "gpuqo_algorithm"
        elif k == "idp_type":
            v = IDP[v] if isinstance(v, str) else v
            k = "gpuqo_idp_type"
        elif k == "memo_kind" and isinstance(v, str):
            v = MEMO[v]
        elif k == "cost_function" and isinstance(v, str):
            v = COST_FUNCTION[v]
        elif not hasattr(c, k) and hasattr(c, "gpuqo_" + k):
            k = "gpuqo_" + k
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    return c
