"""ctypes mirrors of the POD structs in include/gp_presolve_capi.h.

Field order and types must match the C header exactly; the C library and
the CPU oracle both take these structs by pointer.
"""
import ctypes as C

import numpy as np

i32p = C.POINTER(C.c_int32)
f64p = C.POINTER(C.c_double)
u8p = C.POINTER(C.c_uint8)

GP_RULE = {
    "empty_rows": 0,
    "singleton_rows": 1,
    "infeasible_redundant_rows": 2,
    "implied_variable_bounds": 3,
    "duplicate_rows": 4,
    "infeasible_fixed_variables": 5,
    "empty_cols": 6,
    "column_singletons_dual_infer": 7,
    "column_singletons_eq": 8,
    "doubleton_equations": 9,
    "dual_fix": 10,
    "duplicate_columns": 11,
    "redundant_bounds": 12,
}

# PostsolveReductionType (reference presolve_structs.hpp:231-247)
REC = {
    "FixedCol": 0, "FixedColInf": 1, "SubCol": 2, "DuplicateColumn": 3, "DuplicateRow": 4,
    "DeletedRow": 5, "AddedRow": 6, "AddedRows": 7, "LhsChange": 8, "RhsChange": 9,
    "EqToIneq": 10, "BoundChangeNoRow": 11, "BoundChangeTheRow": 12, "DoubletonEquation": 13,
    "FmeCol": 14,
}


class GpCsr(C.Structure):
    _fields_ = [("rows", C.c_int32), ("cols", C.c_int32), ("nnz", C.c_int32),
                ("row_ptr", C.c_void_p), ("col_idx", C.c_void_p), ("val", C.c_void_p)]


class GpLp(C.Structure):
    _fields_ = [("A", GpCsr), ("AT", GpCsr), ("c", C.c_void_p), ("AL", C.c_void_p),
                ("AU", C.c_void_p), ("l", C.c_void_p), ("u", C.c_void_p),
                ("obj_constant", C.c_double)]


_PARAM_FIELDS = [
    ("max_iters", C.c_int32), ("max_time", C.c_double), ("verbose", C.c_int32),
    ("record_postsolve_tape", C.c_int32), ("record_postsolve_tape_cpu", C.c_int32),
    ("feasibility_tol", C.c_double), ("bound_tol", C.c_double), ("zero_tol", C.c_double),
    ("implied_variable_bounds_max_rounds", C.c_int32),
    ("implied_variable_bounds_max_bound_only_rounds", C.c_int32),
    ("implied_variable_bounds_preserve_free_zero_cost", C.c_int32),
    ("doubleton_equations_single_batch_per_iter", C.c_int32),
    ("doubleton_equations_max_fill_in_proxy", C.c_int32),
    ("doubleton_equations_scan", C.c_int32),
    ("doubleton_equations_min_selected_per_batch", C.c_int32),
    ("doubleton_equations_min_selected_ratio", C.c_double),
    ("doubleton_equations_max_batch_rounds", C.c_int32),
    ("doubleton_equations_max_time", C.c_double),
    ("enable_infeasible_fixed_variables", C.c_int32), ("enable_empty_rows", C.c_int32),
    ("enable_singleton_rows", C.c_int32), ("enable_infeasible_redundant_rows", C.c_int32),
    ("enable_implied_variable_bounds", C.c_int32), ("enable_duplicate_rows", C.c_int32),
    ("enable_empty_cols", C.c_int32), ("enable_column_singletons_eq", C.c_int32),
    ("enable_column_singletons_dual_infer", C.c_int32), ("enable_doubleton_equations", C.c_int32),
    ("enable_dual_fix", C.c_int32), ("dual_fix_max_nonzero_cost_col_degree", C.c_int32),
    ("enable_duplicate_columns", C.c_int32), ("duplicate_columns_max_abs_ratio", C.c_double),
    ("duplicate_columns_scale_guard_max_group_size", C.c_int32),
    ("duplicate_columns_scale_guard_min_groups", C.c_int32),
    ("enable_structure_specific_rules", C.c_int32),
    ("enable_zero_cost_redundant_box_bounds", C.c_int32), ("enable_redundant_bounds", C.c_int32),
    ("use_tiered_scheduler", C.c_int32), ("enable_tiered_bootstrap", C.c_int32),
    ("implied_variable_bounds_bound_only_nnz_round_budget", C.c_int64),
    ("enable_bounded_two_row_projection", C.c_int32),
    ("bounded_two_row_projection_max_row_nnz", C.c_int32),
    ("bounded_two_row_projection_safety_tol", C.c_double),
    ("bounded_two_row_projection_min_candidates", C.c_int32),
    ("bounded_two_row_projection_min_candidate_ratio", C.c_double),
    ("enable_orphan_mccormick_projection", C.c_int32),
    ("orphan_mccormick_projection_min_candidates", C.c_int32),
    ("orphan_mccormick_projection_min_candidate_ratio", C.c_double),
    ("enable_linf_components", C.c_int32),
    ("linf_probe_pairs", C.c_int32),
    ("linf_min_pairs", C.c_int32),
    ("linf_max_tail_cols", C.c_int32),
    ("enable_structural_l1_substitution", C.c_int32),
    ("enable_antipodal_components", C.c_int32),
    ("antipodal_probe_pairs", C.c_int32),
    ("antipodal_min_pairs", C.c_int32),
    ("enable_covering_cost_dominance", C.c_int32),
    ("covering_cost_dominance_min_fixed_cols", C.c_int32),
    ("covering_cost_dominance_min_candidate_ratio", C.c_double),
]


class GpParams(C.Structure):
    _fields_ = _PARAM_FIELDS

    def set(self, **kw):
        names = {f[0] for f in _PARAM_FIELDS}
        for k, v in kw.items():
            if k not in names:
                raise KeyError(f"unknown PresolveParams field {k!r}")
            setattr(self, k, int(v) if isinstance(v, bool) else v)
        return self


class GpSummary(C.Structure):
    _fields_ = [("original_rows", C.c_int32), ("original_cols", C.c_int32),
                ("reduced_rows", C.c_int32), ("reduced_cols", C.c_int32),
                ("reduced_nnz", C.c_int32), ("iterations", C.c_int32),
                ("elapsed_seconds", C.c_double), ("obj_constant_delta", C.c_double),
                ("has_infeasible", C.c_int32), ("has_unbounded", C.c_int32)]


class GpServeConfig(C.Structure):
    _fields_ = [("workers", C.c_int32), ("postsolve", C.c_int32), ("hbm_budget_bytes", C.c_double),
                ("seed", C.c_uint64)]


class GpJobResult(C.Structure):
    _fields_ = [("job", C.c_int64), ("tag", C.c_uint64), ("status", C.c_int32), ("worker", C.c_int32),
                ("reduced_rows", C.c_int32), ("reduced_cols", C.c_int32), ("reduced_nnz", C.c_int32),
                ("iterations", C.c_int32), ("has_infeasible", C.c_int32), ("has_unbounded", C.c_int32),
                ("checksum", C.c_uint64), ("t_arrival", C.c_double), ("t_start", C.c_double),
                ("t_end", C.c_double), ("t_ready", C.c_double)]


class GpPostsolveShape(C.Structure):
    _fields_ = [("records", C.c_int64), ("primal_levels", C.c_int32), ("dual_levels", C.c_int32),
                ("tape_bytes", C.c_int64)]


class GpPdhgParams(C.Structure):
    _fields_ = [("max_iters", C.c_int32), ("max_time", C.c_double), ("tol", C.c_double),
                ("check_every", C.c_int32), ("ruiz_iters", C.c_int32), ("power_iters", C.c_int32),
                ("use_graph", C.c_int32), ("verbose", C.c_int32), ("method", C.c_int32)]


class GpPdhgResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("iterations", C.c_int32), ("restarts", C.c_int32),
                ("checks", C.c_int32), ("primal_obj", C.c_double), ("dual_obj", C.c_double),
                ("primal_res", C.c_double), ("dual_res", C.c_double), ("gap", C.c_double),
                ("elapsed_seconds", C.c_double), ("norm_estimate", C.c_double), ("primal_weight", C.c_double),
                ("graph_launches", C.c_int64)]


class GpPipelineResult(C.Structure):
    _fields_ = [("reduced_rows", C.c_int32), ("reduced_cols", C.c_int32), ("reduced_nnz", C.c_int32),
                ("presolve_iterations", C.c_int32), ("obj_constant_delta", C.c_double),
                ("has_infeasible", C.c_int32), ("has_unbounded", C.c_int32), ("solve", GpPdhgResult),
                ("presolve_seconds", C.c_double), ("solve_seconds", C.c_double), ("prepare_seconds", C.c_double),
                ("postsolve_seconds", C.c_double), ("total_seconds", C.c_double)]


class GpLpSpec(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("m", C.c_int32), ("n", C.c_int32), ("len_lo", C.c_int32),
                ("len_hi", C.c_int32), ("mean_len", C.c_double), ("min_len", C.c_int32)]


class GpMpsView(C.Structure):
    _fields_ = [("name", C.c_char_p), ("rows", C.c_int32), ("cols", C.c_int32), ("nnz", C.c_int32),
                ("csr_row_ptr", i32p), ("csr_col_idx", i32p), ("csr_val", f64p),
                ("csc_col_ptr", i32p), ("csc_row_idx", i32p), ("csc_val", f64p),
                ("c", f64p), ("row_lower", f64p), ("row_upper", f64p), ("col_lower", f64p),
                ("col_upper", f64p), ("obj_constant", C.c_double)]


def mps_view_arrays(v) -> dict:
    """Copies of the arrays of a gp_mps_view (or the same layout), keyed as
    an LP dict: A (CSR), AT (the CSC read as the transpose's CSR), c, l, u,
    AL, AU."""
    m, n, nnz = v.rows, v.cols, v.nnz
    ptr = _arr(v.csr_row_ptr, m + 1, np.int32) if m + 1 > 0 else np.zeros(1, np.int32)
    tptr = _arr(v.csc_col_ptr, n + 1, np.int32)
    return {"A_ptr": ptr, "A_idx": _arr(v.csr_col_idx, nnz, np.int32), "A_val": _arr(v.csr_val, nnz, np.float64),
            "AT_ptr": tptr, "AT_idx": _arr(v.csc_row_idx, nnz, np.int32), "AT_val": _arr(v.csc_val, nnz, np.float64),
            "c": _arr(v.c, n, np.float64), "l": _arr(v.col_lower, n, np.float64),
            "u": _arr(v.col_upper, n, np.float64), "AL": _arr(v.row_lower, m, np.float64),
            "AU": _arr(v.row_upper, m, np.float64)}


class GpRecordView(C.Structure):
    _fields_ = [("m0", C.c_int32), ("n0", C.c_int32), ("m1", C.c_int32), ("n1", C.c_int32),
                ("row_red2org", i32p), ("col_red2org", i32p),
                ("fixed_count", C.c_int32), ("fixed_idx", i32p), ("fixed_val", f64p),
                ("obj_constant_old", C.c_double), ("obj_constant_new", C.c_double),
                ("tape_count", C.c_int32), ("tape_types", i32p), ("tape_index_starts", i32p),
                ("tape_value_starts", i32p), ("tape_dual_modes", u8p),
                ("tape_index_count", C.c_int32), ("tape_indices", i32p),
                ("tape_value_count", C.c_int32), ("tape_vals", f64p)]


class GpPlanView(C.Structure):
    _fields_ = [("keep_row", C.c_void_p), ("keep_col", C.c_void_p), ("c", C.c_void_p),
                ("l", C.c_void_p), ("u", C.c_void_p), ("AL", C.c_void_p), ("AU", C.c_void_p),
                ("obj_constant_delta", C.c_double), ("has_change", C.c_int32),
                ("has_row_action", C.c_int32), ("has_col_action", C.c_int32),
                ("has_infeasible", C.c_int32), ("has_unbounded", C.c_int32),
                ("has_new_A", C.c_int32), ("new_A", GpCsr)]


class GpStatsView(C.Structure):
    _fields_ = [("row_nnz", C.c_void_p), ("singleton_row_mask", C.c_void_p),
                ("singleton_row_col", C.c_void_p), ("singleton_row_val", C.c_void_p),
                ("col_nnz", C.c_void_p), ("empty_col_mask", C.c_void_p),
                ("column_singleton_mask", C.c_void_p), ("column_singleton_row", C.c_void_p),
                ("column_singleton_val", C.c_void_p)]


def _arr(ptr, n, dtype):
    if n <= 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).copy()


def record_to_dict(rv: GpRecordView) -> dict:
    """Copies a gp_record_view (host pointers) into numpy arrays."""
    t = rv.tape_count
    return {
        "m0": rv.m0, "n0": rv.n0, "m1": rv.m1, "n1": rv.n1,
        "row_red2org": _arr(rv.row_red2org, rv.m1, np.int32),
        "col_red2org": _arr(rv.col_red2org, rv.n1, np.int32),
        "fixed_idx": _arr(rv.fixed_idx, rv.fixed_count, np.int32),
        "fixed_val": _arr(rv.fixed_val, rv.fixed_count, np.float64),
        "obj_constant_old": rv.obj_constant_old, "obj_constant_new": rv.obj_constant_new,
        "types": _arr(rv.tape_types, t, np.int32),
        "index_starts": _arr(rv.tape_index_starts, t + 1, np.int32) if t else np.zeros(1, np.int32),
        "value_starts": _arr(rv.tape_value_starts, t + 1, np.int32) if t else np.zeros(1, np.int32),
        "dual_modes": _arr(rv.tape_dual_modes, t, np.uint8),
        "indices": _arr(rv.tape_indices, rv.tape_index_count, np.int32),
        "vals": _arr(rv.tape_vals, rv.tape_value_count, np.float64),
    }


class RecordKeepAlive:
    """A gp_record_view built from numpy arrays (host), kept alive together."""

    def __init__(self, rec: dict):
        self._keep = {}

        def p(name, dtype, ctype):
            a = np.ascontiguousarray(rec[name], dtype=dtype)
            if a.size == 0:
                a = np.zeros(1, dtype=dtype)
            self._keep[name] = a
            return a.ctypes.data_as(C.POINTER(ctype))

        v = GpRecordView()
        v.m0, v.n0, v.m1, v.n1 = rec["m0"], rec["n0"], rec["m1"], rec["n1"]
        v.row_red2org = p("row_red2org", np.int32, C.c_int32)
        v.col_red2org = p("col_red2org", np.int32, C.c_int32)
        v.fixed_count = len(rec["fixed_idx"])
        v.fixed_idx = p("fixed_idx", np.int32, C.c_int32)
        v.fixed_val = p("fixed_val", np.float64, C.c_double)
        v.obj_constant_old = rec.get("obj_constant_old", 0.0)
        v.obj_constant_new = rec.get("obj_constant_new", 0.0)
        v.tape_count = len(rec["types"])
        v.tape_types = p("types", np.int32, C.c_int32)
        v.tape_index_starts = p("index_starts", np.int32, C.c_int32)
        v.tape_value_starts = p("value_starts", np.int32, C.c_int32)
        dm = rec.get("dual_modes", np.zeros(len(rec["types"]), np.uint8))
        self._keep["dual_modes_src"] = dm
        rec2 = dict(rec)
        rec2["dual_modes"] = dm
        a = np.ascontiguousarray(dm, dtype=np.uint8)
        if a.size == 0:
            a = np.zeros(1, np.uint8)
        self._keep["dual_modes"] = a
        v.tape_dual_modes = a.ctypes.data_as(u8p)
        v.tape_index_count = len(rec["indices"])
        v.tape_indices = p("indices", np.int32, C.c_int32)
        v.tape_value_count = len(rec["vals"])
        v.tape_vals = p("vals", np.float64, C.c_double)
        self.view = v


def tape_records(rec: dict):
    """Yields (type, indices, vals) per tape record."""
    ist, vst = rec["index_starts"], rec["value_starts"]
    for k, t in enumerate(rec["types"]):
        yield int(t), rec["indices"][ist[k]:ist[k + 1]], rec["vals"][vst[k]:vst[k + 1]]
