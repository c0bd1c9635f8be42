"""Writes reference_kats.json: the known-answer cases of the reference's own
unit tests, transcribed by hand as data (inputs and expected outputs).

Every case names the reference test it comes from (paths relative to the
PolyU-IOR/GPU-Presolver tree).  Inputs that the reference test leaves unset
are omitted here; the runner fills them with neutral defaults (masks all
kept, bounds free, the missing one of A / AT as the transpose of the other,
statistics from the CSR counts).  Expectations are a small assertion
language read by tests/golden_runner.py:

  ("eq", field, values)            exact equality of the whole array
  ("approx", field, index, value)  |got - value| < 1e-12 (index None: scalar;
                                   infinities must match in sign)
  ("bits", field, index, value)    bit-identical double
  ("len", field, n)                array length
  ("flag", name, 0 or 1)           plan flag
  ("le", field, i, field2, j)      got[i] <= got2[j]
  ("gt", field, value)             scalar > value
  ("index_at", idx, starts, k, v)  idx[starts[k]] == v (first index of record k)
  ("one_of", [asserts...], ...)    at least one list of assertions holds

Run:  python tests/golden/make_reference_kats.py
"""
import json
import os

INF = float("inf")
HERE = os.path.dirname(os.path.abspath(__file__))

# PostsolveReductionType / PostsolveDualMode codes
FixedCol, FixedColInf, SubCol, DuplicateColumn, DuplicateRow, DeletedRow = 0, 1, 2, 3, 4, 5
LhsChange, RhsChange, EqToIneq, BoundChangeNoRow, BoundChangeTheRow, DoubletonEquation, FmeCol = \
    8, 9, 10, 11, 12, 13, 14
EXACT, MINIMAL = 1, 2

CASES = []


def case(**kw):
    CASES.append(kw)


def csr(ptr, idx, val):
    return [list(ptr), list(idx), [float(v) for v in val]]


# ---------------------------------------------------------------- kernels --
case(name="row_activity_summary", source="cpp/tests/test_gpu_presolve_kernels.cpp:22-67", kind="activity",
     m=3, n=3, A=csr([0, 0, 1, 3], [2, 0, 1], [4.0, 5.0, 6.0]), l=[-1.0, -INF, 2.0], u=[3.0, 5.0, INF],
     zero_tol=1e-12,
     expect=[("eq", "min_fin", [0.0, 8.0, -5.0]), ("eq", "max_fin", [0.0, 0.0, 45.0]),
             ("eq", "neg_inf", [0, 0, 1]), ("eq", "pos_inf", [0, 1, 0])])
case(name="col_max_abs", source="cpp/tests/test_gpu_presolve_kernels.cpp:69-78", kind="col_max_abs",
     n=3, m=3, AT=csr([0, 1, 2, 3], [2, 2, 1], [5.0, -6.0, 4.0]),
     expect=[("eq", "col_max_abs", [5.0, 6.0, 4.0])])
case(name="nnz_progress", source="cpp/tests/test_gpu_direct_port.cpp:274-276", kind="progress",
     checks=[[0, 0, 0.95, False], [100, 94, 0.95, True], [100, 95, 0.95, False]])

# ------------------------------------------------------------- empty rows --
case(name="empty_rows_removes_feasible", source="cpp/tests/test_rule_empty_rows.cpp:21-55", kind="rule",
     lp=dict(m=3, n=2), plan=dict(keep_row=[1, 1, 1], AL=[-1.0, -2.0, 0.0], AU=[1.0, 3.0, 0.0]),
     stats=dict(row_nnz=[0, 0, 1]),
     steps=[dict(rule="empty_rows", params=dict(feasibility_tol=1e-9),
                 expect=[("eq", "plan.keep_row", [0, 0, 1]), ("flag", "has_infeasible", 0),
                         ("flag", "has_row_action", 1), ("flag", "has_change", 1)])])
case(name="empty_rows_infeasible", source="cpp/tests/test_rule_empty_rows.cpp:57-88", kind="rule",
     lp=dict(m=2, n=1), plan=dict(keep_row=[1, 1], AL=[-1.0, 2.0], AU=[1.0, 3.0]), stats=dict(row_nnz=[0, 0]),
     steps=[dict(rule="empty_rows", params=dict(feasibility_tol=1e-9),
                 expect=[("eq", "plan.keep_row", [1, 1]), ("flag", "has_infeasible", 1)])])

# --------------------------------------------------------- singleton rows --
_srow = dict(lp=dict(m=2, n=2),
             plan=dict(keep_row=[1, 1], l=[0.0, -10.0], u=[100.0, 10.0], AL=[4.0, -6.0], AU=[8.0, -2.0]),
             stats=dict(singleton_row_mask=[1, 1], singleton_row_col=[0, 1], singleton_row_val=[2.0, -2.0]))
case(name="singleton_rows_tighten_and_delete", source="cpp/tests/test_rule_singleton_rows.cpp:22-64", kind="rule",
     **_srow,
     steps=[dict(rule="singleton_rows", params=dict(zero_tol=1e-12, bound_tol=1e-9),
                 expect=[("eq", "plan.keep_row", [0, 0]), ("approx", "plan.l", 0, 2.0),
                         ("approx", "plan.u", 0, 4.0), ("approx", "plan.l", 1, 1.0), ("approx", "plan.u", 1, 3.0),
                         ("flag", "has_row_action", 1), ("flag", "has_change", 1), ("len", "tape.types", 0)])])
case(name="singleton_rows_cpu_tape", source="cpp/tests/test_rule_singleton_rows.cpp:66-101", kind="rule", **_srow,
     steps=[dict(rule="singleton_rows", params=dict(zero_tol=1e-12, bound_tol=1e-9, record_postsolve_tape_cpu=1),
                 expect=[("eq", "tape.types", [DeletedRow, DeletedRow]), ("eq", "tape.index_starts", [0, 2, 4]),
                         ("eq", "tape.value_starts", [0, 3, 6]), ("eq", "tape.dual_modes", [MINIMAL, MINIMAL]),
                         ("eq", "tape.indices", [0, 0, 1, 1]), ("approx", "tape.vals", 0, 4.0),
                         ("approx", "tape.vals", 1, 8.0), ("approx", "tape.vals", 2, 2.0),
                         ("approx", "tape.vals", 3, -6.0), ("approx", "tape.vals", 4, -2.0),
                         ("approx", "tape.vals", 5, -2.0)])])

# ----------------------------------------------- infeasible/redundant rows --
_rr = dict(lp=dict(m=5, n=2, A=csr([0, 2, 4, 6, 8, 9], [0, 1, 0, 1, 0, 1, 0, 1, 0],
                                   [1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, -1.0, 1.0])),
           plan=dict(keep_row=[1, 1, 1, 1, 1], AL=[-INF, -1.0, 2.0, -10.0, -100.0],
                     AU=[10.0, 5.0, 10.0, 10.0, 100.0], l=[0.0, 0.0], u=[4.0, 4.0]),
           stats=dict(row_nnz=[2, 2, 2, 2, 1]))
case(name="redundant_rows_classification", source="cpp/tests/test_rule_infeasible_redundant_rows.cpp:23-67",
     kind="rule", **_rr,
     steps=[dict(rule="infeasible_redundant_rows", params=dict(bound_tol=1e-9),
                 expect=[("eq", "plan.keep_row", [0, 1, 1, 0, 1]), ("approx", "plan.AL", 1, -INF),
                         ("approx", "plan.AU", 2, INF), ("flag", "has_row_action", 1), ("flag", "has_change", 1),
                         ("flag", "has_infeasible", 0), ("len", "tape.types", 0)])])
case(name="redundant_rows_cpu_tape", source="cpp/tests/test_rule_infeasible_redundant_rows.cpp:68-121",
     kind="rule", **_rr,
     steps=[dict(rule="infeasible_redundant_rows", params=dict(bound_tol=1e-9, record_postsolve_tape_cpu=1),
                 expect=[("eq", "tape.types", [DeletedRow, DeletedRow, LhsChange, RhsChange]),
                         ("eq", "tape.index_starts", [0, 1, 2, 3, 4]),
                         ("eq", "tape.value_starts", [0, 2, 4, 8, 12]),
                         ("eq", "tape.dual_modes", [MINIMAL] * 4), ("eq", "tape.indices", [0, 3, 1, 2]),
                         ("approx", "tape.vals", 0, -INF), ("approx", "tape.vals", 1, 10.0),
                         ("approx", "tape.vals", 2, -10.0), ("approx", "tape.vals", 3, 10.0),
                         ("approx", "tape.vals", 4, -1.0), ("approx", "tape.vals", 5, 5.0),
                         ("approx", "tape.vals", 6, -INF), ("approx", "tape.vals", 7, 5.0),
                         ("approx", "tape.vals", 8, 2.0), ("approx", "tape.vals", 9, 10.0),
                         ("approx", "tape.vals", 10, 2.0), ("approx", "tape.vals", 11, INF)])])
case(name="redundant_rows_infeasible_short_circuit",
     source="cpp/tests/test_rule_infeasible_redundant_rows.cpp:123-161", kind="rule",
     lp=dict(m=1, n=2, A=csr([0, 2], [0, 1], [1.0, 1.0])),
     plan=dict(keep_row=[1], AL=[9.0], AU=[20.0], l=[0.0, 0.0], u=[4.0, 4.0]), stats=dict(row_nnz=[2]),
     steps=[dict(rule="infeasible_redundant_rows", params={},
                 expect=[("flag", "has_infeasible", 1), ("eq", "plan.keep_row", [1])])])

# ------------------------------------------------------- implied bounds --
case(name="implied_bounds_tighten", source="cpp/tests/test_rule_implied_variable_bounds.cpp:32-103", kind="rule",
     lp=dict(m=2, n=3, A=csr([0, 2, 4], [0, 1, 0, 2], [1.0] * 4), AT=csr([0, 2, 3, 4], [0, 1, 0, 1], [1.0] * 4)),
     plan=dict(keep_row=[1, 1], keep_col=[1, 1, 1], c=[0.0, 0.0, 0.0], AL=[5.0, -INF], AU=[INF, 6.0],
               l=[0.0, 2.0, 0.0], u=[10.0, 2.0, 0.0]),
     stats=dict(row_nnz=[2, 2]),
     steps=[dict(rule="implied_variable_bounds", params=dict(feasibility_tol=1e-9, zero_tol=1e-12),
                 expect=[("approx", "plan.l", 0, 3.0), ("approx", "plan.u", 0, 6.0), ("flag", "has_row_action", 0),
                         ("flag", "has_col_action", 0), ("flag", "has_change", 1), ("flag", "has_infeasible", 0),
                         ("eq", "tape.types", [BoundChangeTheRow, BoundChangeTheRow]),
                         ("eq", "tape.index_starts", [0, 2, 4]), ("eq", "tape.value_starts", [0, 4, 8]),
                         ("eq", "tape.dual_modes", [MINIMAL, MINIMAL]), ("eq", "tape.indices", [0, 0, 0, 1]),
                         ("approx", "tape.vals", 0, 0.0), ("approx", "tape.vals", 1, 10.0),
                         ("approx", "tape.vals", 2, 3.0), ("approx", "tape.vals", 3, 10.0),
                         ("approx", "tape.vals", 4, 3.0), ("approx", "tape.vals", 5, 10.0),
                         ("approx", "tape.vals", 6, 3.0), ("approx", "tape.vals", 7, 6.0)])])
case(name="implied_bounds_fixed_column", source="cpp/tests/test_rule_implied_variable_bounds.cpp:105-162",
     kind="rule", lp=dict(m=1, n=2, A=csr([0, 2], [0, 1], [1.0, 1.0]), AT=csr([0, 1, 2], [0, 0], [1.0, 1.0])),
     plan=dict(keep_row=[1], keep_col=[1, 1], c=[4.0, 0.0], AL=[6.0], AU=[INF], l=[0.0, 5.0], u=[1.0, 5.0]),
     stats=dict(row_nnz=[2]),
     steps=[dict(rule="implied_variable_bounds", params=dict(feasibility_tol=1e-9, zero_tol=1e-12),
                 expect=[("eq", "plan.keep_col", [0, 1]), ("approx", "plan.l", 0, 1.0), ("approx", "plan.u", 0, 1.0),
                         ("approx", "obj", None, 4.0), ("eq", "tape.types", [FixedCol]),
                         ("eq", "tape.indices", [0, 0]), ("approx", "tape.vals", 0, 1.0),
                         ("approx", "tape.vals", 1, 4.0), ("approx", "tape.vals", 2, 1.0)])])
case(name="implied_bounds_infeasible", source="cpp/tests/test_rule_implied_variable_bounds.cpp:164-206",
     kind="rule", lp=dict(m=1, n=2, A=csr([0, 2], [0, 1], [1.0, 1.0]), AT=csr([0, 1, 2], [0, 0], [1.0, 1.0])),
     plan=dict(keep_row=[1], keep_col=[1, 1], c=[0.0, 0.0], AL=[10.0], AU=[INF], l=[0.0, 0.0], u=[4.0, 0.0]),
     stats=dict(row_nnz=[2]),
     steps=[dict(rule="implied_variable_bounds", params={}, expect=[("flag", "has_infeasible", 1)])])
case(name="implied_bounds_preserve_free_zero_cost", source="cpp/tests/test_rule_implied_variable_bounds.cpp:208-265",
     kind="rule",
     lp=dict(m=2, n=2, A=csr([0, 2, 4], [0, 1, 0, 1], [1.0] * 4), AT=csr([0, 2, 4], [0, 1, 0, 1], [1.0] * 4)),
     plan=dict(keep_row=[1, 1], keep_col=[1, 1], c=[0.0, 0.0], AL=[-5.0, -INF], AU=[INF, 5.0],
               l=[-INF, 0.0], u=[INF, 0.0]),
     stats=dict(row_nnz=[2, 2]),
     steps=[dict(rule="implied_variable_bounds", params=dict(feasibility_tol=1e-9, zero_tol=1e-12),
                 expect=[("approx", "plan.l", 0, -INF), ("approx", "plan.u", 0, INF), ("flag", "has_change", 0),
                         ("len", "tape.types", 0)]),
            dict(rule="implied_variable_bounds",
                 params=dict(feasibility_tol=1e-9, zero_tol=1e-12,
                             implied_variable_bounds_preserve_free_zero_cost=0),
                 expect=[("approx", "plan.l", 0, -5.0), ("approx", "plan.u", 0, 5.0), ("flag", "has_change", 1)])])
case(name="implied_bounds_free_zero_cost_across_rounds",
     source="cpp/tests/test_rule_implied_variable_bounds.cpp:267-330", kind="rule",
     lp=dict(m=2, n=3, A=csr([0, 2, 4], [0, 1, 1, 2], [1.0, -1.0, 1.0, 1.0]),
             AT=csr([0, 1, 3, 4], [0, 0, 1, 1], [1.0, -1.0, 1.0, 1.0]), l=[-INF, 0.0, 0.0], u=[INF, INF, 0.0]),
     plan=dict(keep_row=[1, 1], keep_col=[1, 1, 1], c=[0.0, 0.0, 0.0], AL=[0.0, -INF], AU=[10.0, 1.0],
               l=[-INF, 0.0, 0.0], u=[INF, INF, 0.0]),
     stats=dict(row_nnz=[2, 2]),
     steps=[dict(rule="implied_variable_bounds", params=dict(feasibility_tol=1e-9, zero_tol=1e-12),
                 expect=[("approx", "plan.l", 0, 0.0), ("approx", "plan.u", 0, INF), ("approx", "plan.u", 1, 1.0)]),
            dict(rule="implied_variable_bounds", params=dict(feasibility_tol=1e-9, zero_tol=1e-12),
                 set=dict(has_change=0),
                 expect=[("approx", "plan.l", 0, 0.0), ("approx", "plan.u", 0, INF), ("flag", "has_change", 0)])])

# ------------------------------------------------------- duplicate rows --
case(name="duplicate_rows_delete_and_tighten", source="cpp/tests/test_rule_duplicate_rows.cpp:34-77", kind="rule",
     lp=dict(m=2, n=2, A=csr([0, 2, 4], [0, 1, 0, 1], [1.0, 1.0, 2.0, 2.0])),
     plan=dict(keep_row=[1, 1], AL=[0.0, 4.0], AU=[10.0, 12.0]),
     steps=[dict(rule="duplicate_rows", params={},
                 expect=[("one_of",
                          [("eq", "plan.keep_row", [1, 0]), ("le", "plan.AL", 0, "plan.AU", 0),
                           ("eq", "tape.indices", [0, 1])],
                          [("eq", "plan.keep_row", [0, 1]), ("le", "plan.AL", 1, "plan.AU", 1),
                           ("eq", "tape.indices", [1, 0])]),
                         ("flag", "has_row_action", 1), ("flag", "has_change", 1),
                         ("eq", "tape.types", [DuplicateRow]), ("eq", "tape.index_starts", [0, 2]),
                         ("eq", "tape.value_starts", [0, 5]), ("eq", "tape.dual_modes", [EXACT]),
                         ("len", "tape.vals", 5)])])
case(name="duplicate_rows_disjoint_intervals", source="cpp/tests/test_rule_duplicate_rows.cpp:79-103", kind="rule",
     lp=dict(m=2, n=2, A=csr([0, 2, 4], [0, 1, 0, 1], [1.0] * 4)),
     plan=dict(keep_row=[1, 1], AL=[0.0, 3.0], AU=[1.0, 4.0]),
     steps=[dict(rule="duplicate_rows", params={}, expect=[("flag", "has_infeasible", 1)])])

# ------------------------------------------------------- fixed variables --
_lo, _hi = 2.0, 2.0 + 5.0e-7
_mid = _lo + 0.5 * (_hi - _lo)
case(name="fixed_variables_midpoint_and_shift", source="cpp/tests/test_rule_infeasible_fixed_variables.cpp:45-123",
     kind="rule", lp=dict(m=2, n=2, AT=csr([0, 2, 3], [0, 1, 1], [2.0, -1.0, 3.0])),
     plan=dict(keep_col=[1, 1], keep_row=[1, 1], c=[4.0, 5.0], l=[_lo, 0.0], u=[_hi, 10.0], AL=[5.0, -3.0],
               AU=[7.0, 9.0]),
     steps=[dict(rule="infeasible_fixed_variables", params=dict(bound_tol=1e-6),
                 expect=[("eq", "plan.keep_col", [0, 1]), ("approx", "plan.l", 0, _mid),
                         ("approx", "plan.u", 0, _mid), ("approx", "plan.AL", 0, 5.0 - 2.0 * _mid),
                         ("approx", "plan.AU", 0, 7.0 - 2.0 * _mid), ("approx", "plan.AL", 1, -3.0 + _mid),
                         ("approx", "plan.AU", 1, 9.0 + _mid), ("approx", "obj", None, 4.0 * _mid),
                         ("flag", "has_col_action", 1), ("flag", "has_change", 1), ("len", "tape.types", 0)])])

# ----------------------------------------------------------- empty cols --
case(name="empty_cols_fix", source="cpp/tests/test_rule_empty_cols.cpp:23-63", kind="rule", lp=dict(m=0, n=4),
     plan=dict(keep_col=[1, 1, 1, 0], c=[2.0, -3.0, 0.0, 4.0], l=[1.0, -10.0, -INF, 5.0], u=[9.0, 7.0, 6.0, 5.0]),
     stats=dict(empty_col_mask=[1, 1, 1, 1]),
     steps=[dict(rule="empty_cols", params=dict(bound_tol=1e-9, zero_tol=1e-12),
                 expect=[("eq", "plan.keep_col", [0, 0, 0, 0]), ("approx", "obj", None, 2.0 * 1.0 - 3.0 * 7.0),
                         ("flag", "has_col_action", 1), ("flag", "has_change", 1), ("len", "tape.types", 0)])])
case(name="empty_cols_unbounded", source="cpp/tests/test_rule_empty_cols.cpp:65-91", kind="rule", lp=dict(m=0, n=1),
     plan=dict(keep_col=[1], c=[2.0], l=[-INF], u=[10.0]), stats=dict(empty_col_mask=[1]),
     steps=[dict(rule="empty_cols", params={}, expect=[("flag", "has_unbounded", 1)])])
case(name="empty_cols_infeasible", source="cpp/tests/test_rule_empty_cols.cpp:93-118", kind="rule", lp=dict(m=0, n=1),
     plan=dict(keep_col=[1], c=[0.0], l=[3.0], u=[1.0]), stats=dict(empty_col_mask=[1]),
     steps=[dict(rule="empty_cols", params={}, expect=[("flag", "has_infeasible", 1)])])

# ---------------------------------------------------- column singletons --
case(name="column_singleton_eq_eliminates", source="cpp/tests/test_rule_column_singletons.cpp:24-95", kind="rule",
     lp=dict(m=1, n=2, A=csr([0, 2], [0, 1], [1.0, 1.0]), AT=csr([0, 1, 2], [0, 0], [1.0, 1.0])),
     plan=dict(keep_row=[1], keep_col=[1, 1], c=[4.0, 1.0], l=[0.0, 2.0], u=[10.0, 2.0], AL=[5.0], AU=[5.0]),
     stats=dict(column_singleton_mask=[1, 0], column_singleton_row=[0, -1], column_singleton_val=[1.0, 0.0]),
     steps=[dict(rule="column_singletons_eq", params={},
                 expect=[("eq", "plan.keep_row", [0]), ("eq", "plan.keep_col", [0, 1]),
                         ("approx", "plan.c", 1, -3.0), ("approx", "obj", None, 20.0), ("flag", "has_row_action", 1),
                         ("flag", "has_col_action", 1), ("flag", "has_change", 1), ("eq", "tape.types", [SubCol]),
                         ("eq", "tape.indices", [0, 0, 1, 1]), ("len", "tape.vals", 7),
                         ("approx", "tape.vals", 0, 1.0), ("approx", "tape.vals", 1, 5.0),
                         ("approx", "tape.vals", 2, 0.0), ("approx", "tape.vals", 3, 10.0),
                         ("approx", "tape.vals", 4, 4.0), ("approx", "tape.vals", 5, 1.0),
                         ("approx", "tape.vals", 6, 1.0)])])
_bits = 0.3
_bits += -(100000000.1 * 0.1 / 3.0)
_bits += -(-99999999.9 * 0.2 / 7.0)
case(name="column_singleton_eq_objective_bitwise", source="cpp/tests/test_rule_column_singletons.cpp:97-165",
     kind="rule",
     lp=dict(m=2, n=3, A=csr([0, 2, 4], [0, 2, 1, 2], [3.0, 0.1, 7.0, 0.2]),
             AT=csr([0, 1, 2, 4], [0, 1, 0, 1], [3.0, 7.0, 0.1, 0.2])),
     plan=dict(keep_row=[1, 1], keep_col=[1, 1, 1], c=[100000000.1, -99999999.9, 0.3], l=[-INF, -INF, 0.0],
               u=[INF, INF, 0.0], AL=[5.0, 6.0], AU=[5.0, 6.0]),
     stats=dict(column_singleton_mask=[1, 1, 0], column_singleton_row=[0, 1, -1],
                column_singleton_val=[3.0, 7.0, 0.0]),
     steps=[dict(rule="column_singletons_eq", params={}, expect=[("bits", "plan.c", 2, _bits)])])
case(name="column_singleton_dual_unbounded", source="cpp/tests/test_rule_column_singletons.cpp:167-209", kind="rule",
     lp=dict(m=1, n=2, A=csr([0, 2], [0, 1], [1.0, 1.0])),
     plan=dict(keep_row=[1], keep_col=[1, 1], c=[1.0, 0.0], l=[-INF, 0.0], u=[10.0, 1.0], AL=[-INF], AU=[5.0]),
     stats=dict(column_singleton_mask=[1, 0], column_singleton_row=[0, -1], column_singleton_val=[1.0, 0.0]),
     steps=[dict(rule="column_singletons_dual_infer", params={}, expect=[("flag", "has_unbounded", 1)])])
_eqi = [("eq", "plan.keep_col", [1, 1]), ("flag", "has_row_action", 1), ("flag", "has_col_action", 0),
        ("flag", "has_change", 1), ("eq", "tape.types", [EqToIneq]), ("eq", "tape.index_starts", [0, 1]),
        ("eq", "tape.value_starts", [0, 1]), ("eq", "tape.indices", [0]), ("len", "tape.vals", 1),
        ("approx", "tape.vals", 0, 0.0), ("eq", "tape.dual_modes", [MINIMAL])]
case(name="column_singleton_dual_eq_to_ineq", source="cpp/tests/test_rule_column_singletons.cpp:211-272",
     kind="rule", lp=dict(m=1, n=2, A=csr([0, 2], [0, 1], [1.0, 1.0])),
     plan=dict(keep_row=[1], keep_col=[1, 1], c=[1.0, 0.0], l=[-INF, 0.0], u=[INF, 0.0], AL=[1.0], AU=[5.0]),
     stats=dict(column_singleton_mask=[1, 0], column_singleton_row=[0, -1], column_singleton_val=[1.0, 0.0]),
     steps=[dict(rule="column_singletons_dual_infer", params={},
                 expect=[("approx", "plan.AL", 0, 1.0), ("approx", "plan.AU", 0, 1.0)] + _eqi)])
for k, (obj, coef, lo, hi, lhs, rhs, side) in enumerate([(-1.0, 1.0, 10.0, 30.0, 5.0, 20.0, 20.0),
                                                          (1.0, -1.0, 0.0, 10.0, -20.0, -5.0, -5.0),
                                                          (1.0, 1.0, 0.0, 10.0, 5.0, 20.0, 5.0),
                                                          (-1.0, -1.0, 10.0, 30.0, -20.0, -5.0, -20.0)]):
    case(name=f"column_singleton_dual_one_sided_{k}", source="cpp/tests/test_rule_column_singletons.cpp:274-360",
         kind="rule", lp=dict(m=1, n=2, A=csr([0, 2], [0, 1], [coef, 1.0])),
         plan=dict(keep_row=[1], keep_col=[1, 1], c=[obj, 0.0], l=[lo, 0.0], u=[hi, 0.0], AL=[lhs], AU=[rhs]),
         stats=dict(column_singleton_mask=[1, 0], column_singleton_row=[0, -1], column_singleton_val=[coef, 0.0]),
         steps=[dict(rule="column_singletons_dual_infer", params={},
                     expect=[("approx", "plan.AL", 0, side), ("approx", "plan.AU", 0, side)] + _eqi)])

# --------------------------------------------------- doubleton equations --
_dbl = dict(lp=dict(m=1, n=2, A=csr([0, 2], [0, 1], [1.0, 2.0]), AT=csr([0, 1, 2], [0, 0], [1.0, 2.0])),
            plan=dict(keep_row=[1], keep_col=[1, 1], c=[3.0, 5.0], l=[0.0, -INF], u=[10.0, INF], AL=[4.0],
                      AU=[4.0]))
case(name="doubleton_single_path", source="cpp/tests/test_rule_doubleton_equations.cpp:23-86", kind="rule", **_dbl,
     steps=[dict(rule="doubleton_equations", params=dict(doubleton_equations_scan=0,
                                                         doubleton_equations_max_fill_in_proxy=10),
                 expect=[("eq", "plan.keep_row", [0]), ("eq", "plan.keep_col", [0, 1]),
                         ("approx", "plan.c", 1, -1.0), ("approx", "obj", None, 12.0), ("flag", "has_row_action", 1),
                         ("flag", "has_col_action", 1), ("flag", "has_change", 1)])])
case(name="doubleton_batch_tape", source="cpp/tests/test_rule_doubleton_equations.cpp:88-162", kind="rule", **_dbl,
     steps=[dict(rule="doubleton_equations",
                 params=dict(doubleton_equations_scan=1, doubleton_equations_min_selected_per_batch=1,
                             doubleton_equations_min_selected_ratio=0.0, doubleton_equations_max_batch_rounds=1),
                 expect=[("eq", "tape.types", [DoubletonEquation]), ("eq", "tape.index_starts", [0, 4]),
                         ("eq", "tape.value_starts", [0, 12]), ("eq", "tape.dual_modes", [MINIMAL]),
                         ("eq", "tape.indices", [0, 1, 0, 0]), ("len", "tape.vals", 12)]
                 + [("approx", "tape.vals", i, v) for i, v in
                    enumerate([1.0, 2.0, 4.0, 0.0, 10.0, -INF, INF, -3.0, 2.0, 3.0, 0.0, 0.0])])])


def _fill_cap_lp():
    m = 26
    ptr, idx = [0, 2], [0, 1]
    for row in range(1, 13):
        idx.append(0)
        ptr.append(len(idx))
    for row in range(13, m):
        idx.append(1)
        ptr.append(len(idx))
    at_idx = list(range(0, 13)) + [0] + list(range(13, m))
    return dict(m=m, n=2, A=csr(ptr, idx, [1.0] * 27), AT=csr([0, 13, 27], at_idx, [1.0] * 27))


for cap, changed in ((10, 0), (96, 1)):
    exp = [("flag", "has_change", changed)]
    if changed:
        exp += [("approx", "plan.keep_row", 0, 0), ("eq", "plan.keep_col", [0, 1])]
    case(name=f"doubleton_fill_cap_{cap}", source="cpp/tests/test_rule_doubleton_equations.cpp:164-271",
         kind="rule", lp=_fill_cap_lp(),
         plan=dict(keep_row=[1] * 26, keep_col=[1, 1], c=[0.0, 0.0], l=[-INF, -INF], u=[INF, INF],
                   AL=[0.0] + [-INF] * 25, AU=[0.0] + [INF] * 25),
         steps=[dict(rule="doubleton_equations",
                     params=dict(doubleton_equations_scan=1, doubleton_equations_max_fill_in_proxy=cap,
                                 doubleton_equations_min_selected_per_batch=1,
                                 doubleton_equations_min_selected_ratio=0.0, doubleton_equations_max_batch_rounds=1),
                     expect=exp)])

# -------------------------------------------------------------- dual fix --
case(name="dual_fix_improving_column", source="cpp/tests/test_rule_dual_fix.cpp:23-76", kind="rule",
     lp=dict(m=1, n=1, AT=csr([0, 1], [0], [1.0])),
     plan=dict(keep_col=[1], keep_row=[1], c=[3.0], l=[2.0], u=[INF], AL=[-INF], AU=[10.0]),
     steps=[dict(rule="dual_fix", params=dict(record_postsolve_tape_cpu=1),
                 expect=[("eq", "plan.keep_col", [0]), ("approx", "plan.AU", 0, 8.0), ("approx", "obj", None, 6.0),
                         ("flag", "has_col_action", 1), ("flag", "has_change", 1), ("eq", "tape.types", [FixedCol]),
                         ("eq", "tape.indices", [0, 0]), ("approx", "tape.vals", 0, 2.0),
                         ("approx", "tape.vals", 1, 3.0), ("approx", "tape.vals", 2, 1.0)])])
case(name="dual_fix_unbounded", source="cpp/tests/test_rule_dual_fix.cpp:78-111", kind="rule",
     lp=dict(m=0, n=1, AT=csr([0, 0], [], [])), plan=dict(keep_col=[1], c=[1.0], l=[-INF], u=[10.0]),
     steps=[dict(rule="dual_fix", params={}, expect=[("flag", "has_unbounded", 1)])])
case(name="dual_fix_degree_guard", source="cpp/tests/test_rule_dual_fix.cpp:113-157", kind="rule",
     lp=dict(m=1, n=1, AT=csr([0, 1], [0], [1.0])),
     plan=dict(keep_col=[1], keep_row=[1], c=[3.0], l=[2.0], u=[INF], AL=[-INF], AU=[10.0]),
     steps=[dict(rule="dual_fix", params=dict(dual_fix_max_nonzero_cost_col_degree=0),
                 expect=[("eq", "plan.keep_col", [1]), ("flag", "has_col_action", 0), ("flag", "has_change", 0),
                         ("len", "tape.types", 0)])])
case(name="dual_fix_infinite_zero_cost", source="cpp/tests/test_rule_dual_fix.cpp:159-222", kind="rule",
     lp=dict(m=1, n=1, A=csr([0, 1], [0], [1.0]), AT=csr([0, 1], [0], [1.0])),
     plan=dict(keep_col=[1], keep_row=[1], c=[0.0], l=[-INF], u=[10.0], AL=[-INF], AU=[5.0]),
     steps=[dict(rule="dual_fix", params={},
                 expect=[("eq", "plan.keep_col", [0]), ("eq", "plan.keep_row", [0]), ("flag", "has_row_action", 1),
                         ("flag", "has_col_action", 1), ("flag", "has_change", 1), ("eq", "tape.types", [FixedColInf]),
                         ("eq", "tape.index_starts", [0, 4]), ("eq", "tape.value_starts", [0, 4]),
                         ("eq", "tape.dual_modes", [EXACT]), ("eq", "tape.indices", [-1, 0, 1, 0]),
                         ("approx", "tape.vals", 0, 1.0), ("approx", "tape.vals", 1, 10.0),
                         ("approx", "tape.vals", 2, 5.0), ("approx", "tape.vals", 3, 1.0)])])

# ----------------------------------------------------- duplicate columns --
_dc_lp = dict(m=2, n=2, AT=csr([0, 2, 4], [0, 1, 0, 1], [1.0, 2.0, 2.0, 4.0]))
_dc_plan = dict(keep_row=[1, 1], keep_col=[1, 1], c=[1.0, 2.0], l=[0.0, 1.0], u=[10.0, 3.0], AL=[0.0, 0.0],
                AU=[10.0, 10.0])
case(name="duplicate_columns_merge", source="cpp/tests/test_rule_duplicate_columns.cpp:23-83", kind="rule",
     lp=_dc_lp, plan=_dc_plan,
     steps=[dict(rule="duplicate_columns", params={},
                 expect=[("eq", "plan.keep_col", [1, 0]), ("approx", "plan.l", 0, 2.0), ("approx", "plan.u", 0, 16.0),
                         ("flag", "has_col_action", 1), ("flag", "has_change", 1),
                         ("eq", "tape.types", [DuplicateColumn]), ("eq", "tape.index_starts", [0, 2]),
                         ("eq", "tape.value_starts", [0, 5]), ("eq", "tape.dual_modes", [MINIMAL]),
                         ("eq", "tape.indices", [1, 0]), ("approx", "tape.vals", 0, 2.0),
                         ("approx", "tape.vals", 1, 1.0), ("approx", "tape.vals", 2, 3.0),
                         ("approx", "tape.vals", 3, 0.0), ("approx", "tape.vals", 4, 10.0)])])
case(name="duplicate_columns_fixed", source="cpp/tests/test_rule_duplicate_columns.cpp:85-150", kind="rule",
     lp=dict(m=2, n=2, AT=csr([0, 2, 4], [0, 1, 0, 1], [1.0, 2.0, 1.0, 2.0])),
     plan=dict(keep_row=[1, 1], keep_col=[1, 1], c=[1.0, 2.0], l=[0.0, 3.0], u=[INF, 5.0], AL=[0.0, 0.0],
               AU=[10.0, 20.0]),
     steps=[dict(rule="duplicate_columns", params={},
                 expect=[("eq", "plan.keep_col", [1, 0]), ("approx", "plan.l", 1, 3.0), ("approx", "plan.u", 1, 3.0),
                         ("approx", "plan.AL", 0, -3.0), ("approx", "plan.AU", 0, 7.0),
                         ("approx", "plan.AL", 1, -6.0), ("approx", "plan.AU", 1, 14.0),
                         ("approx", "obj", None, 6.0), ("eq", "tape.types", [FixedCol]),
                         ("eq", "tape.index_starts", [0, 3]), ("eq", "tape.value_starts", [0, 4]),
                         ("eq", "tape.dual_modes", [MINIMAL]), ("eq", "tape.indices", [1, 0, 1]),
                         ("approx", "tape.vals", 0, 3.0), ("approx", "tape.vals", 1, 2.0),
                         ("approx", "tape.vals", 2, 1.0), ("approx", "tape.vals", 3, 2.0)])])
case(name="duplicate_columns_ratio_guard", source="cpp/tests/test_rule_duplicate_columns.cpp:152-197", kind="rule",
     lp=_dc_lp, plan=_dc_plan,
     steps=[dict(rule="duplicate_columns",
                 params=dict(duplicate_columns_max_abs_ratio=1.0, duplicate_columns_scale_guard_min_groups=1),
                 expect=[("eq", "plan.keep_col", [1, 1]), ("flag", "has_col_action", 0), ("flag", "has_change", 0),
                         ("len", "tape.types", 0)])])

# ------------------------------------------------------ redundant bounds --
_rb1 = dict(m=1, n=2, A=csr([0, 2], [0, 1], [1.0, 1.0]), AT=csr([0, 1, 2], [0, 0], [1.0, 1.0]))
case(name="redundant_bounds_upper", source="cpp/tests/test_rule_redundant_bounds.cpp:32-69", kind="rule", lp=_rb1,
     plan=dict(l=[-INF, 0.0], u=[5.0, 5.0], AL=[-INF], AU=[5.0]),
     steps=[dict(rule="redundant_bounds", params=dict(enable_redundant_bounds=1),
                 expect=[("approx", "plan.u", 0, INF), ("flag", "has_col_action", 1), ("flag", "has_change", 1)])])
case(name="redundant_bounds_lower", source="cpp/tests/test_rule_redundant_bounds.cpp:71-108", kind="rule", lp=_rb1,
     plan=dict(l=[0.0, 0.0], u=[INF, 0.0], AL=[0.0], AU=[INF]),
     steps=[dict(rule="redundant_bounds", params=dict(enable_redundant_bounds=1),
                 expect=[("approx", "plan.l", 0, -INF), ("flag", "has_col_action", 1), ("flag", "has_change", 1)])])
case(name="redundant_bounds_both_box", source="cpp/tests/test_rule_redundant_bounds.cpp:110-150", kind="rule",
     lp=dict(m=2, n=2, A=csr([0, 2, 4], [0, 1, 0, 1], [1.0] * 4), AT=csr([0, 2, 4], [0, 1, 0, 1], [1.0] * 4)),
     plan=dict(l=[0.0, 0.0], u=[5.0, 0.0], AL=[-INF, 0.0], AU=[5.0, INF]),
     steps=[dict(rule="redundant_bounds", params={},
                 expect=[("approx", "plan.l", 0, -INF), ("approx", "plan.u", 0, INF), ("flag", "has_col_action", 1),
                         ("flag", "has_change", 1)])])

# ------------------------------------------------------ full presolve ----
_off_all = dict(enable_infeasible_fixed_variables=0, enable_empty_rows=0, enable_singleton_rows=0,
                enable_infeasible_redundant_rows=0, enable_implied_variable_bounds=0, enable_duplicate_rows=0,
                enable_empty_cols=0, enable_column_singletons_eq=0, enable_column_singletons_dual_infer=0,
                enable_doubleton_equations=0, enable_dual_fix=0, enable_duplicate_columns=0,
                enable_redundant_bounds=0)


def _params(**kw):
    return dict(_off_all, **kw)


_closure_lp = dict(m=2, n=3, A=csr([0, 2, 5], [0, 1, 0, 1, 2], [1.0, 1.0, 1.0, 1.0, -1.0]),
                   AT=csr([0, 2, 4, 5], [0, 1, 0, 1, 1], [1.0, 1.0, 1.0, 1.0, -1.0]), c=[0.0, 0.0, 0.0],
                   l=[0.0, 0.0, 0.0], u=[10.0, 1.0, INF], AL=[-INF, -INF], AU=[2.0, 3.0])
for tiered, act in ((0, 0), (1, 1)):
    exp = [("eq", "summary.has_infeasible", 0), ("eq", "summary.has_unbounded", 0),
           ("eq", "summary.reduced_rows", 1), ("eq", "summary.reduced_cols", 3), ("eq", "summary.reduced_nnz", 2),
           ("eq", "record.row_red2org", [0]), ("len", "record.types", 2 if act else 1),
           ("one_of", [("approx", "record.types", 0, BoundChangeTheRow)],
            [("approx", "record.types", 0, BoundChangeNoRow)])]
    if act:
        exp += [("approx", "record.types", 1, DeletedRow),
                ("index_at", "record.indices", "record.index_starts", 1, 1)]
    case(name=f"post_propagation_activity_closure_tiered{tiered}", source="cpp/tests/test_gpu_direct_port.cpp:22-125",
         kind="presolve", lp=_closure_lp,
         params=dict(_params(enable_infeasible_redundant_rows=1, enable_implied_variable_bounds=1),
                     max_iters=2 if tiered else 1, use_tiered_scheduler=tiered, enable_tiered_bootstrap=0,
                     record_postsolve_tape=1, record_postsolve_tape_cpu=act),
         expect=exp,
         postsolve=dict(x_red=[1.0, 1.0, 0.0], y_red=[0.0], z_red=[0.0, 0.0, 0.0], original=True,
                        expect=[("approx", "x", 0, 1.0), ("approx", "x", 1, 1.0), ("approx", "x", 2, 0.0),
                                ("approx", "y", 0, 0.0), ("approx", "y", 1, 0.0), ("approx", "z", 0, 0.0),
                                ("approx", "z", 1, 0.0), ("approx", "z", 2, 0.0)]))
case(name="implied_bounds_dependency_chain_fixed_point", source="cpp/tests/test_gpu_direct_port.cpp:127-187",
     kind="presolve",
     lp=dict(m=2, n=4, A=csr([0, 2, 5], [0, 1, 1, 2, 3], [1.0, 1.0, -1.0, 1.0, 1.0]),
             AT=csr([0, 1, 3, 4, 5], [0, 0, 1, 1, 1], [1.0, 1.0, -1.0, 1.0, 1.0]), c=[0.0] * 4,
             l=[1.0, 0.0, 0.0, 0.0], u=[1.0, INF, INF, 0.0], AL=[-INF, -INF], AU=[1.0, 0.0]),
     params=dict(_params(enable_implied_variable_bounds=1), max_iters=1, use_tiered_scheduler=0,
                 implied_variable_bounds_bound_only_nnz_round_budget=1),
     expect=[("eq", "summary.has_infeasible", 0), ("eq", "summary.has_unbounded", 0),
             ("eq", "summary.reduced_rows", 2), ("eq", "summary.reduced_cols", 2), ("eq", "summary.reduced_nnz", 2)])
_chain = dict(m=2, n=3, A=csr([0, 2, 4], [0, 1, 1, 2], [1.0, 1.0, -1.0, 1.0]),
              AT=csr([0, 1, 3, 4], [0, 0, 1, 1], [1.0, 1.0, -1.0, 1.0]), c=[0.0] * 3, l=[1.0, 0.0, 0.0],
              u=[1.0, INF, INF], AL=[-INF, -INF], AU=[2.0, 0.0])
for budget, u2 in ((4, INF), (0, 1.0)):
    case(name=f"implied_bounds_bound_only_budget_{budget}", source="cpp/tests/test_gpu_direct_port.cpp:189-269",
         kind="presolve", lp=_chain,
         params=dict(_params(enable_implied_variable_bounds=1), max_iters=1, use_tiered_scheduler=0,
                     enable_tiered_bootstrap=0, record_postsolve_tape=0,
                     implied_variable_bounds_bound_only_nnz_round_budget=budget),
         expect=[("eq", "summary.has_infeasible", 0), ("eq", "summary.has_unbounded", 0),
                 ("eq", "summary.reduced_rows", 2), ("eq", "summary.reduced_cols", 3),
                 ("eq", "summary.reduced_nnz", 4), ("approx", "reduced.u", 0, 1.0), ("approx", "reduced.u", 1, 1.0),
                 ("approx", "reduced.u", 2, u2)])
_sched = dict(m=3, n=3, A=csr([0, 1, 2, 2], [0, 1], [1.0, 1.0]), AT=csr([0, 1, 2, 2], [0, 1], [1.0, 1.0]),
              c=[0.0, 0.0, 2.0], l=[5.0, 0.0, 1.0], u=[5.0, 10.0, 3.0], AL=[5.0, -INF, -1.0], AU=[5.0, INF, 1.0])
for tiered in (1, 0):
    case(name=f"scheduler_clears_small_lp_tiered{tiered}", source="cpp/tests/test_gpu_direct_port.cpp:289-361",
         kind="presolve", lp=_sched, params=dict(max_iters=4, use_tiered_scheduler=tiered),
         expect=[("eq", "summary.has_infeasible", 0), ("eq", "summary.has_unbounded", 0),
                 ("eq", "summary.reduced_rows", 0), ("eq", "summary.reduced_cols", 0)])
_dr = 300
case(name="scheduler_doubleton_only", source="cpp/tests/test_gpu_direct_port.cpp:362-429", kind="presolve",
     lp=dict(m=_dr, n=2 * _dr, A=csr([2 * r for r in range(_dr + 1)], list(range(2 * _dr)), [1.0] * (2 * _dr)),
             c=[0.0] * (2 * _dr), l=[0.0] * (2 * _dr), u=[10.0] * (2 * _dr), AL=[1.0] * _dr, AU=[1.0] * _dr),
     params=dict(_params(enable_doubleton_equations=1), max_iters=4),
     expect=[("eq", "summary.has_infeasible", 0), ("eq", "summary.has_unbounded", 0),
             ("eq", "summary.reduced_rows", 0), ("eq", "summary.reduced_cols", _dr),
             ("eq", "summary.reduced_nnz", 0), ("gt", "summary.iterations", 1)])

# ------------------------------------------------ presolve + postsolve ----
case(name="record_drives_fixed_column_postsolve", source="cpp/tests/test_gpu_postsolve.cpp:138-216",
     kind="presolve",
     lp=dict(m=1, n=2, A=csr([0, 2], [0, 1], [2.0, 3.0]), AT=csr([0, 1, 2], [0, 0], [2.0, 3.0]), c=[1.0, 2.0],
             l=[7.0, 0.0], u=[7.0, INF], AL=[10.0], AU=[10.0]),
     params=dict(_params(enable_infeasible_fixed_variables=1), max_iters=1, enable_tiered_bootstrap=0),
     expect=[("eq", "summary.reduced_rows", 1), ("eq", "summary.reduced_cols", 1), ("eq", "record.m0", 1),
             ("eq", "record.n0", 2), ("eq", "record.m1", 1), ("eq", "record.n1", 1), ("eq", "record.col_red2org", [1]),
             ("eq", "record.fixed_idx", [0]), ("approx", "record.fixed_val", 0, 7.0)],
     postsolve=dict(x_red=[4.0], y_red=[1.0], z_red=[0.5], original=True,
                    expect=[("eq", "x", [7.0, 4.0]), ("eq", "y", [1.0]), ("eq", "z", [-1.0, 0.0])]))
case(name="near_fixed_value_recorded", source="cpp/tests/test_gpu_postsolve.cpp:218-296", kind="presolve",
     lp=dict(m=1, n=2, A=csr([0, 2], [0, 1], [1.0, 1.0]), AT=csr([0, 1, 2], [0, 0], [1.0, 1.0]), c=[1.0, 0.0],
             l=[_lo, 0.0], u=[_hi, INF], AL=[0.0], AU=[10.0]),
     params=dict(_params(enable_infeasible_fixed_variables=1), max_iters=1, record_postsolve_tape_cpu=0,
                 enable_tiered_bootstrap=0),
     expect=[("eq", "summary.reduced_rows", 1), ("eq", "summary.reduced_cols", 1), ("eq", "record.col_red2org", [1]),
             ("eq", "record.fixed_idx", [0]), ("len", "record.fixed_val", 1),
             ("approx", "record.fixed_val", 0, _mid)],
     postsolve=dict(x_red=[4.0], y_red=None, z_red=None, original=False,
                    expect=[("approx", "x", 0, _mid), ("approx", "x", 1, 4.0)]))
case(name="empty_column_fixed_value_recorded", source="cpp/tests/test_gpu_postsolve.cpp:298-352", kind="presolve",
     lp=dict(m=0, n=1, A=csr([0], [], []), AT=csr([0, 0], [], []), c=[2.0], l=[3.0], u=[10.0]),
     params=dict(_params(enable_empty_cols=1), max_iters=1, enable_tiered_bootstrap=0),
     expect=[("eq", "summary.reduced_rows", 0), ("eq", "summary.reduced_cols", 0), ("eq", "record.fixed_idx", [0]),
             ("approx", "record.fixed_val", 0, 3.0)],
     postsolve=dict(x_red=None, y_red=None, z_red=None, original=False, expect=[("eq", "x", [3.0])]))
case(name="duplicate_column_globalized_tape", source="cpp/tests/test_gpu_postsolve.cpp:1073-1150", kind="presolve",
     lp=dict(m=2, n=2, A=csr([0, 2, 4], [0, 1, 0, 1], [1.0, 2.0, 2.0, 4.0]),
             AT=csr([0, 2, 4], [0, 1, 0, 1], [1.0, 2.0, 2.0, 4.0]), c=[1.0, 2.0], l=[0.0, 1.0], u=[10.0, 3.0],
             AL=[0.0, 0.0], AU=[10.0, 10.0]),
     params=dict(_params(enable_duplicate_columns=1), max_iters=2, enable_tiered_bootstrap=0),
     expect=[("eq", "summary.reduced_rows", 2), ("eq", "summary.reduced_cols", 1), ("len", "record.types", 1),
             ("eq", "record.types", [DuplicateColumn]), ("eq", "record.indices", [1, 0]),
             ("approx", "record.vals", 0, 2.0)],
     postsolve=dict(x_red=[5.0], y_red=[0.0, 0.0], z_red=[0.4], original=False,
                    expect=[("approx", "x", 0, 1.5), ("approx", "x", 1, 1.75), ("approx", "z", 0, 0.4),
                            ("approx", "z", 1, 0.8)]))


# --------------------------------------------------------- postsolve ----
def rec(m0, n0, m1, n1, row_red2org=(), col_red2org=(), fixed_idx=(), fixed_val=(), types=(), index_starts=(0,),
        value_starts=(0,), dual_modes=(), indices=(), vals=()):
    return dict(m0=m0, n0=n0, m1=m1, n1=n1, row_red2org=list(row_red2org), col_red2org=list(col_red2org),
                fixed_idx=list(fixed_idx), fixed_val=list(fixed_val), types=list(types),
                index_starts=list(index_starts), value_starts=list(value_starts), dual_modes=list(dual_modes),
                indices=list(indices), vals=[float(v) for v in vals])


case(name="postsolve_scatter_fixed_restore", source="cpp/tests/test_gpu_postsolve.cpp:24-59", kind="postsolve",
     record=rec(3, 4, 2, 2, [0, 2], [0, 2], [1, 3], [5.0, -2.0]), x_red=[10.0, 20.0], y_red=[1.5, -3.0],
     z_red=[0.25, 0.75], original=None,
     expect=[("eq", "x", [10.0, 5.0, 20.0, -2.0]), ("eq", "y", [1.5, 0.0, -3.0]), ("eq", "z", [0.25, 0.0, 0.75, 0.0])])
case(name="postsolve_null_duals", source="cpp/tests/test_gpu_postsolve.cpp:61-88", kind="postsolve",
     record=rec(1, 1, 1, 1, [0], [0]), x_red=[2.0], y_red=None, z_red=None, original=None,
     expect=[("eq", "x", [2.0]), ("eq", "y", [0.0]), ("eq", "z", [0.0])])
case(name="postsolve_unretrieved_fixed_dual", source="cpp/tests/test_gpu_postsolve.cpp:90-136", kind="postsolve",
     record=rec(1, 2, 1, 1, [0], [1], [0], [2.0]), x_red=[7.0], y_red=[3.0], z_red=[1.5],
     original=dict(m=1, n=2, AT=csr([0, 1, 1], [0], [2.0]), c=[10.0, 5.0]),
     expect=[("eq", "x", [2.0, 7.0]), ("approx", "z", 0, 0.0), ("approx", "z", 1, 1.5)])
case(name="postsolve_duplicate_column_replay", source="cpp/tests/test_gpu_postsolve.cpp:354-388", kind="postsolve",
     record=rec(0, 2, 0, 1, [], [0], types=[DuplicateColumn], index_starts=[0, 2], value_starts=[0, 5],
                dual_modes=[MINIMAL], indices=[1, 0], vals=[2.0, 0.0, 4.0, 1.0, 1.0]),
     x_red=[5.0], y_red=None, z_red=[0.5], original=None,
     expect=[("approx", "x", 0, 1.0), ("approx", "x", 1, 2.0), ("approx", "z", 0, 0.5), ("approx", "z", 1, 1.0)])
case(name="postsolve_doubleton_replay", source="cpp/tests/test_gpu_postsolve.cpp:390-427", kind="postsolve",
     record=rec(1, 2, 0, 1, [], [1], types=[DoubletonEquation], index_starts=[0, 4], value_starts=[0, 12],
                dual_modes=[MINIMAL], indices=[0, 1, 0, 0],
                vals=[2.0, 3.0, 8.0, 0.0, 10.0, 0.0, 10.0, 0.0, 10.0, 4.0, 0.0, 0.0]),
     x_red=[2.0], y_red=None, z_red=[0.5], original=None,
     expect=[("approx", "x", 0, 1.0), ("approx", "x", 1, 2.0), ("approx", "y", 0, 2.0), ("approx", "z", 0, 0.0),
             ("approx", "z", 1, 0.5)])
case(name="postsolve_doubleton_original_model", source="cpp/tests/test_gpu_postsolve.cpp:429-498",
     kind="postsolve",
     record=rec(2, 2, 1, 1, [1], [1], types=[DoubletonEquation], index_starts=[0, 4], value_starts=[0, 12],
                dual_modes=[MINIMAL], indices=[0, 1, 0, 0],
                vals=[2.0, -2.0, 0.0, 0.0, INF, 0.0, 10.0, 5.0, 10.0, 10.0, 2.0, 0.0]),
     x_red=[5.0], y_red=[1.0], z_red=[4.0],
     original=dict(m=2, n=2, A=csr([0, 2, 3], [0, 1, 0], [2.0, -2.0, 3.0]), AT=csr([0, 2, 3], [0, 1, 0],
                   [2.0, 3.0, -2.0]), c=[10.0, 0.0], l=[0.0, 5.0], u=[INF, 10.0], AL=[0.0, -INF], AU=[0.0, INF]),
     expect=[("approx", "x", 0, 5.0), ("approx", "x", 1, 5.0), ("approx", "y", 0, 3.0), ("approx", "y", 1, 1.0),
             ("approx", "z", 0, 4.0), ("approx", "z", 1, 0.0)])
case(name="postsolve_subcol_replay", source="cpp/tests/test_gpu_postsolve.cpp:500-535", kind="postsolve",
     record=rec(1, 3, 0, 2, [], [1, 2], types=[SubCol], index_starts=[0, 5], value_starts=[0, 8],
                dual_modes=[MINIMAL], indices=[0, 0, 2, 1, 2], vals=[2.0, 13.0, -10.0, 10.0, 6.0, 1.0, 1.0, 3.0]),
     x_red=[2.0, 3.0], y_red=None, z_red=[0.5, 0.7], original=None,
     expect=[("approx", "x", 0, 1.0), ("approx", "y", 0, 3.0), ("approx", "z", 0, 0.0)])
case(name="postsolve_fixed_col_reduced_cost", source="cpp/tests/test_gpu_postsolve.cpp:537-572", kind="postsolve",
     record=rec(1, 2, 1, 1, [0], [1], types=[FixedCol], index_starts=[0, 2], value_starts=[0, 3],
                dual_modes=[MINIMAL], indices=[0, 0], vals=[4.0, 10.0, 2.0]),
     x_red=[8.0], y_red=[3.0], z_red=[0.5], original=None,
     expect=[("approx", "x", 0, 4.0), ("approx", "z", 0, 4.0)])
case(name="postsolve_fme_and_eq_to_ineq", source="cpp/tests/test_gpu_postsolve.cpp:574-614", kind="postsolve",
     record=rec(2, 2, 1, 1, [1], [1], types=[FmeCol, EqToIneq], index_starts=[0, 5, 6], value_starts=[0, 5, 6],
                dual_modes=[MINIMAL, MINIMAL], indices=[1, 0, 0, 1, 1, 1], vals=[1.0, 1.0, 1.0, 3.0, 2.0, 1.25]),
     x_red=[2.0], y_red=[0.75], z_red=[0.5], original=None,
     expect=[("approx", "x", 0, 7.0), ("approx", "y", 0, 0.0), ("approx", "y", 1, 2.0), ("approx", "z", 0, 0.0)])
case(name="postsolve_duplicate_row_split", source="cpp/tests/test_gpu_postsolve.cpp:616-678", kind="postsolve",
     record=rec(2, 1, 1, 1, [0], [0], types=[DuplicateRow], index_starts=[0, 2], value_starts=[0, 5],
                dual_modes=[MINIMAL], indices=[0, 1], vals=[1.0, 0.0, 10.0, 2.0, INF]),
     x_red=[2.0], y_red=[5.0], z_red=[0.0],
     original=dict(m=2, n=1, A=csr([0, 1, 2], [0, 0], [1.0, 1.0]), AT=csr([0, 2], [0, 1], [1.0, 1.0]), c=[0.0],
                   l=[0.0], u=[INF], AL=[0.0, 2.0], AU=[10.0, INF]),
     expect=[("approx", "y", 0, 0.0), ("approx", "y", 1, 5.0)])
case(name="postsolve_deleted_row_original_model", source="cpp/tests/test_gpu_postsolve.cpp:680-739",
     kind="postsolve",
     record=rec(1, 1, 0, 1, [], [0], types=[DeletedRow], index_starts=[0, 2], value_starts=[0, 3],
                dual_modes=[MINIMAL], indices=[0, 0], vals=[4.0, 4.0, 2.0]),
     x_red=[2.0], y_red=None, z_red=[0.0],
     original=dict(m=1, n=1, A=csr([0, 1], [0], [2.0]), AT=csr([0, 1], [0], [2.0]), c=[6.0], l=[0.0], u=[10.0],
                   AL=[4.0], AU=[4.0]),
     expect=[("approx", "y", 0, 3.0), ("approx", "z", 0, 0.0)])
case(name="postsolve_row_bound_change_refines_duals", source="cpp/tests/test_gpu_postsolve.cpp:741-804",
     kind="postsolve",
     record=rec(1, 1, 1, 1, [0], [0], types=[BoundChangeTheRow], index_starts=[0, 2], value_starts=[0, 4],
                dual_modes=[MINIMAL], indices=[0, 0], vals=[0.0, 10.0, 2.0, 10.0]),
     x_red=[2.0], y_red=[0.0], z_red=[6.0],
     original=dict(m=1, n=1, A=csr([0, 1], [0], [2.0]), AT=csr([0, 1], [0], [2.0]), c=[6.0], l=[0.0], u=[10.0],
                   AL=[4.0], AU=[INF]),
     expect=[("approx", "y", 0, 3.0), ("approx", "z", 0, 0.0)])
case(name="postsolve_bound_change_no_row_projection", source="cpp/tests/test_gpu_postsolve.cpp:806-866",
     kind="postsolve",
     record=rec(1, 1, 1, 1, [0], [0], types=[BoundChangeNoRow], index_starts=[0, 1], value_starts=[0, 0],
                dual_modes=[MINIMAL], indices=[0], vals=[]),
     x_red=[0.0], y_red=[1.0], z_red=[99.0],
     original=dict(m=1, n=1, A=csr([0, 1], [0], [2.0]), AT=csr([0, 1], [0], [2.0]), c=[5.0], l=[0.0], u=[INF],
                   AL=[-INF], AU=[INF]),
     expect=[("approx", "z", 0, 3.0)])


if __name__ == "__main__":
    path = os.path.join(HERE, "reference_kats.json")
    with open(path, "w") as f:
        json.dump({"cases": CASES}, f, indent=1)
    print(f"wrote {len(CASES)} cases to {path}")
