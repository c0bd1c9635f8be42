"""Fail-closed screens of the structure-specific rules, on the LPs of the
reference's own tests for those rules:

  covering cost dominance   cpp/tests/test_rule_covering_cost_dominance.cpp:55-191
  Linf components           cpp/tests/test_rule_linf_components.cpp:58-183, 299-314
  antipodal components      cpp/tests/test_rule_antipodal_components.cpp:45-175
  bounded two-row / orphan McCormick projection
                            cpp/tests/test_rule_projected_auxiliaries.cpp:122-294
  structural L1 (outer pair block)
                            cpp/tests/test_rule_structural_l1_substitution.cpp:92-128

Where the reference's rule acts, presolve must stop with GP_ERR_UNSUPPORTED
(status 5); where the reference's detector rejects the LP (its fail-closed
variants), or the rule is switched off or left to a scheduler that does not
run it, presolve must run and give the oracle's record.  The covering,
McCormick and bounded projection screens evaluate the rule's own acceptance
test, so they refuse exactly the LPs the rule would change.  The Linf and
antipodal screens evaluate the rule's sampled probes only: LPs that pass the
probes but fail the rule's later validation (the fail-closed variants of
those two tests) are refused as well, which these tests pin."""
import numpy as np
import pytest

import gp_ffi
import lp_synth
import oracle_ffi as O
from helpers import assert_same_record

pytestmark = pytest.mark.gpu

INF = float("inf")
GENERIC_OFF = dict(enable_infeasible_fixed_variables=0, enable_empty_rows=0, enable_singleton_rows=0,
                   enable_infeasible_redundant_rows=0, enable_implied_variable_bounds=0, enable_duplicate_rows=0,
                   enable_empty_cols=0, enable_column_singletons_eq=0, enable_column_singletons_dual_infer=0,
                   enable_doubleton_equations=0, enable_dual_fix=0, enable_duplicate_columns=0,
                   enable_redundant_bounds=0)
STRUCT_OFF = dict(enable_covering_cost_dominance=0, enable_linf_components=0, enable_antipodal_components=0,
                  enable_bounded_two_row_projection=0, enable_orphan_mccormick_projection=0,
                  enable_structural_l1_substitution=0)


def _lp(rows, n, c, AL, AU, l, u):
    """LP dict from rows given as lists of (column, value)."""
    m = len(rows)
    ptr = np.zeros(m + 1, np.int32)
    idx, val = [], []
    for i, r in enumerate(rows):
        for j, v in r:
            idx.append(j)
            val.append(v)
        ptr[i + 1] = len(idx)
    idx, val = np.array(idx, np.int32), np.array(val, np.float64)
    tp, ti, tv = lp_synth.transpose_csr(m, n, ptr, idx, val)
    f = lambda a: np.array(a, np.float64)  # noqa: E731
    return {"m": m, "n": n, "A_ptr": ptr, "A_idx": idx, "A_val": val, "AT_ptr": tp, "AT_idx": ti, "AT_val": tv,
            "c": f(c), "l": f(l), "u": f(u), "AL": f(AL), "AU": f(AU), "obj": 0.0}


def _refused(lp, **kw):
    with pytest.raises(gp_ffi.GpError, match="gp status 5"):
        gp_ffi.presolve(lp, gp_ffi.params(**kw), True)


def _runs_like_oracle(lp, **kw):
    s_gpu, rec_gpu, _ = gp_ffi.presolve(lp, gp_ffi.params(**kw), True)
    s_cpu, rec_cpu, _ = O.presolve(lp, O.params(**kw), True)
    for k in ("reduced_rows", "reduced_cols", "reduced_nnz", "iterations"):
        assert s_gpu[k] == s_cpu[k], k
    assert_same_record(rec_gpu, rec_cpu)


# ---- covering cost dominance ------------------------------------------------
def _covering(variant):
    a = [1.0] * 11
    costs = [1.0, 1.0, 1.0, 3.0, 3.0, 4.0]
    AU = [INF, INF, INF]
    upper = [1.0] * 6
    if variant == "non_unit":
        a[1] = 0.5
    elif variant == "non_covering_row":
        AU[0] = 1.0
    elif variant == "missing_witness":
        costs[2] = 2.0
    elif variant == "non_binary":
        upper[5] = 2.0
    cols = [[0, 3, 4, 5], [1, 3, 4, 5], [2, 4, 5]]
    rows, k = [], 0
    for r in cols:
        rows.append([(j, a[k + q]) for q, j in enumerate(r)])
        k += len(r)
    return _lp(rows, 6, costs, [1.0] * 3, AU, [0.0] * 6, upper)


COVERING = dict(GENERIC_OFF, max_iters=1, enable_antipodal_components=0, enable_structural_l1_substitution=0,
                covering_cost_dominance_min_fixed_cols=1, covering_cost_dominance_min_candidate_ratio=0.2)


def test_covering_valid_is_refused():
    _refused(_covering("valid"), **COVERING)


@pytest.mark.parametrize("variant", ["non_unit", "non_covering_row", "missing_witness", "non_binary"])
def test_covering_rejected_variants_run(variant):
    _runs_like_oracle(_covering(variant), **COVERING)


@pytest.mark.parametrize("off", [dict(enable_tiered_bootstrap=0), dict(enable_structure_specific_rules=0),
                                 dict(enable_covering_cost_dominance=0), dict(use_tiered_scheduler=0)])
def test_covering_not_scheduled_runs(off):
    _runs_like_oracle(_covering("valid"), **dict(COVERING, **off))


def test_covering_ratio_gate():
    """Two of six columns are dominated (cost above the row count): a 0.5
    ratio gate turns the rule off."""
    _runs_like_oracle(_covering("valid"), **dict(COVERING, covering_cost_dominance_min_candidate_ratio=0.5))


# ---- Linf components --------------------------------------------------------
def _linf(wrong_tail=False, wrong_coefficient=False):
    n = 8 if wrong_tail else 7
    t2 = 7 if wrong_tail else 6
    k = -3.0 if wrong_coefficient else -2.0
    rows = [[(0, -2.0), (3, -2.0), (6, 1.0)], [(1, k), (4, k), (t2, 1.0)], [(2, -3.0), (5, -3.0), (6, 1.0)],
            [(0, 1.0), (1, -1.0), (3, -1.0), (4, 1.0)], [(0, 2.0), (1, 3.0), (2, -1.0), (3, -2.0), (4, -3.0), (5, 1.0)]]
    c = [0.0] * n
    c[6] = 1.0
    lo = [2.0, 2.0, 4.0, 1.0, 1.0, 1.0, 0.0]
    up = [10.0, 7.0, 9.0, 6.0, 4.0, 8.0, INF]
    if wrong_tail:
        c[7] = 1.0
        lo.append(0.0)
        up.append(INF)
    return _lp(rows, n, c, [5.0, 5.0, 7.0, 0.0, 10.0], [INF, INF, INF, 0.0, 20.0], lo, up)


LINF = dict(GENERIC_OFF, max_iters=1, enable_covering_cost_dominance=0, enable_antipodal_components=0,
            enable_bounded_two_row_projection=0, enable_orphan_mccormick_projection=0,
            enable_structural_l1_substitution=0, linf_min_pairs=1, linf_probe_pairs=3)


@pytest.mark.parametrize("wrong_tail,wrong_coefficient", [(False, False), (True, False), (False, True)])
def test_linf_probe_accepting_lps_are_refused(wrong_tail, wrong_coefficient):
    _refused(_linf(wrong_tail, wrong_coefficient), **LINF)


def test_linf_layout_gate():
    """Four pairs are required: the layout screen rejects the three-pair LP."""
    _runs_like_oracle(_linf(), **dict(LINF, linf_min_pairs=4))


# ---- antipodal components ---------------------------------------------------
def _antipodal(conflicting_shift=False):
    rows = [[(0, 1.0), (1, -1.0), (3, -1.0), (4, 1.0)],
            [(0, 1.0), (1, 1.0), (2, -1.0), (3, -1.0), (4, -1.0), (5, 1.0)]]
    lo = [2.0, 4.0, 0.0, 1.0, 2.0 if conflicting_shift else 3.0, 0.0]
    return _lp(rows, 6, [1.0, 2.0, 3.0, 1.0, 2.0, 3.0], [0.0, 5.0], [0.0, 5.0], lo,
               [10.0, 9.0, 6.0, 8.0, 8.0, 7.0])


ANTIPODAL = dict(GENERIC_OFF, max_iters=1, enable_covering_cost_dominance=0, enable_structural_l1_substitution=0,
                 antipodal_min_pairs=1, antipodal_probe_pairs=3)


@pytest.mark.parametrize("conflicting_shift", [False, True])
def test_antipodal_probe_accepting_lps_are_refused(conflicting_shift):
    _refused(_antipodal(conflicting_shift), **ANTIPODAL)


def test_antipodal_owned_by_tiered_bootstrap():
    _runs_like_oracle(_antipodal(), **dict(ANTIPODAL, enable_tiered_bootstrap=0))


def test_antipodal_unmatched_mate_runs():
    """Column 3 no longer mirrors column 0 (cost 2 instead of 1): a sampled
    pair fails and the screen lets the LP through."""
    lp = _antipodal()
    lp["c"][3] = 2.0
    _runs_like_oracle(lp, **ANTIPODAL)


# ---- bounded two-row and orphan McCormick projections ------------------------
def _pair_block(rhs=4.0, recourse_cost=0.0):
    rows = [[(0, -1.0), (1, -1.0), (4, 1.0)], [(0, -1.0), (4, 1.0)], [(1, -1.0), (4, 1.0)],
            [(2, 1.0), (3, -1.0), (4, 2.0), (5, -2.0)], [(2, -1.0), (3, 1.0), (4, 2.0), (5, 2.0)]]
    return _lp(rows, 6, [-1.0, -2.0, 0.0, 0.0, 0.0, recourse_cost], [-1.0, -INF, -INF, -INF, -INF],
               [INF, 0.0, 0.0, 2.0, rhs], [0.0] * 6, [1.0] * 6)


def _mccormick(corrupt_upper_row=False, nonunit_endpoint=False):
    rows = [[(0, -1.0), (1, -1.0), (2, 1.0)], [(0, -0.5 if corrupt_upper_row else -1.0), (2, 1.0)],
            [(1, -1.0), (2, 1.0)]]
    return _lp(rows, 3, [0.0] * 3, [-1.0, -INF, -INF], [INF, 0.0, 0.0], [0.0] * 3,
               [1.0, 2.0 if nonunit_endpoint else 1.0, 1.0])


PROJECTED = dict(GENERIC_OFF, max_iters=1, enable_covering_cost_dominance=0, enable_antipodal_components=0,
                 enable_structural_l1_substitution=0, enable_zero_cost_redundant_box_bounds=0,
                 bounded_two_row_projection_min_candidates=1, bounded_two_row_projection_min_candidate_ratio=0.0,
                 orphan_mccormick_projection_min_candidates=1, orphan_mccormick_projection_min_candidate_ratio=0.0)


def test_bounded_projection_block_is_refused():
    _refused(_pair_block(), **dict(PROJECTED, enable_orphan_mccormick_projection=0))


@pytest.mark.parametrize("rhs,cost", [(3.0, 0.0), (4.0, 0.25)])
def test_bounded_projection_rejected_variants_run(rhs, cost):
    _runs_like_oracle(_pair_block(rhs, cost), **dict(PROJECTED, enable_orphan_mccormick_projection=0))


def test_projections_owned_by_tiered_bootstrap():
    _runs_like_oracle(_pair_block(), **dict(PROJECTED, enable_tiered_bootstrap=0))


def test_mccormick_exact_is_refused():
    _refused(_mccormick(), **dict(PROJECTED, enable_bounded_two_row_projection=0))


@pytest.mark.parametrize("corrupt,nonunit", [(True, False), (False, True)])
def test_mccormick_rejected_variants_run(corrupt, nonunit):
    _runs_like_oracle(_mccormick(corrupt, nonunit), **dict(PROJECTED, enable_bounded_two_row_projection=0))


# ---- structural L1 -----------------------------------------------------------
def _outer(break_signature=False):
    v0 = 2.0 if break_signature else 1.0
    rows = [[(0, v0), (1, -1.0)], [(0, 1.0), (1, 1.0)], [(0, -1.0), (1, 1.0)],
            [(0, 1.0), (2, -1.0)], [(0, 1.0), (2, 1.0)], [(0, -1.0), (2, 1.0)],
            [(3, 1.0)], [(3, -1.0), (4, 1.0)], [(3, 1.0), (4, 1.0)], [(4, 1.0), (5, -1.0)],
            [(5, -1.0), (6, 1.0)], [(5, -1.0), (7, 1.0)], [(5, -1.0), (8, 1.0)], [(5, -1.0), (9, 1.0)]]
    AU = [INF, INF, INF, INF, INF, INF, 0.0, INF, INF, 0.0, INF, INF, INF, INF]
    return _lp(rows, 10, [1.0, 2.0, 3.0, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 1.0], [0.0] * 14, AU,
               [0.0, -INF, -INF, -INF, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0], [10.0] + [INF] * 9)


L1 = dict(GENERIC_OFF, max_iters=1, **dict(STRUCT_OFF, enable_structural_l1_substitution=1))


def test_structural_outer_block_is_refused():
    _refused(_outer(), **L1)


def test_structural_broken_signature_runs():
    _runs_like_oracle(_outer(True), **L1)


# ---- LPs without those structures -------------------------------------------
@pytest.mark.parametrize("seed", [0, 1])
def test_general_lps_pass_every_screen(seed):
    """The synthetic LPs of the test workloads have none of the structures:
    with every structure rule switched on (the reference's defaults), the
    presolve runs and equals the oracle's."""
    _runs_like_oracle(lp_synth.config0(seed=seed))
