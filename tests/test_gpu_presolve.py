"""Full presolve and postsolve on the GPU against the oracle: summary,
record (maps, fixed values, tape) and reduced LP bit-exact on dyadic inputs;
postsolve (x, y, z) within 1e-12."""
import numpy as np
import pytest

import gp_ffi
import lp_synth
import oracle_ffi as O
from helpers import assert_same_lp, assert_same_record

pytestmark = pytest.mark.gpu

SUMMARY_KEYS = ("original_rows", "original_cols", "reduced_rows", "reduced_cols", "reduced_nnz", "iterations",
                "has_infeasible", "has_unbounded")
TOL = 1e-12


def _run_both(lp, **kw):
    sg, rg, lg = gp_ffi.presolve(lp, gp_ffi.params(**kw), True)
    so, ro, lo = O.presolve(lp, O.params(**kw), True)
    for k in SUMMARY_KEYS:
        assert sg[k] == so[k], f"summary {k}: gpu {sg[k]} oracle {so[k]}"
    assert sg["obj_constant_delta"] == so["obj_constant_delta"]
    assert_same_record(rg, ro)
    assert_same_lp(lg, lo)
    return so, ro, lo


def _postsolve_both(lp, rec, red, seed):
    x, y, z = lp_synth.reduced_solution(seed, red["m"], red["n"], red["l"], red["u"])
    for original in (lp, None):
        got = gp_ffi.postsolve(rec, x, y, z, original)
        want = O.postsolve(rec, x, y, z, original)
        for name, a, b in zip("xyz", got, want):
            np.testing.assert_allclose(a, b, rtol=TOL, atol=TOL, err_msg=f"postsolve {name}")


CASES = {
    "c0": (lambda: lp_synth.config0(seed=0), {}),
    "c0_fixed_sched": (lambda: lp_synth.config0(seed=0), dict(use_tiered_scheduler=0)),
    "c0_tape_cpu": (lambda: lp_synth.config0(seed=0), dict(record_postsolve_tape_cpu=1)),
    "c0_dbl_eager": (lambda: lp_synth.config0(seed=0), dict(doubleton_equations_min_selected_per_batch=1,
                                                            doubleton_equations_min_selected_ratio=0.0)),
    "c0_dbl_scalar": (lambda: lp_synth.config0(seed=0), dict(doubleton_equations_scan=0)),
    "c0_redundant": (lambda: lp_synth.config0(seed=0), dict(enable_redundant_bounds=1)),
    "c0_zero_cost_box": (lambda: lp_synth.config0(seed=0), dict(enable_zero_cost_redundant_box_bounds=1)),
    "c0_no_tape": (lambda: lp_synth.config0(seed=0), dict(record_postsolve_tape=0)),
    "r_zipf": (lambda: lp_synth.generate(20000, 30000, seed=5, row_len=(2, 14), zipf=1.5), {}),
    "r_dense_eq": (lambda: lp_synth.generate(3000, 4000, seed=6, mean_row_len=3.0, frac_doubleton=0.2,
                                            frac_eq=0.6), {}),
    "r_small": (lambda: lp_synth.generate(50, 70, seed=8, mean_row_len=4.0), {}),
    "c2_small": (lambda: lp_synth.config2(seed=2, blocks=100, link_rows=10, link_nnz=1500), {}),
}


@pytest.mark.parametrize("case", list(CASES))
def test_presolve_postsolve(case):
    make, kw = CASES[case]
    lp = make()
    kw = dict(kw, enable_structure_specific_rules=0)
    so, ro, lo = _run_both(lp, **kw)
    _postsolve_both(lp, ro, lo, seed=3)


@pytest.fixture
def closure_always():
    """Batched singleton closure on every fast phase, in both libraries."""
    before = (gp_ffi.batched_closure_min_nnz(), O.batched_closure_min_nnz())
    gp_ffi.set_batched_closure_min_nnz(0)
    O.set_batched_closure_min_nnz(0)
    yield
    gp_ffi.set_batched_closure_min_nnz(before[0])
    O.set_batched_closure_min_nnz(before[1])


CLOSURE_CASES = {
    "c0": (lambda: lp_synth.config0(seed=0), {}),
    "c0_no_bootstrap": (lambda: lp_synth.config0(seed=0), dict(enable_tiered_bootstrap=0)),
    "c0_no_bootstrap_tape_cpu": (lambda: lp_synth.config0(seed=0),
                                 dict(enable_tiered_bootstrap=0, record_postsolve_tape_cpu=1)),
    "r_zipf_no_bootstrap": (lambda: lp_synth.generate(20000, 30000, seed=5, row_len=(2, 14), zipf=1.5),
                            dict(enable_tiered_bootstrap=0)),
    "r_dense_eq_no_bootstrap": (lambda: lp_synth.generate(3000, 4000, seed=6, mean_row_len=3.0, frac_doubleton=0.2,
                                                          frac_eq=0.6), dict(enable_tiered_bootstrap=0)),
    "r_long_rows": (lambda: lp_synth.generate(20000, 30000, seed=9, row_len=(2, 64), zipf=1.5), {}),
}


@pytest.mark.parametrize("case", list(CLOSURE_CASES))
def test_batched_singleton_closure(case, closure_always):
    """The fast phase's batched singleton closure (the reference runs it from
    10M nnz; here forced on at every size) against the oracle's."""
    make, kw = CLOSURE_CASES[case]
    lp = make()
    so, ro, lo = _run_both(lp, enable_structure_specific_rules=0, **kw)
    _postsolve_both(lp, ro, lo, seed=5)


def test_structure_switch_is_identical():
    lp = lp_synth.config0(seed=0)
    a = gp_ffi.presolve(lp, gp_ffi.params(enable_structure_specific_rules=1), True)
    b = gp_ffi.presolve(lp, gp_ffi.params(enable_structure_specific_rules=0), True)
    assert_same_record(a[1], b[1])
    assert_same_lp(a[2], b[2])


def test_empty_lp():
    lp = lp_synth.generate(0, 0, seed=1)
    _run_both(lp)


def test_config2_full_size():
    """The block-structured LP (2,000 blocks of 40 x 60 plus 200 linking rows of
    3,000 nnz: 80,200 x 120,000, ~1.3M nnz) with the structure-specific rules
    off, end to end."""
    lp = lp_synth.config2(seed=2)
    so, ro, lo = _run_both(lp, enable_structure_specific_rules=0)
    _postsolve_both(lp, ro, lo, seed=5)


def test_config1_full_size():
    """The single large LP workload (1M x 1.5M, ~7.5M nnz) end to end."""
    lp = lp_synth.config1(seed=1)
    so, ro, lo = _run_both(lp, enable_structure_specific_rules=0)
    _postsolve_both(lp, ro, lo, seed=4)
