"""A phase whose predecessor changed nothing takes over that phase's plan
and statistics instead of building them again (the reference builds a fresh
plan per phase, gpu_presolve.cu:2219-2452; a plan that no rule touched is a
fresh plan of the same LP).  With the checking hook on, every reused pair is
compared bit for bit with a freshly built one inside the presolve, and the
records must still equal the oracle's."""
import pytest

import gp_ffi
import lp_synth
import oracle_ffi as O
from helpers import assert_same_lp, assert_same_record

pytestmark = pytest.mark.gpu

CASES = {
    "c0": (lambda: lp_synth.config0(seed=0), {}),
    "c0_fixed_sched": (lambda: lp_synth.config0(seed=0), dict(use_tiered_scheduler=0)),
    "c0_redundant": (lambda: lp_synth.config0(seed=0), dict(enable_redundant_bounds=1)),
    "c0_tape_cpu": (lambda: lp_synth.config0(seed=0), dict(record_postsolve_tape_cpu=1)),
    "r_zipf": (lambda: lp_synth.generate(20000, 30000, seed=5, row_len=(2, 14), zipf=1.5), {}),
    "r_dense_eq": (lambda: lp_synth.generate(3000, 4000, seed=6, mean_row_len=3.0, frac_doubleton=0.2,
                                            frac_eq=0.6), {}),
    "r_long_dup": (lambda: lp_synth.generate(600, 3000, seed=21, mean_row_len=40.0, min_row_len=33,
                                            frac_dup_rows=0.05), {}),
    "c2_small": (lambda: lp_synth.config2(seed=2, blocks=100, link_rows=10, link_nnz=1500), {}),
}


@pytest.fixture
def checked_reuse():
    gp_ffi.check_plan_reuse(True)
    yield
    gp_ffi.check_plan_reuse(False)


@pytest.mark.parametrize("case", list(CASES))
def test_reused_plans_equal_fresh_plans(case, checked_reuse):
    make, kw = CASES[case]
    lp = make()
    kw = dict(kw, enable_structure_specific_rules=0)
    before = gp_ffi.plan_reuses()
    sg, rg, lg = gp_ffi.presolve(lp, gp_ffi.params(**kw), True)
    reused = gp_ffi.plan_reuses() - before
    so, ro, lo = O.presolve(lp, O.params(**kw), True)
    assert (sg["reduced_rows"], sg["reduced_cols"], sg["reduced_nnz"]) == \
           (so["reduced_rows"], so["reduced_cols"], so["reduced_nnz"])
    assert_same_record(rg, ro)
    assert_same_lp(lg, lo)
    assert reused > 0, "no phase reused its predecessor's plan"
