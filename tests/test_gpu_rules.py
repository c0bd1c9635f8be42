"""Each rule of the rule operator API on the GPU against the oracle, called
directly on phase plans as the reference's rule tests do.  Plans, flags,
tapes and rebuilt matrices must agree bit for bit (dyadic inputs)."""
import copy

import numpy as np
import pytest

import gp_ffi
import lp_synth
import oracle_ffi as O
from helpers import assert_same_array, fresh_plan, stats

pytestmark = pytest.mark.gpu

ROW_RULES = ["empty_rows", "singleton_rows", "infeasible_redundant_rows", "implied_variable_bounds",
             "duplicate_rows"]
COL_RULES = ["infeasible_fixed_variables", "empty_cols", "column_singletons_dual_infer", "column_singletons_eq",
             "doubleton_equations", "dual_fix", "duplicate_columns", "redundant_bounds"]

LPS = {
    "c0": lambda: lp_synth.config0(seed=0),
    "r1": lambda: lp_synth.generate(400, 600, seed=21, mean_row_len=5.0),
    "r2": lambda: lp_synth.generate(3000, 4500, seed=22, row_len=(2, 14), zipf=1.5),
    "r3": lambda: lp_synth.generate(1500, 2000, seed=23, mean_row_len=3.0, frac_doubleton=0.2,
                                     frac_eq=0.6),
}
_cache = {}


def get_lp(name):
    if name not in _cache:
        _cache[name] = LPS[name]()
    return _cache[name]


PARAM_SETS = {
    "default": dict(enable_structure_specific_rules=0),
    "tape_cpu": dict(enable_structure_specific_rules=0, record_postsolve_tape_cpu=1),
    "dbl_eager": dict(enable_structure_specific_rules=0, doubleton_equations_min_selected_per_batch=1,
                      doubleton_equations_min_selected_ratio=0.0),
    "dbl_scalar": dict(enable_structure_specific_rules=0, doubleton_equations_scan=0),
    "rb_one_sided": dict(enable_structure_specific_rules=0, enable_redundant_bounds=1),
}


def _compare(tag, pg, po, fg, fo, tg, to, ag, ao):
    for k in ("keep_row", "keep_col", "c", "l", "u", "AL", "AU"):
        assert_same_array(f"{tag} plan {k}", pg[k], po[k])
    for k in ("has_change", "has_row_action", "has_col_action", "has_infeasible", "has_unbounded", "has_new_A"):
        assert fg[k] == fo[k], f"{tag} flag {k}: gpu {fg[k]} oracle {fo[k]}"
    assert fg["obj_constant_delta"] == fo["obj_constant_delta"], \
        f"{tag} obj delta {fg['obj_constant_delta']!r} vs {fo['obj_constant_delta']!r}"
    for k in ("types", "index_starts", "value_starts", "dual_modes", "indices", "vals"):
        assert_same_array(f"{tag} tape {k}", tg[k], to[k])
    assert (ag is None) == (ao is None), f"{tag} new_A presence"
    if ag is not None:
        for name, a, b in zip(("ptr", "idx", "val"), ag, ao):
            assert_same_array(f"{tag} new_A {name}", a, b)


@pytest.mark.parametrize("lp_name", list(LPS))
@pytest.mark.parametrize("pset", list(PARAM_SETS))
@pytest.mark.parametrize("rule", ROW_RULES + COL_RULES)
def test_rule_fresh_plan(rule, pset, lp_name):
    lp = get_lp(lp_name)
    st = stats(lp)
    pg, po = fresh_plan(lp), fresh_plan(lp)
    fg, tg, ag = gp_ffi.rule_apply(rule, pg, lp, st, gp_ffi.params(**PARAM_SETS[pset]))
    fo, to, ao = O.rule_apply(rule, po, lp, st, O.params(**PARAM_SETS[pset]))
    _compare(f"{rule}/{pset}/{lp_name}", pg, po, fg, fo, tg, to, ag, ao)


@pytest.mark.parametrize("lp_name", list(LPS))
@pytest.mark.parametrize("pset", ["default", "tape_cpu", "dbl_eager"])
def test_rule_chain(pset, lp_name):
    """The rules applied one after another to one plan, as a phase does."""
    lp = get_lp(lp_name)
    st = stats(lp)
    for rules in (ROW_RULES, COL_RULES):
        pg, po = fresh_plan(lp), fresh_plan(lp)
        for rule in rules:
            fg, tg, ag = gp_ffi.rule_apply(rule, pg, lp, st, gp_ffi.params(**PARAM_SETS[pset]))
            fo, to, ao = O.rule_apply(rule, po, lp, st, O.params(**PARAM_SETS[pset]))
            _compare(f"chain {rule}/{pset}/{lp_name}", pg, po, fg, fo, tg, to, ag, ao)
            for k in ("has_change", "has_row_action", "has_col_action", "has_infeasible", "has_unbounded"):
                pg[k] = po[k] = fo[k]
            pg["obj_constant_delta"] = po["obj_constant_delta"] = fo["obj_constant_delta"]
            if fo["has_infeasible"] or fo["has_unbounded"] or fo["has_new_A"]:
                break


def test_dual_fix_candidate_mask():
    lp = get_lp("r1")
    st = stats(lp)
    cand = (np.arange(lp["n"]) % 3 == 0).astype(np.uint8)
    pg, po = fresh_plan(lp), fresh_plan(lp)
    fg, tg, ag = gp_ffi.rule_apply("dual_fix", pg, lp, st, gp_ffi.params(), cand)
    fo, to, ao = O.rule_apply("dual_fix", po, lp, st, O.params(), cand)
    _compare("dual_fix cand", pg, po, fg, fo, tg, to, ag, ao)
