"""Prepared postsolve (gp_postsolve_prepare / gp_postsolve_plan_run): every
run equals the one-shot gp_postsolve of the same reduced point bit for bit,
eagerly and replayed from the captured hipGraph, across several runs with
other device work in between."""
import numpy as np
import pytest

import gp_ffi
import lp_synth
from helpers import assert_same_array

pytestmark = pytest.mark.gpu

CASES = {
    "c0": lambda: lp_synth.config0(seed=0),
    "zipf": lambda: lp_synth.generate(20000, 30000, seed=5, row_len=(2, 14), zipf=1.5),
}


def _presolved(make):
    lp = make()
    _, rec, red = gp_ffi.presolve(lp, gp_ffi.params(enable_structure_specific_rules=0), True)
    return lp, rec, red


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("use_graph", [True, False])
def test_plan_runs_match_postsolve(case, use_graph):
    lp, rec, red = _presolved(CASES[case])
    dlp = gp_ffi.DeviceLp(lp)
    plan = gp_ffi.PostsolvePlan(rec, original_dev=dlp, use_graph=use_graph)
    points = [lp_synth.reduced_solution(seed, red["m"], red["n"], red["l"], red["u"]) for seed in (1, 2, 3)]
    for rep in range(2):
        for seed, (x, y, z) in zip((1, 2, 3), points):
            want = gp_ffi.postsolve(rec, x, y, z, original_dev=dlp)  # eager work between the replays
            got = plan.run(x, y, z)
            for name, a, b in zip("xyz", got, want):
                assert_same_array(f"{case} run {rep} seed {seed} {name}", a, b)
    assert plan.graph_runs() == (6 if use_graph else 0)
    plan.free()


def test_plan_missing_inputs_and_no_original():
    lp, rec, red = _presolved(CASES["c0"])
    x, y, z = lp_synth.reduced_solution(7, red["m"], red["n"], red["l"], red["u"])
    plan = gp_ffi.PostsolvePlan(rec, use_graph=True)
    for args in ((x, None, None), (x, y, z), (None, y, z)):
        got = plan.run(*args)
        want = gp_ffi.postsolve(rec, *args)
        for name, a, b in zip("xyz", got, want):
            assert_same_array(name, a, b)
    assert plan.graph_runs() == 1  # only the run with all three inputs replays the graph
    plan.free()


@pytest.mark.parametrize("case", list(CASES))
def test_plan_shape(case):
    """gp_postsolve_plan_shape: records and tape bytes as the record holds
    them; a pass has at least one level when the tape is not empty and no
    more levels than records."""
    lp, rec, red = _presolved(CASES[case])
    plan = gp_ffi.PostsolvePlan(rec, original=lp, use_graph=False)
    sh = plan.shape()
    plan.free()
    R = len(rec["types"])
    assert sh["records"] == R
    want = 4 * (R + len(rec["index_starts"]) + len(rec["value_starts"]) + len(rec["indices"])) + \
        len(rec["dual_modes"]) + 8 * len(rec["vals"])
    assert sh["tape_bytes"] == want
    assert 0 <= sh["dual_levels"] <= R and 0 <= sh["primal_levels"] <= R
    assert R == 0 or sh["dual_levels"] + sh["primal_levels"] >= 1
