"""Properties the mathematics fixes for two oracle functions whose exact
choices the reference's tests do not pin, checked by brute force on small
random LPs:

* the doubleton batch's matching (rule_doubleton_equations.cu:876-913:
  rounds of minimum-row-index claims on both columns of every candidate
  row): the rows it substitutes never share a column, and on these LPs
  (conflict chains shorter than its 8 rounds) they are exactly the
  lexicographically first maximal set of column-disjoint doubleton rows,
  i.e. scanning the rows in index order and taking every row whose two
  columns are still free;
* duplicate rows with three or more parallel members
  (rule_duplicate_rows.cu:243-381): one member survives, and its interval
  is the intersection of every member's interval expressed in the
  survivor's scale (a row r = f k constrains a_k x to [AL_r, AU_r] / f);
  an empty intersection marks the LP infeasible.

All coefficients are dyadic, so these intersections are exact."""
import numpy as np
import pytest

import oracle_ffi as O
from helpers import lp_from_rows

ALL_OFF = dict(enable_structure_specific_rules=0, enable_infeasible_fixed_variables=0, enable_empty_rows=0,
               enable_singleton_rows=0, enable_infeasible_redundant_rows=0, enable_implied_variable_bounds=0,
               enable_duplicate_rows=0, enable_empty_cols=0, enable_column_singletons_eq=0,
               enable_column_singletons_dual_infer=0, enable_doubleton_equations=0, enable_dual_fix=0,
               enable_duplicate_columns=0, enable_redundant_bounds=0, use_tiered_scheduler=0, max_iters=1)
DOUBLETON = 13  # PostsolveReductionType::DoubletonEquation


def _doubleton_lp(seed):
    """n columns; doubleton equality rows over random column pairs (many
    share columns), plus longer inequality rows; coefficients in +-{1, 2}."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(12, 30))
    rows, AL, AU = [], [], []
    pairs = []
    for _ in range(int(rng.integers(8, 24))):
        if rng.random() < 0.75:
            a, b = sorted(rng.choice(n, size=2, replace=False).tolist())
            pairs.append(len(rows))
            rows.append([(a, float(rng.choice([-2.0, -1.0, 1.0, 2.0]))), (b, float(rng.choice([-2.0, -1.0, 1.0, 2.0])))])
            rhs = float(rng.integers(-4, 5))
            AL.append(rhs)
            AU.append(rhs)
        else:
            cols = sorted(rng.choice(n, size=int(rng.integers(3, 6)), replace=False).tolist())
            rows.append([(j, float(rng.choice([-1.0, 1.0]))) for j in cols])
            AL.append(-np.inf)
            AU.append(float(rng.integers(5, 20)))
    c = rng.choice([-1.0, 0.0, 1.0], size=n)
    return lp_from_rows(rows, n, c, AL, AU, [-64.0] * n, [64.0] * n), rows, pairs


def _greedy(rows, candidates):
    used, take = set(), []
    for i in candidates:
        a, b = rows[i][0][0], rows[i][1][0]
        if a not in used and b not in used:
            take.append(i)
            used.update((a, b))
    return take


@pytest.mark.parametrize("seed", range(12))
def test_doubleton_batch_is_the_lexicographic_greedy_matching(seed):
    lp, rows, pairs = _doubleton_lp(seed)
    prm = O.params(**dict(ALL_OFF, enable_doubleton_equations=1, doubleton_equations_single_batch_per_iter=1,
                          doubleton_equations_max_batch_rounds=1, doubleton_equations_min_selected_per_batch=1,
                          doubleton_equations_min_selected_ratio=0.0))
    s, rec, red = O.presolve(lp, prm, True)
    if s["has_infeasible"]:
        pytest.skip("a substitution emptied a column box (the matching is not observable then)")
    t = np.asarray(rec["types"])
    starts = np.asarray(rec["index_starts"])
    idx = np.asarray(rec["indices"])
    chosen, cols = [], []
    for k in np.nonzero(t == DOUBLETON)[0]:
        e, kc, row = idx[starts[k]:starts[k] + 3]
        chosen.append(int(row))
        cols += [int(e), int(kc)]
        assert sorted((int(e), int(kc))) == sorted(j for j, _ in rows[int(row)])
    assert len(cols) == len(set(cols)), "two substituted rows share a column"
    assert sorted(chosen) == _greedy(rows, pairs)


def _dup_lp(seed, infeasible=False):
    """Groups of 3-5 parallel rows (a base row times +-{1/2, 1, 2, 4}) with
    random one- or two-sided intervals, among unrelated rows."""
    rng = np.random.default_rng(seed)
    n = 40
    rows, AL, AU, groups = [], [], [], []
    used = set()
    for g in range(int(rng.integers(2, 5))):
        cols = sorted(rng.choice(n, size=int(rng.integers(2, 5)), replace=False).tolist())
        key = tuple(cols)
        if key in used:
            continue
        used.add(key)
        base = [float(rng.choice([-2.0, -1.0, 1.0, 2.0])) for _ in cols]
        center = float(rng.integers(-8, 9))
        members = []
        for _ in range(int(rng.integers(3, 6))):
            f = float(rng.choice([-2.0, -1.0, -0.5, 0.5, 1.0, 2.0, 4.0]))
            lo, hi = center - float(rng.integers(0, 6)), center + float(rng.integers(0, 6))
            kind = rng.integers(0, 3)
            lo, hi = (lo, hi) if kind == 0 else ((-np.inf, hi) if kind == 1 else (lo, np.inf))
            if infeasible and not members:
                lo, hi = center + 20.0, center + 30.0  # disjoint from every other member (all contain center)
            a, b = (lo * f, hi * f) if f > 0 else (hi * f, lo * f)
            members.append((len(rows), f))
            rows.append([(j, v * f) for j, v in zip(cols, base)])
            AL.append(a)
            AU.append(b)
        groups.append(members)
    for _ in range(6):  # unrelated rows with distinct column sets
        cols = sorted(rng.choice(n, size=6, replace=False).tolist())
        rows.append([(j, float(rng.choice([-1.0, 1.0, 3.0]))) for j in cols])
        AL.append(-np.inf)
        AU.append(50.0)
    return lp_from_rows(rows, n, np.zeros(n), AL, AU, [-100.0] * n, [100.0] * n), groups


@pytest.mark.parametrize("seed", range(12))
def test_duplicate_rows_keep_the_intersection(seed):
    lp, groups = _dup_lp(seed)
    prm = O.params(**dict(ALL_OFF, enable_duplicate_rows=1))
    s, rec, red = O.presolve(lp, prm, True)
    assert not s["has_infeasible"]
    r2o = list(np.asarray(rec["row_red2org"]))
    for members in groups:
        alive = [r for r, _ in members if r in r2o]
        assert len(alive) == 1, (members, alive)
        keep = alive[0]
        fk = dict(members)[keep]
        lo, hi = -np.inf, np.inf
        for r, f in members:
            # a_keep x = (fk / f) (a_r x), a_r x in [AL_r, AU_r]
            q = fk / f
            a, b = lp["AL"][r] * q, lp["AU"][r] * q
            if q < 0:
                a, b = b, a
            lo, hi = max(lo, a), min(hi, b)
        i = r2o.index(keep)
        assert (red["AL"][i], red["AU"][i]) == (lo, hi), (keep, red["AL"][i], red["AU"][i], lo, hi)


@pytest.mark.parametrize("seed", range(4))
def test_duplicate_rows_empty_intersection_is_infeasible(seed):
    lp, _ = _dup_lp(100 + seed, infeasible=True)
    s, _, _ = O.presolve(lp, O.params(**dict(ALL_OFF, enable_duplicate_rows=1)), True)
    assert s["has_infeasible"]


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(6))
def test_gpu_matches_on_the_property_lps(seed):
    """The GPU rules give the same records on these LPs (so the properties
    above hold for them too)."""
    import gp_ffi
    from helpers import assert_same_record

    for lp, kw in ((_doubleton_lp(seed)[0], dict(enable_doubleton_equations=1, doubleton_equations_max_batch_rounds=1,
                                                  doubleton_equations_single_batch_per_iter=1,
                                                  doubleton_equations_min_selected_per_batch=1,
                                                  doubleton_equations_min_selected_ratio=0.0)),
                   (_dup_lp(seed)[0], dict(enable_duplicate_rows=1))):
        prm = dict(ALL_OFF, **kw)
        sg, rg, _ = gp_ffi.presolve(lp, gp_ffi.params(**prm), True)
        so, ro, _ = O.presolve(lp, O.params(**prm), True)
        assert (sg["reduced_rows"], sg["reduced_cols"], sg["has_infeasible"]) == \
               (so["reduced_rows"], so["reduced_cols"], so["has_infeasible"])
        assert_same_record(rg, ro)
