"""Pins the CPU oracle with facts that do not come from the oracle itself:
closed forms (row activity), library routines (scipy transposes and
submatrices), and LP theory (presolve preserves the optimal value; postsolve
returns a feasible optimal point of the original LP, solved independently
by HiGHS).  A dropped term, a wrong sign or a transposed index in any rule
changes at least one of these."""
import numpy as np
import pytest
import scipy.sparse as sp

import lp_check as K
import lp_synth
import oracle_ffi as O


def _random_csr(m, n, seed, density=0.01):
    rng = np.random.default_rng(seed)
    A = sp.random(m, n, density=density, format="csr", random_state=rng,
                  data_rvs=lambda k: np.round(rng.uniform(0.5, 2.0, k) * 8) / 8 * rng.choice([-1.0, 1.0], k))
    A.sort_indices()
    return A


def test_row_activity_closed_form():
    m, n = 3000, 2000
    A = _random_csr(m, n, 1, 0.004)
    A.data[::11] = 0.0  # explicit zeros are skipped
    rng = np.random.default_rng(2)
    l = np.round(rng.uniform(-4, 0, n) * 1024) / 1024
    u = np.round(rng.uniform(0, 4, n) * 1024) / 1024
    l[rng.random(n) < 0.2] = -np.inf
    u[rng.random(n) < 0.2] = np.inf
    mn, mx, ni, pi = O.row_activity(m, n, A.indptr, A.indices, A.data, l, u, 1e-10)
    rows = np.repeat(np.arange(m), np.diff(A.indptr))
    a, lj, uj = A.data, l[A.indices], u[A.indices]
    live = np.abs(a) > 1e-10
    with np.errstate(invalid="ignore"):
        tmin = np.where(a >= 0, a * lj, a * uj)
        tmax = np.where(a >= 0, a * uj, a * lj)
    fmin = live & np.isfinite(tmin)
    fmax = live & np.isfinite(tmax)
    want_mn = np.bincount(rows[fmin], weights=tmin[fmin], minlength=m)
    want_mx = np.bincount(rows[fmax], weights=tmax[fmax], minlength=m)
    want_ni = np.bincount(rows[live & ~np.isfinite(tmin) & (tmin < 0)], minlength=m)
    want_pi = np.bincount(rows[live & ~np.isfinite(tmax) & (tmax > 0)], minlength=m)
    # dyadic data: every partial sum is exact, so the order of summation is irrelevant
    np.testing.assert_array_equal(mn, want_mn)
    np.testing.assert_array_equal(mx, want_mx)
    np.testing.assert_array_equal(ni, want_ni)
    np.testing.assert_array_equal(pi, want_pi)


@pytest.mark.parametrize("m,n,seed", [(1, 1, 1), (500, 300, 2), (4000, 6000, 3)])
def test_transpose_matches_scipy(m, n, seed):
    A = _random_csr(m, n, seed, 0.01 if m > 1 else 1.0)
    rows, cols, ptr, idx, val = O.transpose_csr(m, n, A.indptr, A.indices, A.data)
    T = A.T.tocsr()
    T.sort_indices()
    assert (rows, cols) == (n, m)
    np.testing.assert_array_equal(ptr, T.indptr)
    np.testing.assert_array_equal(idx, T.indices)
    np.testing.assert_array_equal(val, T.data)


@pytest.mark.parametrize("seed,frac", [(1, 0.2), (2, 0.0), (3, 1.0), (4, 0.5)])
def test_compaction_is_the_kept_submatrix(seed, frac):
    m, n = 2000, 1500
    A = _random_csr(m, n, seed, 0.01)
    rng = np.random.default_rng(seed)
    kr = (rng.random(m) >= frac).astype(np.uint8)
    kc = (rng.random(n) >= frac).astype(np.uint8)
    rows, cols, ptr, idx, val = O.compact_csr(m, n, A.indptr, A.indices, A.data, kr, kc)
    S = A[kr.astype(bool)][:, kc.astype(bool)].tocsr()
    S.sort_indices()
    assert (rows, cols) == S.shape
    if cols == 0:
        np.testing.assert_array_equal(ptr, np.zeros(rows + 1, np.int32))
        return
    np.testing.assert_array_equal(ptr, S.indptr)
    np.testing.assert_array_equal(idx, S.indices)
    np.testing.assert_array_equal(val, S.data)


EQUIV_LPS = {
    "c0": lambda: lp_synth.config0(seed=0),
    "r1": lambda: lp_synth.generate(400, 600, seed=21, mean_row_len=5.0),
    "r3": lambda: lp_synth.generate(1500, 2000, seed=23, mean_row_len=3.0, frac_doubleton=0.2, frac_eq=0.6),
    "zipf": lambda: lp_synth.generate(20000, 30000, seed=5, row_len=(2, 14), zipf=1.5),
}
EQUIV_PARAMS = {
    "default": {},
    "dbl_eager": dict(doubleton_equations_min_selected_per_batch=1, doubleton_equations_min_selected_ratio=0.0),
    "dbl_scalar": dict(doubleton_equations_scan=0),
    "tape_cpu": dict(record_postsolve_tape_cpu=1),
    "redundant": dict(enable_redundant_bounds=1, enable_zero_cost_redundant_box_bounds=1),
}
_lp_cache = {}


def _lp(name):
    if name not in _lp_cache:
        _lp_cache[name] = EQUIV_LPS[name]()
    return _lp_cache[name]


@pytest.mark.parametrize("pname", list(EQUIV_PARAMS))
@pytest.mark.parametrize("lname", list(EQUIV_LPS))
def test_presolve_preserves_optimum_and_postsolve_recovers_it(lname, pname):
    """optimum(original) == optimum(reduced) + obj shift; postsolve of the
    reduced optimum is feasible for the original with the same objective.

    Tolerances: the implied-bound rule fixes a forced column at its relaxed
    lower bound (rule_implied_variable_bounds.cu:582-617), up to
    0.5 * feasibility_tol * |x| off: the reduced LP may then need HiGHS's
    primal tolerance relaxed to 1e-5, and back-substitution through pivots
    of magnitude >= 1/2 can grow the recovered point's violation to 1e-4.
    A wrong term in a rule shows up as O(1) errors."""
    lp = _lp(lname)
    s, rec, red = O.presolve(lp, O.params(enable_structure_specific_rules=0, **EQUIV_PARAMS[pname]), True)
    assert not s["has_infeasible"] and not s["has_unbounded"]
    st0, f0, *_ = K.solve(lp)
    st1, f1, x1, y1, z1 = K.solve(red)
    if st1 == 2:
        st1, f1, x1, y1, z1 = K.solve(red, feasibility_tol=1e-5)
    assert st0 == 0 and st1 == 0, (st0, st1)
    scale = 1.0 + abs(f0)
    assert abs(f1 - f0) <= 1e-5 * scale, (f0, f1)
    x, y, z = O.postsolve(rec, x1, y1, z1, lp)
    assert K.primal_violation(lp, x) <= 1e-4
    assert abs(K.objective(lp, x) - f0) <= 1e-5 * scale


@pytest.fixture
def closure_always():
    """Batched singleton closure on every fast phase (threshold 0 nnz)."""
    before = O.batched_closure_min_nnz()
    O.set_batched_closure_min_nnz(0)
    yield
    O.set_batched_closure_min_nnz(before)


def test_closure_threshold_default():
    assert O.batched_closure_min_nnz() == 10_000_000


CLOSURE_PARAMS = {
    "default": {},
    "no_bootstrap": dict(enable_tiered_bootstrap=0),
    "no_bootstrap_tape_cpu": dict(enable_tiered_bootstrap=0, record_postsolve_tape_cpu=1),
}


@pytest.mark.parametrize("pname", list(CLOSURE_PARAMS))
@pytest.mark.parametrize("lname", list(EQUIV_LPS))
def test_closure_preserves_optimum_and_postsolve_recovers_it(lname, pname, closure_always):
    """The batched singleton closure (column-singleton EQ, dual fix and
    singleton rows on one plan with degree-refreshed statistics) keeps the
    LP's optimum, and postsolve of the reduced optimum recovers a feasible
    optimal point; same tolerances as above.  Without the bootstrap the
    first fast phase's closure does the whole singleton cascade, so the
    record must differ from the closure-free schedule's (it really ran)."""
    lp = _lp(lname)
    prm = dict(enable_structure_specific_rules=0, **CLOSURE_PARAMS[pname])
    s, rec, red = O.presolve(lp, O.params(**prm), True)
    assert not s["has_infeasible"] and not s["has_unbounded"]
    if pname != "default":
        O.set_batched_closure_min_nnz(10_000_000)
        _, rec_plain, _ = O.presolve(lp, O.params(**prm), True)
        O.set_batched_closure_min_nnz(0)
        assert (rec["types"].tolist() != rec_plain["types"].tolist()
                or rec["col_red2org"].tolist() != rec_plain["col_red2org"].tolist()
                or rec["indices"].tolist() != rec_plain["indices"].tolist())
    st0, f0, *_ = K.solve(lp)
    st1, f1, x1, y1, z1 = K.solve(red)
    if st1 == 2:
        st1, f1, x1, y1, z1 = K.solve(red, feasibility_tol=1e-5)
    assert st0 == 0 and st1 == 0, (st0, st1)
    scale = 1.0 + abs(f0)
    assert abs(f1 - f0) <= 1e-5 * scale, (f0, f1)
    x, y, z = O.postsolve(rec, x1, y1, z1, lp)
    assert K.primal_violation(lp, x) <= 1e-4
    assert abs(K.objective(lp, x) - f0) <= 1e-5 * scale
    if pname == "no_bootstrap_tape_cpu" and lname in ("c0", "r1"):
        assert K.dual_residual(lp, y, z) <= 1e-9


@pytest.mark.parametrize("lname", ["c0", "r1"])
def test_dual_recovery_exact_with_cpu_tape(lname):
    """With every rule taping (record_postsolve_tape_cpu), the recovered
    duals satisfy c - A^T y - z = 0 on these LPs.  With the default
    parameters several rules leave no tape and their duals come only from
    the final projection; the residual is then reported, not asserted."""
    lp = _lp(lname)
    s, rec, red = O.presolve(lp, O.params(enable_structure_specific_rules=0, record_postsolve_tape_cpu=1), True)
    st1, f1, x1, y1, z1 = K.solve(red)
    assert st1 == 0
    assert K.dual_residual(red, y1, z1) <= 1e-9
    x, y, z = O.postsolve(rec, x1, y1, z1, lp)
    assert K.dual_residual(lp, y, z) <= 1e-9


def _row_key(ptr, idx, val, i):
    b, e = ptr[i], ptr[i + 1]
    if b == e:
        return None
    return (tuple(idx[b:e]), tuple(np.round(val[b:e] / val[b], 12)))


def test_duplicate_rows_leave_no_parallel_pair():
    lp = lp_synth.generate(3000, 4000, seed=31, mean_row_len=4.0, frac_dup_rows=0.1)
    from helpers import fresh_plan, stats
    plan = fresh_plan(lp)
    flags, tape, _ = O.rule_apply("duplicate_rows", plan, lp, stats(lp), O.params())
    assert flags["has_change"]
    ptr, idx, val = lp["A_ptr"], lp["A_idx"], lp["A_val"]
    kept = {}
    for i in range(lp["m"]):
        if plan["keep_row"][i]:
            k = _row_key(ptr, idx, val, i)
            if k is not None:
                assert k not in kept, f"rows {kept[k]} and {i} are parallel and both kept"
                kept[k] = i
    for i in np.nonzero(plan["keep_row"] == 0)[0]:
        assert _row_key(ptr, idx, val, i) in kept, f"deleted row {i} has no kept parallel row"
