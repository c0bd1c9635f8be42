"""The solver hand-off: gp_pdhg_solve (restarted PDHG, all vectors on the
device) against HiGHS (scipy.optimize.linprog) as the independent check.
The reference hands its reduced LP to an external GPU solver and holds no
solver of its own, so there is no oracle restatement here: a returned point
is checked by the KKT quantities the reference's postsolve checker uses
(cpp/tools/gpu_presolver_check_postsolve_solution.cpp:212-246) recomputed on
the host, and by HiGHS's optimal objective.  The tolerances below follow
from the solver's relative KKT tolerance tol: the reported residuals are
<= tol by construction of the stop test, the objective is within a few tol
of the optimum (relative), and the largest row or bound violation is within
tol * (1 + ||b||)."""
import ctypes as C

import numpy as np
import pytest

import gp_ffi
import lp_check as K
import lp_synth

pytestmark = pytest.mark.gpu


def _lp_from_dense(A, AL, AU, c, l, u, obj=0.0):
    A = np.asarray(A, np.float64)
    m, n = A.shape
    rows, cols = np.nonzero(A)
    ptr, idx, val = lp_synth.csr_from_coo(m, n, rows, cols, A[rows, cols])
    tptr, tidx, tval = lp_synth.transpose_csr(m, n, ptr, idx, val)
    f = lambda v: np.asarray(v, np.float64)
    return {"m": m, "n": n, "A_ptr": ptr, "A_idx": idx, "A_val": val, "AT_ptr": tptr, "AT_idx": tidx,
            "AT_val": tval, "c": f(c), "AL": f(AL), "AU": f(AU), "l": f(l), "u": f(u), "obj": obj}


def _host_checks(lp, res, x, y, z, tol):
    """The returned point, re-evaluated on the host in FP64."""
    A = K.matrix(lp)
    b = np.maximum(np.where(np.isfinite(lp["AL"]), np.abs(lp["AL"]), 0.0),
                   np.where(np.isfinite(lp["AU"]), np.abs(lp["AU"]), 0.0))
    l, u = np.asarray(lp["l"]), np.asarray(lp["u"])
    assert np.all(x >= l) and np.all(x <= u), "x leaves its bounds"
    assert K.primal_violation(lp, x) <= 2 * tol * (1.0 + np.linalg.norm(b))
    pobj = K.objective(lp, x)
    assert abs(pobj - res["primal_obj"]) <= 1e-9 * (1.0 + abs(pobj))
    # z is c - A^T y projected on the cone of the column bounds: zero on free
    # columns, >= 0 without an upper bound, <= 0 without a lower bound.
    assert np.all(z[~np.isfinite(l) & ~np.isfinite(u)] == 0.0)
    assert np.all(z[~np.isfinite(u)] >= 0.0) and np.all(z[~np.isfinite(l)] <= 0.0)
    g = np.asarray(lp["c"]) - A.T @ y
    assert np.linalg.norm(g - z) <= 2 * tol * (1.0 + np.linalg.norm(lp["c"]))
    return pobj


def test_textbook_lp():
    """max x1 + x2 s.t. x1 + 2 x2 <= 4, 3 x1 + x2 <= 6, x >= 0: the optimum
    is the vertex (8/5, 6/5) with value 14/5 and duals (2/5, 1/5)."""
    lp = _lp_from_dense([[1, 2], [3, 1]], [-np.inf, -np.inf], [4, 6], [-1, -1], [0, 0], [np.inf, np.inf])
    res, x, y, z = gp_ffi.pdhg_solve(lp, tol=1e-10)
    assert res["status"] == 0
    np.testing.assert_allclose(x, [1.6, 1.2], atol=1e-8)
    np.testing.assert_allclose(y, [-0.4, -0.2], atol=1e-8)  # c - A^T y - z = 0 with z = 0
    np.testing.assert_allclose(z, [0.0, 0.0], atol=1e-8)
    assert abs(res["primal_obj"] + 2.8) <= 1e-9


def test_bounds_only():
    """No rows: every column goes to the bound its cost points at."""
    lp = _lp_from_dense(np.zeros((0, 4)), [], [], [1.0, -2.0, 0.5, -1.0], [-1, 0, 2, -3], [3, 5, 4, 7])
    res, x, y, z = gp_ffi.pdhg_solve(lp, tol=1e-10)
    assert res["status"] == 0
    np.testing.assert_array_equal(x, [-1.0, 5.0, 2.0, 7.0])
    np.testing.assert_array_equal(z, [1.0, -2.0, 0.5, -1.0])
    assert len(y) == 0


CASES = {
    "planted_200x300": lambda: lp_synth.generate(200, 300, seed=11, mean_row_len=4.0),
    "planted_eq_heavy": lambda: lp_synth.generate(300, 400, seed=12, mean_row_len=3.0, frac_eq=0.6),
    "config0": lambda: lp_synth.config0(seed=0),
}


@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("case", list(CASES))
def test_matches_highs(case, method):
    lp = CASES[case]()
    tol = 1e-6
    st, f, *_ = K.solve(lp)
    assert st == 0
    res, x, y, z = gp_ffi.pdhg_solve(lp, tol=tol, method=method)
    assert res["status"] == 0, res
    assert max(res["primal_res"], res["dual_res"], res["gap"]) <= tol
    pobj = _host_checks(lp, res, x, y, z, tol)
    assert abs(pobj - f) <= 10 * tol * (1.0 + abs(f)), (pobj, f)


@pytest.mark.parametrize("method", [0, 1])
def test_deterministic_and_graph_equals_direct(method):
    """Same LP, same parameters: the same bits, whether each check period is
    a replayed graph or its kernels launched one by one."""
    lp = lp_synth.generate(200, 300, seed=11, mean_row_len=4.0)
    runs = [gp_ffi.pdhg_solve(lp, tol=1e-6, use_graph=g, method=method) for g in (1, 1, 0)]
    for res, *vecs in runs[1:]:
        assert res["iterations"] == runs[0][0]["iterations"]
        assert res["restarts"] == runs[0][0]["restarts"]
        for a, b in zip(vecs, runs[0][1:]):
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    assert runs[0][0]["graph_launches"] == runs[0][0]["checks"] > 0
    assert runs[2][0]["graph_launches"] == 0


def test_iteration_limit():
    lp = lp_synth.config0(seed=0)
    res, x, y, z = gp_ffi.pdhg_solve(lp, tol=1e-12, max_iters=128, check_every=64)
    assert res["status"] == 1 and res["iterations"] == 128 and res["checks"] == 2


def test_invalid_parameters():
    lp = lp_synth.generate(50, 70, seed=8, mean_row_len=4.0)
    dlp = gp_ffi.DeviceLp(lp)
    for bad in (dict(tol=0.0), dict(check_every=0), dict(power_iters=0), dict(max_iters=-1), dict(method=2)):
        with pytest.raises(gp_ffi.GpError, match="status 1"):
            gp_ffi.pdhg_solve_view(dlp.g, gp_ffi.pdhg_params(**bad))


@pytest.mark.parametrize("method", [0, 1])
def test_presolve_solve_postsolve(method):
    """The hand-off: presolve on the device, the reduced LP solved in place
    (its device view goes straight into gp_pdhg_solve), the reduced solution
    postsolved on the device against the original LP.  The recovered x is
    checked against HiGHS on the original LP."""
    lp = lp_synth.config0(seed=0)
    tol = 1e-7
    dlp = gp_ffi.DeviceLp(lp)
    h = gp_ffi.Presolved(dlp, gp_ffi.params(enable_structure_specific_rules=0), True)
    try:
        view = h.reduced_view()
        m1, n1 = view.A.rows, view.A.cols
        res, (dx, dy, dz) = gp_ffi.pdhg_solve_view(view, gp_ffi.pdhg_params(tol=tol, method=method))
        assert res["status"] == 0, res
        rec = h.record()
        red = h.reduced()
    finally:
        h.free()
    x, y, z = dx.numpy(), dy.numpy(), dz.numpy()
    assert (len(x), len(y)) == (n1, m1)
    _host_checks(red, res, x, y, z, tol)
    xo, yo, zo = gp_ffi.postsolve(rec, x, y, z, original_dev=dlp)
    st, f, *_ = K.solve(lp)
    assert st == 0
    assert abs(K.objective(lp, xo) - f) <= 10 * tol * (1.0 + abs(f))
    assert K.primal_violation(lp, xo) <= 1e-4


@pytest.mark.parametrize("method", [0, 1])
def test_free_columns_and_equalities(method):
    """Free columns, equality rows and a one-sided row: min x0 + 2 x1 - x2
    s.t. x0 + x1 = 3, x1 - x2 = 1, x0 + x2 >= 1, x0, x2 free, 0 <= x1 <= 5.
    The optimum (from the equalities, x0 = 3 - x1, x2 = x1 - 1, so the
    objective is 3 - x1 + 2 x1 - x1 + 1 = 4 on the whole feasible segment
    x0 + x2 = 2 >= 1) is 4; HiGHS agrees."""
    lp = _lp_from_dense([[1, 1, 0], [0, 1, -1], [1, 0, 1]], [3, 1, 1], [3, 1, np.inf], [1, 2, -1],
                        [-np.inf, 0, -np.inf], [np.inf, 5, np.inf])
    st, f, *_ = K.solve(lp)
    assert st == 0 and abs(f - 4.0) <= 1e-9
    res, x, y, z = gp_ffi.pdhg_solve(lp, tol=1e-9, method=method)
    assert res["status"] == 0
    assert abs(res["primal_obj"] - 4.0) <= 1e-6
    _host_checks(lp, res, x, y, z, 1e-9)
    assert z[0] == 0.0 and z[2] == 0.0  # free columns have no reduced cost


def test_infeasible_lp_stops_at_the_limit():
    """x0 + x1 <= 1 and x0 + x1 >= 3 with x >= 0 has no solution: the solver
    must stop at its iteration limit (status 1) with a primal residual far
    from zero, not report an optimum."""
    lp = _lp_from_dense([[1, 1], [1, 1]], [-np.inf, 3], [1, np.inf], [1, 1], [0, 0], [np.inf, np.inf])
    res, x, y, z = gp_ffi.pdhg_solve(lp, tol=1e-6, max_iters=4096)
    assert res["status"] == 1 and res["iterations"] == 4096
    assert res["primal_res"] > 1e-3


@pytest.mark.parametrize("method", [0, 1])
def test_hand_off_on_the_block_structured_lp(method):
    """The hand-off on a block-structured LP (100 blocks of 40 x 60 and 10
    linking rows): GPU presolve, the reduced view solved in place, GPU
    postsolve; the recovered objective against HiGHS on the original LP.
    This reduced LP is hard for restarted PDHG (about 47k iterations to a
    relative KKT error of 1e-4; 1e-5 takes far longer), so the check runs at
    1e-4: objective within 10 tol (relative), row and bound violations within
    10 tol (1 + ||b||) after postsolve."""
    lp = lp_synth.config2(seed=2, blocks=100, link_rows=10, link_nnz=1500)
    tol = 1e-4
    dlp = gp_ffi.DeviceLp(lp)
    h = gp_ffi.Presolved(dlp, gp_ffi.params(enable_structure_specific_rules=0), True)
    try:
        res, (dx, dy, dz) = gp_ffi.pdhg_solve_view(h.reduced_view(), gp_ffi.pdhg_params(tol=tol, method=method))
        rec = h.record()
    finally:
        h.free()
    assert res["status"] == 0, res
    xo, yo, zo = gp_ffi.postsolve(rec, dx.numpy(), dy.numpy(), dz.numpy(), original_dev=dlp)
    st, f, *_ = K.solve(lp)
    assert st == 0
    assert abs(K.objective(lp, xo) - f) <= 10 * tol * (1.0 + abs(f))
    b = np.maximum(np.where(np.isfinite(lp["AL"]), np.abs(lp["AL"]), 0.0),
                   np.where(np.isfinite(lp["AU"]), np.abs(lp["AU"]), 0.0))
    assert K.primal_violation(lp, xo) <= 10 * tol * (1.0 + np.linalg.norm(b))


@pytest.mark.parametrize("case", ["config0", "block"])
def test_pipeline_equals_the_steps(case):
    """gp_presolve_solve (postsolve prepared on a second stream while the
    solver runs) gives exactly what gp_presolve, gp_pdhg_solve on the
    reduced view and gp_postsolve give one after another."""
    lp = lp_synth.config0(seed=0) if case == "config0" else \
        lp_synth.config2(seed=2, blocks=100, link_rows=10, link_nnz=1500)
    tol = 1e-6 if case == "config0" else 1e-4
    prm = gp_ffi.params(enable_structure_specific_rules=0)
    sprm = gp_ffi.pdhg_params(tol=tol)
    dlp = gp_ffi.DeviceLp(lp)
    res, (px, py, pz) = gp_ffi.presolve_solve(dlp, prm, sprm)
    h = gp_ffi.Presolved(dlp, prm, True)
    try:
        sres, (dx, dy, dz) = gp_ffi.pdhg_solve_view(h.reduced_view(), sprm)
        rec = h.record()
        s = h.summary()
    finally:
        h.free()
    xo, yo, zo = gp_ffi.postsolve(rec, dx.numpy(), dy.numpy(), dz.numpy(), original_dev=dlp)
    assert (res["reduced_rows"], res["reduced_cols"], res["reduced_nnz"]) == \
           (s["reduced_rows"], s["reduced_cols"], s["reduced_nnz"])
    assert res["solve"]["status"] == 0 and res["solve"]["iterations"] == sres["iterations"]
    for a, b in zip((px.numpy(), py.numpy(), pz.numpy()), (xo, yo, zo)):
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    assert res["prepare_seconds"] > 0.0 and res["total_seconds"] >= res["presolve_seconds"]


def test_pipeline_on_an_infeasible_lp():
    """Presolve finds the LP infeasible (an empty row whose bounds exclude
    zero): nothing is solved, the outputs are zero and the flag comes back."""
    lp = _lp_from_dense([[1, 1], [0, 0]], [1, 1], [3, 2], [1, 1], [0, 0], [5, 5])
    dlp = gp_ffi.DeviceLp(lp)
    res, (px, py, pz) = gp_ffi.presolve_solve(dlp, gp_ffi.params(enable_structure_specific_rules=0),
                                              gp_ffi.pdhg_params())
    assert res["has_infeasible"] == 1 and res["solve"]["status"] == -1
    assert not px.numpy().any() and not py.numpy().any() and not pz.numpy().any()
