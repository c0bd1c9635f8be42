"""Host FP64 checks of LP solutions: solving with scipy's HiGHS, primal
feasibility, objective, and the dual residual c - A^T y - z of a recovered
(x, y, z) (the checks of the reference's postsolve checker tool,
cpp/tools/gpu_presolver_check_postsolve_solution.cpp:212-246)."""
import numpy as np
import scipy.optimize as so
import scipy.sparse as sp


def matrix(lp):
    return sp.csr_matrix((lp["A_val"], lp["A_idx"], lp["A_ptr"]), shape=(lp["m"], lp["n"]))


def solve(lp, feasibility_tol=None):
    """HiGHS on min c'x s.t. AL <= Ax <= AU, l <= x <= u.  Returns
    (status, objective including obj, x, y, z) with y, z in the convention
    c - A^T y - z = 0.  feasibility_tol overrides HiGHS's primal tolerance."""
    m, n = lp["m"], lp["n"]
    A = matrix(lp)
    AL, AU = np.asarray(lp["AL"]), np.asarray(lp["AU"])
    eq = np.isfinite(AL) & np.isfinite(AU) & (AL == AU)
    up = ~eq & np.isfinite(AU)
    lo = ~eq & np.isfinite(AL)
    blocks, rhs = [], []
    if up.any():
        blocks.append(A[up])
        rhs.append(AU[up])
    if lo.any():
        blocks.append(-A[lo])
        rhs.append(-AL[lo])
    kw = {}
    if blocks:
        kw["A_ub"] = sp.vstack(blocks).tocsr()
        kw["b_ub"] = np.concatenate(rhs)
    if eq.any():
        kw["A_eq"] = A[eq]
        kw["b_eq"] = AU[eq]
    bounds = [(None if not np.isfinite(a) else a, None if not np.isfinite(b) else b)
              for a, b in zip(lp["l"], lp["u"])]
    if n == 0:
        feasible = bool(np.all(AL <= 0) and np.all(AU >= 0))
        return (0 if feasible else 2), float(lp.get("obj", 0.0)), np.zeros(0), np.zeros(m), np.zeros(0)
    if feasibility_tol is not None:
        kw["options"] = {"primal_feasibility_tolerance": feasibility_tol}
    r = so.linprog(lp["c"], bounds=bounds, method="highs", **kw)
    if r.status != 0:
        return r.status, None, None, None, None
    y = np.zeros(m)
    k = 0
    if up.any():
        y[up] += r.ineqlin.marginals[k:k + up.sum()]
        k += up.sum()
    if lo.any():
        y[lo] -= r.ineqlin.marginals[k:k + lo.sum()]
    if eq.any():
        y[eq] = r.eqlin.marginals
    z = r.lower.marginals + r.upper.marginals
    return 0, float(r.fun) + float(lp.get("obj", 0.0)), r.x, y, z


def primal_violation(lp, x):
    """Largest bound or row violation of x."""
    A = matrix(lp)
    ax = A @ x
    l, u = np.asarray(lp["l"]), np.asarray(lp["u"])
    viol = [np.max(np.maximum(l - x, 0.0), initial=0.0), np.max(np.maximum(x - u, 0.0), initial=0.0)]
    if lp["m"]:
        viol += [np.max(np.maximum(np.asarray(lp["AL"]) - ax, 0.0)), np.max(np.maximum(ax - np.asarray(lp["AU"]), 0.0))]
    return float(max(viol))


def dual_residual(lp, y, z):
    """max |c - A^T y - z|."""
    if lp["n"] == 0:
        return 0.0
    r = np.asarray(lp["c"]) - matrix(lp).T @ y - z
    return float(np.max(np.abs(r)))


def objective(lp, x):
    return float(np.dot(lp["c"], x)) + float(lp.get("obj", 0.0))
