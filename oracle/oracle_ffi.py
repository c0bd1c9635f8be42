"""ctypes binding of the CPU oracle (oracle/build/liboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
It runs the serial restatement of the reference's presolve/postsolve on
host numpy arrays and returns results in the same dict layout as the HIP
library's binding, so a test compares the two field by field.
"""
import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(_HERE), "gpu-presolver_amd"))
import gp_types as T  # noqa: E402

# ORACLE_LIB selects another build of the same sources (the CPU baseline's
# OpenMP leg loads build/liboracle_omp.so).
LIB_PATH = os.environ.get("ORACLE_LIB") or os.path.join(_HERE, "build", "liboracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.orc_last_error.restype = C.c_char_p
        L.orc_params_default.argtypes = [C.POINTER(T.GpParams)]
        L.orc_row_activity.argtypes = [C.POINTER(T.GpCsr), vp, vp, C.c_double, vp, vp, vp, vp]
        L.orc_col_max_abs.argtypes = [C.POINTER(T.GpCsr), vp]
        L.orc_compact_csr.argtypes = [C.POINTER(T.GpCsr), vp, vp, C.POINTER(T.GpCsr)]
        L.orc_transpose_csr.argtypes = [C.POINTER(T.GpCsr), C.POINTER(T.GpCsr)]
        L.orc_rule_apply.argtypes = [C.c_int32, C.POINTER(T.GpPlanView), C.POINTER(T.GpLp),
                                     C.POINTER(T.GpStatsView), C.POINTER(T.GpParams), vp, C.POINTER(vp)]
        L.orc_tape_view.argtypes = [vp, C.POINTER(T.GpRecordView)]
        L.orc_tape_free.argtypes = [vp]
        L.orc_presolve.argtypes = [C.POINTER(T.GpLp), C.POINTER(T.GpParams), C.c_int32, C.POINTER(vp)]
        L.orc_summary.argtypes = [vp, C.POINTER(T.GpSummary)]
        L.orc_record.argtypes = [vp, C.POINTER(T.GpRecordView)]
        L.orc_reduced_lp.argtypes = [vp, C.POINTER(T.GpLp)]
        L.orc_result_free.argtypes = [vp]
        L.orc_postsolve.argtypes = [C.POINTER(T.GpRecordView), vp, vp, vp, C.POINTER(T.GpLp), vp, vp, vp]
        L.orc_free.argtypes = [vp]
        L.orc_has_good_nnz_progress.argtypes = [C.c_int32, C.c_int32, C.c_double]
        L.orc_has_good_nnz_progress.restype = C.c_int32
        L.orc_set_batched_closure_min_nnz.argtypes = [C.c_int32]
        L.orc_set_batched_closure_min_nnz.restype = None
        L.orc_batched_closure_min_nnz.argtypes = []
        L.orc_batched_closure_min_nnz.restype = C.c_int32
        L.orc_batched_closure_runs.argtypes = []
        L.orc_batched_closure_runs.restype = C.c_int64
        L.orc_dup_row_hashes.argtypes = [C.POINTER(T.GpCsr), vp, C.c_double, vp]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())


def params(**kw):
    p = T.GpParams()
    lib().orc_params_default(C.byref(p))
    return p.set(**kw)


def has_good_nnz_progress(before, after, ratio):
    return bool(lib().orc_has_good_nnz_progress(before, after, ratio))


def set_batched_closure_min_nnz(nnz):
    """Working-LP nnz from which the fast phase opens with the batched
    singleton closure (10,000,000 by default, as in the reference)."""
    lib().orc_set_batched_closure_min_nnz(int(nnz))


def batched_closure_min_nnz():
    return int(lib().orc_batched_closure_min_nnz())


def batched_closure_runs():
    """Times the oracle's batched singleton closure has run in this process."""
    return int(lib().orc_batched_closure_runs())


def dup_row_hashes(m, n, ptr, idx, val, keep_row, ctol):
    """The duplicate-row hash of every row, as the rule computes it."""
    ptr = np.ascontiguousarray(ptr, np.int32)
    idx = np.ascontiguousarray(idx, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    keep = np.ascontiguousarray(keep_row, np.uint8)
    out = np.zeros(max(m, 1), np.uint64)
    g = T.GpCsr(m, n, len(idx), ptr.ctypes.data, idx.ctypes.data if idx.size else None,
                val.ctypes.data if val.size else None)
    _check(lib().orc_dup_row_hashes(C.byref(g), keep.ctypes.data, ctol, out.ctypes.data))
    return out[:m]


def _ptr(a):
    return a.ctypes.data if a is not None and a.size else None


class HostLp:
    """Keeps numpy arrays of an LP dict alive behind a gp_lp of host pointers."""

    def __init__(self, lp):
        self.keep = {}
        def arr(k, dt):
            if lp.get(k) is None:
                return None
            a = np.ascontiguousarray(lp[k], dtype=dt)
            if a.size == 0:
                a = np.zeros(1, dtype=dt)
            self.keep[k] = a
            return a.ctypes.data
        g = T.GpLp()
        m, n = lp["m"], lp["n"]
        nnz = len(lp["A_idx"])
        g.A = T.GpCsr(m, n, nnz, arr("A_ptr", np.int32), arr("A_idx", np.int32), arr("A_val", np.float64))
        g.AT = T.GpCsr(n, m, nnz, arr("AT_ptr", np.int32), arr("AT_idx", np.int32), arr("AT_val", np.float64))
        g.c = arr("c", np.float64)
        g.AL = arr("AL", np.float64)
        g.AU = arr("AU", np.float64)
        g.l = arr("l", np.float64)
        g.u = arr("u", np.float64)
        g.obj_constant = float(lp.get("obj", 0.0))
        self.g = g


def csr_struct(m, n, ptr, idx, val):
    keep = [np.ascontiguousarray(ptr, np.int32), np.ascontiguousarray(idx, np.int32),
            np.ascontiguousarray(val, np.float64)]
    keep = [k if k.size else np.zeros(1, k.dtype) for k in keep]
    return T.GpCsr(m, n, len(idx), keep[0].ctypes.data, keep[1].ctypes.data, keep[2].ctypes.data), keep


def row_activity(m, n, ptr, idx, val, l, u, ztol):
    A, keep = csr_struct(m, n, ptr, idx, val)
    mn = np.zeros(m); mx = np.zeros(m); ni = np.zeros(m, np.int32); pi = np.zeros(m, np.int32)
    l = np.ascontiguousarray(l, np.float64); u = np.ascontiguousarray(u, np.float64)
    _check(lib().orc_row_activity(C.byref(A), l.ctypes.data, u.ctypes.data, ztol,
                                   mn.ctypes.data, mx.ctypes.data, ni.ctypes.data, pi.ctypes.data))
    return mn, mx, ni, pi


def col_max_abs(n, m, ptr, idx, val):
    AT, keep = csr_struct(n, m, ptr, idx, val)
    out = np.zeros(n)
    _check(lib().orc_col_max_abs(C.byref(AT), out.ctypes.data))
    return out


def _take_csr(g):
    ptr = np.ctypeslib.as_array(C.cast(g.row_ptr, T.i32p), shape=(g.rows + 1,)).copy()
    idx = np.ctypeslib.as_array(C.cast(g.col_idx, T.i32p), shape=(max(g.nnz, 1),))[:g.nnz].copy()
    val = np.ctypeslib.as_array(C.cast(g.val, T.f64p), shape=(max(g.nnz, 1),))[:g.nnz].copy()
    for p in (g.row_ptr, g.col_idx, g.val):
        lib().orc_free(p)
    return ptr, idx, val


def compact_csr(m, n, ptr, idx, val, keep_row, keep_col):
    A, keep = csr_struct(m, n, ptr, idx, val)
    kr = np.ascontiguousarray(keep_row, np.uint8); kc = np.ascontiguousarray(keep_col, np.uint8)
    out = T.GpCsr()
    _check(lib().orc_compact_csr(C.byref(A), kr.ctypes.data, kc.ctypes.data, C.byref(out)))
    return (out.rows, out.cols) + _take_csr(out)


def transpose_csr(m, n, ptr, idx, val):
    A, keep = csr_struct(m, n, ptr, idx, val)
    out = T.GpCsr()
    _check(lib().orc_transpose_csr(C.byref(A), C.byref(out)))
    return (out.rows, out.cols) + _take_csr(out)


def rule_apply(rule, plan, lp, stats, prm, candidate=None):
    """Runs one rule on host arrays.  plan: dict with keep_row, keep_col, c, l, u,
    AL, AU (numpy, modified in place) and flags; stats: dict of numpy arrays.
    Returns (plan_flags dict, tape dict, new_A or None)."""
    hl = HostLp(lp)
    pv = T.GpPlanView()
    for k in ("keep_row", "keep_col", "c", "l", "u", "AL", "AU"):
        a = plan.get(k)
        setattr(pv, k, a.ctypes.data if a is not None else None)
    pv.obj_constant_delta = plan.get("obj_constant_delta", 0.0)
    for k in ("has_change", "has_row_action", "has_col_action", "has_infeasible", "has_unbounded"):
        setattr(pv, k, int(plan.get(k, 0)))
    sv = T.GpStatsView()
    for k, _ in T.GpStatsView._fields_:
        a = stats.get(k) if stats else None
        setattr(sv, k, a.ctypes.data if a is not None else None)
    tape = C.c_void_p()
    cand = np.ascontiguousarray(candidate, np.uint8) if candidate is not None else None
    _check(lib().orc_rule_apply(T.GP_RULE[rule] if isinstance(rule, str) else rule, C.byref(pv), C.byref(hl.g),
                                C.byref(sv), C.byref(prm), cand.ctypes.data if cand is not None else None,
                                C.byref(tape)))
    flags = {k: getattr(pv, k) for k in ("obj_constant_delta", "has_change", "has_row_action",
                                          "has_col_action", "has_infeasible", "has_unbounded", "has_new_A")}
    rv = T.GpRecordView()
    _check(lib().orc_tape_view(tape, C.byref(rv)))
    rv.m0 = rv.n0 = rv.m1 = rv.n1 = 0
    t = T.record_to_dict(rv)
    lib().orc_tape_free(tape)
    new_A = None
    if pv.has_new_A:
        new_A = _take_csr(pv.new_A)
    return flags, t, new_A


def presolve(lp, prm, keep_reduced=True):
    """Returns (summary dict, record dict, reduced LP dict or None)."""
    hl = HostLp(lp)
    h = C.c_void_p()
    _check(lib().orc_presolve(C.byref(hl.g), C.byref(prm), int(keep_reduced), C.byref(h)))
    try:
        s = T.GpSummary()
        _check(lib().orc_summary(h, C.byref(s)))
        summary = {k: getattr(s, k) for k, _ in T.GpSummary._fields_}
        rv = T.GpRecordView()
        _check(lib().orc_record(h, C.byref(rv)))
        rec = T.record_to_dict(rv)
        red = None
        if keep_reduced:
            g = T.GpLp()
            _check(lib().orc_reduced_lp(h, C.byref(g)))
            red = lp_from_struct(g)
        return summary, rec, red
    finally:
        lib().orc_result_free(h)


def lp_from_struct(g):
    m, n, nnz = g.A.rows, g.A.cols, g.A.nnz
    def a(p, k, t):
        if k <= 0 or not p:
            return np.zeros(0, dtype=np.float64 if t is T.f64p else np.int32)
        return np.ctypeslib.as_array(C.cast(p, t), shape=(k,)).copy()
    return {"m": m, "n": n,
            "A_ptr": a(g.A.row_ptr, m + 1, T.i32p) if m >= 0 else np.zeros(1, np.int32),
            "A_idx": a(g.A.col_idx, nnz, T.i32p), "A_val": a(g.A.val, nnz, T.f64p),
            "AT_ptr": a(g.AT.row_ptr, n + 1, T.i32p), "AT_idx": a(g.AT.col_idx, nnz, T.i32p),
            "AT_val": a(g.AT.val, nnz, T.f64p),
            "c": a(g.c, n, T.f64p), "l": a(g.l, n, T.f64p), "u": a(g.u, n, T.f64p),
            "AL": a(g.AL, m, T.f64p), "AU": a(g.AU, m, T.f64p), "obj": g.obj_constant}


def postsolve(rec, x_red, y_red, z_red, original=None):
    rk = T.RecordKeepAlive(rec)
    n0, m0 = rec["n0"], rec["m0"]
    x = np.zeros(max(n0, 1)); y = np.zeros(max(m0, 1)); z = np.zeros(max(n0, 1))
    def p(a):
        if a is None:
            return None
        a = np.ascontiguousarray(a, np.float64)
        return a.ctypes.data if a.size else None
    keep = [np.ascontiguousarray(v, np.float64) if v is not None else None for v in (x_red, y_red, z_red)]
    hl = HostLp(original) if original is not None else None
    _check(lib().orc_postsolve(C.byref(rk.view), p(keep[0]), p(keep[1]), p(keep[2]),
                                C.byref(hl.g) if hl else None, x.ctypes.data, y.ctypes.data, z.ctypes.data))
    return x[:n0], y[:m0], z[:n0]
