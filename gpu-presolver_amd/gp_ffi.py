"""ctypes binding of libgpu_presolver_amd.so (the HIP library, C-ABI in
include/gp_presolve_capi.h).

Device memory is managed through the HIP runtime directly (libamdhip64), so
the binding needs neither torch nor a CUDA-compat layer.  The functions take
and return host numpy arrays with the same layout as the CPU oracle's
binding, which lets a test compare the two field by field.  There is no CPU
fallback: if the library is not built or no GPU is present, every call
raises.
"""
import ctypes as C
import os

import numpy as np

import gp_types as T

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libgpu_presolver_amd.so")
_lib = None
_hip = None

H2D, D2H, D2D = 1, 2, 3


class GpError(RuntimeError):
    pass


def hip():
    """The HIP runtime (device memory, copies, events, streams)."""
    global _hip
    if _hip is None:
        H = C.CDLL("libamdhip64.so", mode=C.RTLD_GLOBAL)
        vp = C.c_void_p
        H.hipMalloc.argtypes = [C.POINTER(vp), C.c_size_t]
        H.hipFree.argtypes = [vp]
        H.hipMemcpy.argtypes = [vp, vp, C.c_size_t, C.c_int]
        H.hipMemset.argtypes = [vp, C.c_int, C.c_size_t]
        H.hipGetErrorString.restype = C.c_char_p
        H.hipGetErrorString.argtypes = [C.c_int]
        H.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        H.hipEventCreate.argtypes = [C.POINTER(vp)]
        H.hipEventRecord.argtypes = [vp, vp]
        H.hipEventSynchronize.argtypes = [vp]
        H.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), vp, vp]
        H.hipEventDestroy.argtypes = [vp]
        H.hipStreamCreate.argtypes = [C.POINTER(vp)]
        H.hipStreamDestroy.argtypes = [vp]
        H.hipStreamSynchronize.argtypes = [vp]
        _hip = H
    return _hip


def hip_check(rc, what):
    if rc != 0:
        raise GpError(f"{what}: {hip().hipGetErrorString(rc).decode()}")


def device_count():
    n = C.c_int(0)
    rc = hip().hipGetDeviceCount(C.byref(n))
    return n.value if rc == 0 else 0


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GpError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        hip()
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.gp_last_error.restype = C.c_char_p
        L.gp_version.restype = C.c_char_p
        L.gp_params_default.argtypes = [C.POINTER(T.GpParams)]
        L.gp_device_free.argtypes = [vp]
        L.gp_row_activity.argtypes = [C.POINTER(T.GpCsr), vp, vp, C.c_double, vp, vp, vp, vp, vp]
        L.gp_row_activity_noalloc.argtypes = [C.POINTER(T.GpCsr), vp, vp, C.c_double, vp, vp, vp, vp, vp]
        L.gp_col_max_abs.argtypes = [C.POINTER(T.GpCsr), vp, vp]
        L.gp_compact_csr.argtypes = [C.POINTER(T.GpCsr), vp, vp, C.POINTER(T.GpCsr)]
        L.gp_transpose_csr.argtypes = [C.POINTER(T.GpCsr), C.POINTER(T.GpCsr)]
        L.gp_rule_apply.argtypes = [C.c_int32, C.POINTER(T.GpPlanView), C.POINTER(T.GpLp),
                                    C.POINTER(T.GpStatsView), C.POINTER(T.GpParams), vp, C.POINTER(vp)]
        L.gp_tape_view.argtypes = [vp, C.POINTER(T.GpRecordView)]
        L.gp_tape_free.argtypes = [vp]
        L.gp_presolve.argtypes = [C.POINTER(T.GpLp), C.POINTER(T.GpParams), C.c_int32, C.POINTER(vp)]
        L.gp_presolve_summary.argtypes = [vp, C.POINTER(T.GpSummary)]
        L.gp_presolve_record.argtypes = [vp, C.POINTER(T.GpRecordView)]
        L.gp_presolve_reduced_lp.argtypes = [vp, C.POINTER(T.GpLp)]
        L.gp_presolve_free.argtypes = [vp]
        L.gp_postsolve.argtypes = [C.POINTER(T.GpRecordView), vp, vp, vp, C.POINTER(T.GpLp), vp, vp, vp]
        L.gp_has_good_nnz_progress.argtypes = [C.c_int32, C.c_int32, C.c_double]
        L.gp_has_good_nnz_progress.restype = C.c_int32
        L.gp_set_batched_closure_min_nnz.argtypes = [C.c_int32]
        L.gp_set_batched_closure_min_nnz.restype = None
        L.gp_batched_closure_min_nnz.argtypes = []
        L.gp_batched_closure_min_nnz.restype = C.c_int32
        L.gp_server_create.argtypes = [vp, C.POINTER(T.GpServeConfig), C.POINTER(vp)]
        L.gp_server_submit.argtypes = [vp, C.POINTER(T.GpLp), C.c_uint64, C.c_double, C.POINTER(C.c_int64)]
        L.gp_server_wait.argtypes = [vp, C.POINTER(T.GpJobResult), C.c_int32, C.c_double]
        L.gp_server_wait.restype = C.c_int32
        L.gp_server_now.argtypes = [vp]
        L.gp_server_now.restype = C.c_double
        L.gp_server_destroy.argtypes = [vp]
        L.gp_server_destroy.restype = None
        L.gp_serve_trace.argtypes = [C.POINTER(T.GpLp), C.c_int32, vp, vp, C.c_int32, vp,
                                     C.POINTER(T.GpServeConfig), C.POINTER(T.GpJobResult)]
        L.gp_synthetic_point.argtypes = [C.c_uint64, C.POINTER(T.GpLp), vp, vp, vp]
        L.gp_plain_job.argtypes = [C.POINTER(T.GpLp), vp, C.c_uint64, C.c_uint64, C.c_int32,
                                   C.POINTER(T.GpJobResult)]
        L.gp_plain_job_spec.argtypes = [C.POINTER(T.GpLpSpec), vp, C.c_uint64, C.c_uint64, C.c_int32,
                                        C.POINTER(T.GpJobResult)]
        L.gp_generate_lp.argtypes = [C.POINTER(T.GpLpSpec), C.POINTER(T.GpLp)]
        L.gp_free_lp.argtypes = [C.POINTER(T.GpLp)]
        L.gp_server_submit_spec.argtypes = [vp, C.POINTER(T.GpLpSpec), C.c_uint64, C.c_double,
                                            C.POINTER(C.c_int64)]
        L.gp_serve_spec_trace.argtypes = [C.POINTER(T.GpLpSpec), C.c_int32, vp, vp,
                                          C.POINTER(T.GpServeConfig), C.POINTER(T.GpJobResult)]
        L.gp_checksum.argtypes = [vp, C.c_int32, vp, C.c_int32, vp, C.POINTER(C.c_uint64)]
        L.gp_mps_read.argtypes = [C.c_char_p, C.POINTER(vp)]
        L.gp_mps_parse.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(vp)]
        L.gp_mps_get_view.argtypes = [vp, C.POINTER(T.GpMpsView)]
        L.gp_mps_upload.argtypes = [vp, C.POINTER(T.GpLp)]
        L.gp_mps_free.argtypes = [vp]
        L.gp_mps_free.restype = None
        L.gp_test_check_plan_reuse.argtypes = [C.c_int32]
        L.gp_test_check_plan_reuse.restype = None
        L.gp_test_plan_reuses.argtypes = []
        L.gp_test_plan_reuses.restype = C.c_int64
        L.gp_pdhg_params_default.argtypes = [C.POINTER(T.GpPdhgParams)]
        L.gp_pdhg_params_default.restype = None
        L.gp_pdhg_solve.argtypes = [C.POINTER(T.GpLp), C.POINTER(T.GpPdhgParams), vp, vp, vp,
                                    C.POINTER(T.GpPdhgResult)]
        L.gp_presolve_solve.argtypes = [C.POINTER(T.GpLp), C.POINTER(T.GpParams), C.POINTER(T.GpPdhgParams), vp, vp,
                                        vp, C.POINTER(T.GpPipelineResult)]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise GpError(f"gp status {rc}: " + lib().gp_last_error().decode())


def params(**kw):
    p = T.GpParams()
    lib().gp_params_default(C.byref(p))
    return p.set(**kw)


def has_good_nnz_progress(before, after, ratio):
    return bool(lib().gp_has_good_nnz_progress(before, after, ratio))


def check_plan_reuse(on):
    """Testing hook: every reused phase plan is compared with a fresh one."""
    lib().gp_test_check_plan_reuse(1 if on else 0)


def plan_reuses():
    """Phases on this thread that reused their predecessor's plan so far."""
    return int(lib().gp_test_plan_reuses())


def set_batched_closure_min_nnz(nnz):
    """Working-LP nnz from which the fast phase opens with the batched
    singleton closure (10,000,000 by default, as in the reference)."""
    lib().gp_set_batched_closure_min_nnz(int(nnz))


def batched_closure_min_nnz():
    return int(lib().gp_batched_closure_min_nnz())


# ---- device memory ---------------------------------------------------------
class DeviceArray:
    """hipMalloc'ed copy of a numpy array (freed with the object)."""

    def __init__(self, host=None, *, size=0, dtype=np.float64):
        if host is not None:
            host = np.ascontiguousarray(host)
            self.dtype, self.size = host.dtype, host.size
        else:
            self.dtype, self.size = np.dtype(dtype), int(size)
        self.nbytes = self.size * self.dtype.itemsize
        p = C.c_void_p()
        hip_check(hip().hipMalloc(C.byref(p), max(self.nbytes, 8)), "hipMalloc")
        self.ptr = p.value
        if host is not None and self.nbytes:
            hip_check(hip().hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, H2D), "hipMemcpy H2D")

    def numpy(self):
        out = np.empty(self.size, dtype=self.dtype)
        if self.nbytes:
            hip_check(hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H), "hipMemcpy D2H")
        return out

    def __del__(self):
        if getattr(self, "ptr", None) and _hip is not None:
            hip().hipFree(self.ptr)
            self.ptr = None


def download(ptr, n, dtype):
    """Copies n elements from a device pointer to a new numpy array."""
    out = np.empty(n, dtype=dtype)
    if n > 0:
        hip_check(hip().hipMemcpy(out.ctypes.data, ptr, out.nbytes, D2H), "hipMemcpy D2H")
    return out


def _take_csr(g, free=True):
    ptr = download(g.row_ptr, g.rows + 1, np.int32)
    idx = download(g.col_idx, g.nnz, np.int32)
    val = download(g.val, g.nnz, np.float64)
    if free:
        for p in (g.row_ptr, g.col_idx, g.val):
            if p:
                _check(lib().gp_device_free(p))
    return ptr, idx, val


class DeviceCsr:
    def __init__(self, rows, cols, ptr, idx, val):
        self.keep = [DeviceArray(np.asarray(ptr, np.int32)), DeviceArray(np.asarray(idx, np.int32)),
                     DeviceArray(np.asarray(val, np.float64))]
        self.g = T.GpCsr(rows, cols, len(idx), self.keep[0].ptr, self.keep[1].ptr, self.keep[2].ptr)


class DeviceLp:
    """An LP dict (see lp_synth) uploaded to HBM behind a gp_lp of device pointers."""

    def __init__(self, lp):
        m, n = lp["m"], lp["n"]
        self.m, self.n = m, n
        self.A = DeviceCsr(m, n, lp["A_ptr"], lp["A_idx"], lp["A_val"])
        self.AT = DeviceCsr(n, m, lp["AT_ptr"], lp["AT_idx"], lp["AT_val"])
        # Absent vectors stay null pointers, as a caller may leave them.
        self.vec = {k: DeviceArray(np.asarray(lp[k], np.float64)) for k in ("c", "AL", "AU", "l", "u")
                    if lp.get(k) is not None}
        g = T.GpLp()
        g.A, g.AT = self.A.g, self.AT.g
        for k in ("c", "AL", "AU", "l", "u"):
            setattr(g, k, self.vec[k].ptr if k in self.vec else None)
        g.obj_constant = float(lp.get("obj", 0.0))
        self.g = g

    def at_host(self):
        """The (possibly row-sorted in place) AT back on the host."""
        return (self.AT.keep[0].numpy(), self.AT.keep[1].numpy(), self.AT.keep[2].numpy())


# ---- primitives ------------------------------------------------------------
def row_activity(m, n, ptr, idx, val, l, u, ztol, stream=None):
    A = DeviceCsr(m, n, ptr, idx, val)
    dl, du = DeviceArray(np.asarray(l, np.float64)), DeviceArray(np.asarray(u, np.float64))
    mn, mx = DeviceArray(size=m), DeviceArray(size=m)
    ni, pi = DeviceArray(size=m, dtype=np.int32), DeviceArray(size=m, dtype=np.int32)
    _check(lib().gp_row_activity(C.byref(A.g), dl.ptr, du.ptr, ztol, mn.ptr, mx.ptr, ni.ptr, pi.ptr, stream))
    if stream is not None:
        hip_check(hip().hipStreamSynchronize(stream), "hipStreamSynchronize")
    return mn.numpy(), mx.numpy(), ni.numpy(), pi.numpy()


def col_max_abs(n, m, ptr, idx, val):
    AT = DeviceCsr(n, m, ptr, idx, val)
    out = DeviceArray(size=n)
    _check(lib().gp_col_max_abs(C.byref(AT.g), out.ptr, None))
    return out.numpy()


def compact_csr(m, n, ptr, idx, val, keep_row, keep_col):
    A = DeviceCsr(m, n, ptr, idx, val)
    kr, kc = DeviceArray(np.asarray(keep_row, np.uint8)), DeviceArray(np.asarray(keep_col, np.uint8))
    out = T.GpCsr()
    _check(lib().gp_compact_csr(C.byref(A.g), kr.ptr, kc.ptr, C.byref(out)))
    return (out.rows, out.cols) + _take_csr(out)


def radix_sort(keys, vals, key_bits):
    """Stable sort of (keys, vals) by the low key_bits bits of the keys
    (uint32 or uint64) through gpk_radix_sort_u32_i32 / _u64_i32 on the null
    stream; returns the sorted (keys, vals)."""
    keys = np.ascontiguousarray(keys)
    vals = np.ascontiguousarray(vals, dtype=np.int32)
    n = keys.size
    L = lib()
    fn = L.gpk_radix_sort_u64_i32 if keys.dtype == np.uint64 else L.gpk_radix_sort_u32_i32
    vp = C.c_void_p
    fn.argtypes = [vp, vp, vp, vp, C.c_int32, C.c_int32, vp, vp]
    L.gpk_radix_scratch_ints.restype = C.c_int64
    L.gpk_radix_scratch_ints.argtypes = [C.c_int64]
    dk, dv = DeviceArray(keys), DeviceArray(vals)
    ka, va = DeviceArray(size=max(n, 1), dtype=keys.dtype), DeviceArray(size=max(n, 1), dtype=np.int32)
    scratch = DeviceArray(size=max(1, L.gpk_radix_scratch_ints(n)), dtype=np.int32)
    hip_check(fn(dk.ptr, dv.ptr, ka.ptr, va.ptr, n, key_bits, scratch.ptr, None), "radix sort")
    hip_check(hip().hipDeviceSynchronize(), "radix sort synchronize")
    return dk.numpy(), dv.numpy()


def transpose_csr(m, n, ptr, idx, val):
    A = DeviceCsr(m, n, ptr, idx, val)
    out = T.GpCsr()
    _check(lib().gp_transpose_csr(C.byref(A.g), C.byref(out)))
    return (out.rows, out.cols) + _take_csr(out)


# ---- rule operator API -----------------------------------------------------
_PLAN_VEC = ("keep_row", "keep_col", "c", "l", "u", "AL", "AU")


def rule_apply(rule, plan, lp, stats, prm, candidate=None):
    """Runs one rule on the GPU.  plan: dict with keep_row, keep_col, c, l, u,
    AL, AU (numpy, updated in place from the device) and flags; stats: dict of
    numpy arrays.  Returns (plan_flags dict, tape dict, new_A or None)."""
    dl = DeviceLp(lp)
    dev = {}
    pv = T.GpPlanView()
    for k in _PLAN_VEC:
        a = plan.get(k)
        if a is not None:
            dev[k] = DeviceArray(a)
            setattr(pv, k, dev[k].ptr)
    pv.obj_constant_delta = plan.get("obj_constant_delta", 0.0)
    for k in ("has_change", "has_row_action", "has_col_action", "has_infeasible", "has_unbounded"):
        setattr(pv, k, int(plan.get(k, 0)))
    sv = T.GpStatsView()
    sdev = {}
    for k, _ in T.GpStatsView._fields_:
        a = stats.get(k) if stats else None
        if a is not None:
            sdev[k] = DeviceArray(a)
            setattr(sv, k, sdev[k].ptr)
    cand = DeviceArray(np.asarray(candidate, np.uint8)) if candidate is not None else None
    tape = C.c_void_p()
    _check(lib().gp_rule_apply(T.GP_RULE[rule] if isinstance(rule, str) else rule, C.byref(pv), C.byref(dl.g),
                               C.byref(sv), C.byref(prm), cand.ptr if cand is not None else None, C.byref(tape)))
    for k, d in dev.items():
        plan[k][...] = d.numpy()
    flags = {k: getattr(pv, k) for k in ("obj_constant_delta", "has_change", "has_row_action",
                                          "has_col_action", "has_infeasible", "has_unbounded", "has_new_A")}
    rv = T.GpRecordView()
    _check(lib().gp_tape_view(tape, C.byref(rv)))
    t = T.record_to_dict(rv)
    lib().gp_tape_free(tape)
    new_A = _take_csr(pv.new_A) if pv.has_new_A else None
    # The rules sort the caller's AT rows in place (doubleton batch,
    # duplicate columns); report it so a test can compare with the oracle.
    t["_AT_after"] = dl.at_host()
    return flags, t, new_A


# ---- presolve / postsolve --------------------------------------------------
def lp_from_struct(g):
    m, n, nnz = g.A.rows, g.A.cols, g.A.nnz
    return {"m": m, "n": n,
            "A_ptr": download(g.A.row_ptr, m + 1, np.int32), "A_idx": download(g.A.col_idx, nnz, np.int32),
            "A_val": download(g.A.val, nnz, np.float64),
            "AT_ptr": download(g.AT.row_ptr, n + 1, np.int32), "AT_idx": download(g.AT.col_idx, nnz, np.int32),
            "AT_val": download(g.AT.val, nnz, np.float64),
            "c": download(g.c, n, np.float64), "l": download(g.l, n, np.float64),
            "u": download(g.u, n, np.float64), "AL": download(g.AL, m, np.float64),
            "AU": download(g.AU, m, np.float64), "obj": g.obj_constant}


class Presolved:
    """A presolve handle (device-resident reduced LP and host record)."""

    def __init__(self, dlp, prm, keep_reduced=True):
        self.dlp = dlp
        self.h = C.c_void_p()
        _check(lib().gp_presolve(C.byref(dlp.g), C.byref(prm), int(keep_reduced), C.byref(self.h)))
        self.keep_reduced = keep_reduced

    def summary(self):
        s = T.GpSummary()
        _check(lib().gp_presolve_summary(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in T.GpSummary._fields_}

    def record(self):
        rv = T.GpRecordView()
        _check(lib().gp_presolve_record(self.h, C.byref(rv)))
        return T.record_to_dict(rv)

    def reduced_view(self):
        """Device view (gp_lp) of the reduced LP, valid until free()."""
        g = T.GpLp()
        _check(lib().gp_presolve_reduced_lp(self.h, C.byref(g)))
        return g

    def reduced(self):
        return lp_from_struct(self.reduced_view())

    def free(self):
        if self.h:
            lib().gp_presolve_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        if _lib is not None:
            self.free()


def presolve(lp, prm, keep_reduced=True):
    """Returns (summary dict, record dict, reduced LP dict or None)."""
    dlp = DeviceLp(lp)
    h = Presolved(dlp, prm, keep_reduced)
    try:
        return h.summary(), h.record(), (h.reduced() if keep_reduced else None)
    finally:
        h.free()


class PreparedPostsolve:
    """A postsolve whose inputs (reduced x/y/z, record view, original LP) are
    already in place, so run() times gp_postsolve alone."""

    def __init__(self, rec, x_red, y_red, z_red, original_dev):
        self.rk = T.RecordKeepAlive(rec)
        n0, m0 = rec["n0"], rec["m0"]
        self.out = (DeviceArray(size=n0), DeviceArray(size=m0), DeviceArray(size=n0))
        self.ins = [DeviceArray(np.asarray(v, np.float64)) for v in (x_red, y_red, z_red)]
        self.orig = original_dev

    def run(self):
        dx, dy, dz = self.out
        _check(lib().gp_postsolve(C.byref(self.rk.view), *[d.ptr for d in self.ins], C.byref(self.orig.g),
                                  dx.ptr, dy.ptr, dz.ptr))

    def run_plan(self, plan):
        """The same postsolve through a PostsolvePlan of this record."""
        dx, dy, dz = self.out
        _check(lib().gp_postsolve_plan_run(plan.h, *[d.ptr for d in self.ins], dx.ptr, dy.ptr, dz.ptr))


def postsolve(rec, x_red, y_red, z_red, original=None, original_dev=None):
    """original: LP dict (uploaded here) or original_dev: a resident DeviceLp."""
    rk = T.RecordKeepAlive(rec)
    n0, m0 = rec["n0"], rec["m0"]
    dx, dy, dz = DeviceArray(size=n0), DeviceArray(size=m0), DeviceArray(size=n0)
    ins = [DeviceArray(np.asarray(v, np.float64)) if v is not None else None for v in (x_red, y_red, z_red)]
    orig = original_dev if original_dev is not None else (DeviceLp(original) if original is not None else None)
    _check(lib().gp_postsolve(C.byref(rk.view), *[d.ptr if d is not None else None for d in ins],
                              C.byref(orig.g) if orig is not None else None, dx.ptr, dy.ptr, dz.ptr))
    return dx.numpy(), dy.numpy(), dz.numpy()


class PostsolvePlan:
    """gp_postsolve_prepare / gp_postsolve_plan_run: the record's tape, maps
    and replay schedules built once, any number of reduced points postsolved
    with them (replayed from a captured hipGraph when use_graph is set and
    x, y, z are all given)."""

    def __init__(self, rec, original=None, original_dev=None, use_graph=True):
        L = lib()
        L.gp_postsolve_prepare.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(C.c_void_p)]
        L.gp_postsolve_plan_run.argtypes = [C.c_void_p] + [C.c_void_p] * 6
        L.gp_postsolve_plan_graph_runs.restype = C.c_int64
        L.gp_postsolve_plan_graph_runs.argtypes = [C.c_void_p]
        L.gp_postsolve_plan_free.argtypes = [C.c_void_p]
        self._rk = T.RecordKeepAlive(rec)  # the record view must outlive the plan
        self._orig = original_dev if original_dev is not None else (
            DeviceLp(original) if original is not None else None)
        self.n0, self.m0 = rec["n0"], rec["m0"]
        self.h = C.c_void_p()
        _check(L.gp_postsolve_prepare(C.addressof(self._rk.view),
                                      C.addressof(self._orig.g) if self._orig is not None else None,
                                      1 if use_graph else 0, C.byref(self.h)))

    def run(self, x_red, y_red, z_red):
        dx, dy, dz = DeviceArray(size=self.n0), DeviceArray(size=self.m0), DeviceArray(size=self.n0)
        ins = [DeviceArray(np.asarray(v, np.float64)) if v is not None else None for v in (x_red, y_red, z_red)]
        _check(lib().gp_postsolve_plan_run(self.h, *[d.ptr if d is not None else None for d in ins],
                                           dx.ptr, dy.ptr, dz.ptr))
        return dx.numpy(), dy.numpy(), dz.numpy()

    def graph_runs(self):
        return lib().gp_postsolve_plan_graph_runs(self.h)

    def shape(self):
        """Tape records, primal and dual dependency levels, tape bytes."""
        L = lib()
        L.gp_postsolve_plan_shape.argtypes = [C.c_void_p, C.POINTER(T.GpPostsolveShape)]
        out = T.GpPostsolveShape()
        _check(L.gp_postsolve_plan_shape(self.h, C.byref(out)))
        return {"records": out.records, "primal_levels": out.primal_levels, "dual_levels": out.dual_levels,
                "tape_bytes": out.tape_bytes}

    def free(self):
        if self.h:
            lib().gp_postsolve_plan_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            self.free()


# ---- solver hand-off ---------------------------------------------------------
def pdhg_params(**kw):
    p = T.GpPdhgParams()
    lib().gp_pdhg_params_default(C.byref(p))
    names = {f[0] for f in T.GpPdhgParams._fields_}
    for k, v in kw.items():
        if k not in names:
            raise KeyError(f"unknown gp_pdhg_params field {k!r}")
        setattr(p, k, v)
    return p


def pdhg_solve_view(g, prm=None, out=None):
    """gp_pdhg_solve on a device gp_lp (a DeviceLp's .g or a presolve
    handle's reduced_view()); returns (result dict, (x, y, z) DeviceArrays)."""
    m, n = g.A.rows, g.A.cols
    if out is None:
        out = (DeviceArray(size=n), DeviceArray(size=m), DeviceArray(size=n))
    r = T.GpPdhgResult()
    prm = prm if prm is not None else pdhg_params()
    _check(lib().gp_pdhg_solve(C.byref(g), C.byref(prm), out[0].ptr, out[1].ptr, out[2].ptr, C.byref(r)))
    return {k: getattr(r, k) for k, _ in T.GpPdhgResult._fields_}, out


def presolve_solve(dlp, prm, sprm, out=None):
    """gp_presolve_solve on a resident DeviceLp: presolve, the reduced LP
    solved in HBM while the postsolve is prepared, then postsolved.  Returns
    (result dict with the solve's dict under "solve", (x, y, z) DeviceArrays
    in the original LP's space)."""
    m, n = dlp.m, dlp.n
    if out is None:
        out = (DeviceArray(size=n), DeviceArray(size=m), DeviceArray(size=n))
    r = T.GpPipelineResult()
    _check(lib().gp_presolve_solve(C.byref(dlp.g), C.byref(prm), C.byref(sprm), out[0].ptr, out[1].ptr, out[2].ptr,
                                   C.byref(r)))
    d = {k: getattr(r, k) for k, _ in T.GpPipelineResult._fields_ if k != "solve"}
    d["solve"] = {k: getattr(r.solve, k) for k, _ in T.GpPdhgResult._fields_}
    return d, out


def pdhg_solve(lp, **kw):
    """Uploads an LP dict, solves it on the device; returns (result, x, y, z)."""
    dlp = DeviceLp(lp)
    res, (x, y, z) = pdhg_solve_view(dlp.g, pdhg_params(**kw))
    return res, x.numpy(), y.numpy(), z.numpy()


# ---- timing ----------------------------------------------------------------
class Stream:
    """A HIP stream at the default priority."""

    def __init__(self):
        s = C.c_void_p()
        hip_check(hip().hipStreamCreate(C.byref(s)), "hipStreamCreate")
        self.ptr = s.value

    def sync(self):
        hip_check(hip().hipStreamSynchronize(self.ptr), "hipStreamSynchronize")

    def __del__(self):
        if getattr(self, "ptr", None) and _hip is not None:
            hip().hipStreamDestroy(self.ptr)
            self.ptr = None


class Event:
    def __init__(self):
        e = C.c_void_p()
        hip_check(hip().hipEventCreate(C.byref(e)), "hipEventCreate")
        self.ptr = e.value

    def record(self, stream):
        hip_check(hip().hipEventRecord(self.ptr, stream.ptr if isinstance(stream, Stream) else stream),
                  "hipEventRecord")

    def elapsed_ms(self, later):
        hip_check(hip().hipEventSynchronize(later.ptr), "hipEventSynchronize")
        ms = C.c_float()
        hip_check(hip().hipEventElapsedTime(C.byref(ms), self.ptr, later.ptr), "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        if getattr(self, "ptr", None) and _hip is not None:
            hip().hipEventDestroy(self.ptr)
            self.ptr = None


# ---- serving and batching ----------------------------------------------------
JOB_FIELDS = [k for k, _ in T.GpJobResult._fields_]


def serve_config(workers, postsolve=True, hbm_budget_bytes=0.0, seed=0):
    return T.GpServeConfig(int(workers), int(bool(postsolve)), float(hbm_budget_bytes), int(seed))


def _results_dict(res):
    return [{k: getattr(r, k) for k in JOB_FIELDS} for r in res]


def serve_trace(dlps, lp_of_job, prm, cfg, arrival_s=None):
    """Runs job k = dlps[lp_of_job[k]] arriving at arrival_s[k] (None: all at
    once) through the serving runtime; returns a list of result dicts."""
    jobs = len(lp_of_job)
    lps = (T.GpLp * len(dlps))(*[d.g for d in dlps])
    loj = np.ascontiguousarray(lp_of_job, dtype=np.int32)
    arr = None if arrival_s is None else np.ascontiguousarray(arrival_s, dtype=np.float64)
    res = (T.GpJobResult * max(jobs, 1))()
    _check(lib().gp_serve_trace(lps, len(dlps), loj.ctypes.data, None if arr is None else arr.ctypes.data, jobs,
                                C.byref(prm), C.byref(cfg), res))
    return _results_dict(res[:jobs])


class Server:
    """The streaming form: submit jobs, wait for results."""

    def __init__(self, prm, cfg):
        self.h = C.c_void_p()
        _check(lib().gp_server_create(C.byref(prm), C.byref(cfg), C.byref(self.h)))
        self._keep = []

    def submit(self, dlp, tag, t_arrival=-1.0):
        self._keep.append(dlp)
        job = C.c_int64()
        _check(lib().gp_server_submit(self.h, C.byref(dlp.g), int(tag), float(t_arrival), C.byref(job)))
        return job.value

    def wait(self, max_results=64, timeout_s=1.0):
        buf = (T.GpJobResult * max_results)()
        n = lib().gp_server_wait(self.h, buf, max_results, timeout_s)
        return _results_dict(buf[:n])

    def now(self):
        return lib().gp_server_now(self.h)

    def close(self):
        if self.h:
            lib().gp_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        if _lib is not None:
            self.close()


def plain_job(dlp, prm, seed, tag, postsolve=True):
    """One serving job (presolve, seeded point seed ^ tag, postsolve,
    checksum) through the plain synchronous entry points on this thread."""
    out = T.GpJobResult()
    _check(lib().gp_plain_job(C.byref(dlp.g), C.byref(prm), int(seed), int(tag), int(bool(postsolve)),
                              C.byref(out)))
    return _results_dict([out])[0]


def lp_spec(seed, m, n=None, row_len=None, mean_len=0.0, min_len=1):
    """gp_lp_spec: general rows of U[row_len] entries, or max(min_len,
    Poisson(mean_len)) when mean_len > 0; n defaults to round(1.5 m)."""
    n = int(round(1.5 * m)) if n is None else int(n)
    lo, hi = row_len if row_len is not None else (0, 0)
    return T.GpLpSpec(int(seed), int(m), n, int(lo), int(hi), float(mean_len), int(min_len))


class GeneratedLp:
    """An LP generated on the device from a gp_lp_spec (library memory,
    released by gp_free_lp); .g is its gp_lp, as DeviceLp.g."""

    def __init__(self, spec):
        self.spec = spec
        self.g = T.GpLp()
        _check(lib().gp_generate_lp(C.byref(spec), C.byref(self.g)))
        self.m, self.n = self.g.A.rows, self.g.A.cols

    def host(self):
        """The LP as a host dict (the layout of lp_synth)."""
        g = self.g
        d = {"m": g.A.rows, "n": g.A.cols, "obj": g.obj_constant}
        for name, csr in (("A", g.A), ("AT", g.AT)):
            d[name + "_ptr"] = download(csr.row_ptr, csr.rows + 1, np.int32)
            d[name + "_idx"] = download(csr.col_idx, csr.nnz, np.int32)
            d[name + "_val"] = download(csr.val, csr.nnz, np.float64)
        for k, n in (("c", g.A.cols), ("l", g.A.cols), ("u", g.A.cols), ("AL", g.A.rows), ("AU", g.A.rows)):
            d[k] = download(getattr(g, k), n, np.float64)
        return d

    def free(self):
        if self.g is not None and _lib is not None:
            _check(lib().gp_free_lp(C.byref(self.g)))
            self.g = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def plain_job_spec(spec, prm, seed, tag, postsolve=True):
    out = T.GpJobResult()
    _check(lib().gp_plain_job_spec(C.byref(spec), C.byref(prm), int(seed), int(tag), int(bool(postsolve)),
                                   C.byref(out)))
    return _results_dict([out])[0]


def serve_spec_trace(specs, prm, cfg, arrival_s=None):
    """Job k generates specs[k] on its worker, arrives at arrival_s[k] (None:
    all at once) and is presolved (+ postsolved); returns result dicts."""
    jobs = len(specs)
    arr_specs = (T.GpLpSpec * max(jobs, 1))(*specs)
    arr = None if arrival_s is None else np.ascontiguousarray(arrival_s, dtype=np.float64)
    res = (T.GpJobResult * max(jobs, 1))()
    _check(lib().gp_serve_spec_trace(arr_specs, jobs, None if arr is None else arr.ctypes.data, C.byref(prm),
                                     C.byref(cfg), res))
    return _results_dict(res[:jobs])


def _mps_dict(view):
    m, n, nnz = view.rows, view.cols, view.nnz
    a = T.mps_view_arrays(view)
    return dict(a, name=view.name.decode(), m=m, n=n, obj=view.obj_constant)


class MpsModel:
    """A model read by the library's MPS reader (host memory); .lp is the
    LP dict (layout of lp_synth) plus its name."""

    def __init__(self, text=None, path=None):
        self.h = C.c_void_p()
        if path is not None:
            _check(lib().gp_mps_read(os.fsencode(path), C.byref(self.h)))
        else:
            data = text.encode() if isinstance(text, str) else bytes(text)
            _check(lib().gp_mps_parse(data, len(data), C.byref(self.h)))
        v = T.GpMpsView()
        _check(lib().gp_mps_get_view(self.h, C.byref(v)))
        self.lp = _mps_dict(v)

    def upload(self):
        """The model as a device LP (library memory, released with
        gp_free_lp); returned as a GeneratedLp-like holder."""
        return _UploadedLp(self)

    def close(self):
        if self.h and _lib is not None:
            lib().gp_mps_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        self.close()


class _UploadedLp(GeneratedLp):
    def __init__(self, model):
        self.spec = None
        self.g = T.GpLp()
        _check(lib().gp_mps_upload(model.h, C.byref(self.g)))
        self.m, self.n = self.g.A.rows, self.g.A.cols


def synthetic_point(seed, reduced_g):
    """The serving runtime's seeded reduced point for a reduced LP (gp_lp of
    device pointers); returns host copies of x, y, z."""
    n, m = reduced_g.A.cols, reduced_g.A.rows
    x, y, z = DeviceArray(size=n), DeviceArray(size=m), DeviceArray(size=n)
    _check(lib().gp_synthetic_point(int(seed), C.byref(reduced_g), x.ptr, y.ptr, z.ptr))
    return x.numpy(), y.numpy(), z.numpy()


def checksum(x, y, z):
    """gp_checksum of host vectors (uploaded first)."""
    dx, dy, dz = (DeviceArray(np.asarray(v, np.float64)) for v in (x, y, z))
    out = C.c_uint64()
    _check(lib().gp_checksum(dx.ptr, dx.size, dy.ptr, dy.size, dz.ptr, C.byref(out)))
    return out.value
