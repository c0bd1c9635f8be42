"""Failure handling of the GPU library, through the C-ABI: a call that fails
part way leaves the thread's next call unaffected, and LP arrays that break
the CSR contract are rejected with GP_ERR_INVALID_ARGUMENT before any kernel
reads them."""
import ctypes as C

import numpy as np
import pytest

import gp_ffi
import lp_synth
from helpers import assert_same_lp, assert_same_record

pytestmark = pytest.mark.gpu

GP_ERR_INVALID_ARGUMENT = 1
GP_ERR_INTERNAL = 3


def _status(fn, *args):
    """Runs one C-ABI call and returns (status, last error text)."""
    rc = fn(*args)
    return rc, gp_ffi.lib().gp_last_error().decode()


def _presolve_status(lp, prm):
    dlp = gp_ffi.DeviceLp(lp)
    h = C.c_void_p()
    rc, msg = _status(gp_ffi.lib().gp_presolve, C.byref(dlp.g), C.byref(prm), 1, C.byref(h))
    if h:
        gp_ffi.lib().gp_presolve_free(h)
    return rc, msg


@pytest.mark.parametrize("after", [1, 2, 5])
def test_failed_presolve_leaves_next_call_unchanged(after):
    """The armed record update throws right after queuing the record's host
    copies; the copies are dropped, and the next presolve on this thread gives
    the same record and reduced LP as before the failure."""
    L = gp_ffi.lib()
    L.gp_test_fail_after_record_updates.argtypes = [C.c_int32]
    L.gp_test_fail_after_record_updates.restype = None
    lp = lp_synth.config0(seed=0)
    prm = gp_ffi.params(enable_structure_specific_rules=0)
    s0, r0, l0 = gp_ffi.presolve(lp, prm, True)
    L.gp_test_fail_after_record_updates(after)
    try:
        rc, msg = _presolve_status(lp, prm)
    finally:
        L.gp_test_fail_after_record_updates(0)
    assert rc == GP_ERR_INTERNAL and "injected failure" in msg, (rc, msg)
    for _ in range(2):
        s1, r1, l1 = gp_ffi.presolve(lp, prm, True)
        assert s1["reduced_rows"] == s0["reduced_rows"] and s1["reduced_cols"] == s0["reduced_cols"]
        assert_same_record(r1, r0)
        assert_same_lp(l1, l0)


def _with_rows(lp, fix):
    """A copy of lp whose A (and AT, rebuilt by a plain transpose) is edited
    by fix(ptr, idx, val) -> (ptr, idx, val)."""
    out = dict(lp)
    ptr, idx, val = fix(lp["A_ptr"].copy(), lp["A_idx"].copy(), lp["A_val"].copy())
    out["A_ptr"], out["A_idx"], out["A_val"] = ptr, idx, val
    m, n = lp["m"], lp["n"]
    rows = np.repeat(np.arange(m), np.diff(ptr))
    order = np.lexsort((rows, idx))
    at_ptr = np.zeros(n + 1, np.int32)
    np.add.at(at_ptr, idx + 1, 1)
    out["AT_ptr"] = np.cumsum(at_ptr).astype(np.int32)
    out["AT_idx"] = rows[order].astype(np.int32)
    out["AT_val"] = val[order]
    return out


def _first_long_row(ptr):
    return int(np.nonzero(np.diff(ptr) >= 3)[0][0])


def test_repeated_column_in_a_row_is_rejected():
    lp = lp_synth.generate(200, 300, seed=21, mean_row_len=5.0)

    def dup(ptr, idx, val):
        i = _first_long_row(ptr)
        idx[ptr[i] + 1] = idx[ptr[i]]  # the row now holds its first column twice
        return ptr, idx, val

    bad = _with_rows(lp, dup)
    rc, msg = _presolve_status(bad, gp_ffi.params(enable_structure_specific_rules=0))
    assert rc == GP_ERR_INVALID_ARGUMENT and "repeats a column" in msg, (rc, msg)


def test_repeated_column_in_an_unsorted_row_is_rejected():
    lp = lp_synth.generate(200, 300, seed=22, mean_row_len=5.0)

    def dup(ptr, idx, val):
        i = _first_long_row(ptr)
        b = ptr[i]
        # last column first, then the rest, then the first column again: not
        # sorted, and the repeat is not between neighbours
        row = idx[b:ptr[i + 1]].copy()
        idx[b:ptr[i + 1]] = np.concatenate([row[-1:], row[:-1]])
        idx[ptr[i + 1] - 1] = idx[b + 1]
        return ptr, idx, val

    bad = _with_rows(lp, dup)
    rc, msg = _presolve_status(bad, gp_ffi.params(enable_structure_specific_rules=0))
    assert rc == GP_ERR_INVALID_ARGUMENT and "repeats a column" in msg, (rc, msg)


def test_column_index_out_of_range_is_rejected():
    lp = lp_synth.generate(200, 300, seed=23, mean_row_len=5.0)
    bad = dict(lp)
    idx = lp["A_idx"].copy()
    idx[-1] = lp["n"] + 7
    bad["A_idx"] = idx
    rc, msg = _presolve_status(bad, gp_ffi.params(enable_structure_specific_rules=0))
    assert rc == GP_ERR_INVALID_ARGUMENT and "outside" in msg, (rc, msg)


def test_bad_row_pointers_are_rejected():
    lp = lp_synth.generate(200, 300, seed=24, mean_row_len=5.0)
    bad = dict(lp)
    ptr = lp["A_ptr"].copy()
    ptr[5], ptr[6] = ptr[6], ptr[5] - 1
    bad["A_ptr"] = ptr
    rc, msg = _presolve_status(bad, gp_ffi.params(enable_structure_specific_rules=0))
    assert rc == GP_ERR_INVALID_ARGUMENT and "row pointers" in msg, (rc, msg)


def test_postsolve_rejects_original_with_repeated_column():
    lp = lp_synth.generate(200, 300, seed=25, mean_row_len=5.0)
    s, rec, red = gp_ffi.presolve(lp, gp_ffi.params(enable_structure_specific_rules=0), True)
    x, y, z = lp_synth.reduced_solution(3, red["m"], red["n"], red["l"], red["u"])

    def dup(ptr, idx, val):
        i = _first_long_row(ptr)
        idx[ptr[i] + 1] = idx[ptr[i]]
        return ptr, idx, val

    bad = _with_rows(lp, dup)
    with pytest.raises(gp_ffi.GpError, match="repeats a column"):
        gp_ffi.postsolve(rec, x, y, z, bad)
