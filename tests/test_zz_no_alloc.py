"""gp_row_activity_noalloc (the allocation-free row activity) against
gp_row_activity: bit-identical outputs on two calls in a row, the second on a
smaller LP (the fixed staging array keeps the first call's pairs beyond the
second LP's columns), on the null stream and on a caller's stream; a column
count beyond the fixed array is rejected before anything is launched."""
import ctypes as C

import numpy as np
import pytest

import gp_ffi
import lp_synth

pytestmark = pytest.mark.gpu

GP_ERR_INVALID_ARGUMENT = 1  # gp_status (gp_presolve_capi.h)
MAX_COLS = 2_097_152  # GP_ROW_ACTIVITY_NOALLOC_MAX_COLS


def _outputs(m):
    return (gp_ffi.DeviceArray(size=m), gp_ffi.DeviceArray(size=m),
            gp_ffi.DeviceArray(size=m, dtype=np.int32), gp_ffi.DeviceArray(size=m, dtype=np.int32))


def _call(fn, A, dl, du, ztol, outs, stream):
    mn, mx, ni, pi = outs
    return fn(C.byref(A.g), dl.ptr, du.ptr, ztol, mn.ptr, mx.ptr, ni.ptr, pi.ptr, stream)


def _both(lp, stream=None):
    m, n = lp["m"], lp["n"]
    A = gp_ffi.DeviceCsr(m, n, lp["A_ptr"], lp["A_idx"], lp["A_val"])
    l = lp["l"].copy()
    u = lp["u"].copy()
    l[::7] = -np.inf  # infinite bounds on both sides reach the counters
    dl, du = gp_ffi.DeviceArray(l), gp_ffi.DeviceArray(u)
    ref, got = _outputs(m), _outputs(m)
    s = None if stream is None else stream.ptr
    gp_ffi._check(_call(gp_ffi.lib().gp_row_activity, A, dl, du, 1e-9, ref, s))
    gp_ffi._check(_call(gp_ffi.lib().gp_row_activity_noalloc, A, dl, du, 1e-9, got, s))
    if stream is not None:
        stream.sync()
    out = []
    for r, g in zip(ref, got):
        a, b = r.numpy(), g.numpy()
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
        out.append(a)
    return out


@pytest.mark.parametrize("use_stream", [False, True])
def test_noalloc_matches_row_activity_on_two_calls(use_stream):
    stream = gp_ffi.Stream() if use_stream else None
    big = lp_synth.generate(20_000, 30_000, seed=61, mean_row_len=6.0)
    small = lp_synth.generate(300, 450, seed=62, mean_row_len=6.0)
    first = _both(big, stream)
    second = _both(small, stream)
    assert first[0].shape == (20_000,) and second[0].shape == (300,)
    for outs in (first, second):  # both infinite-term counters are exercised
        assert (outs[2] > 0).any() and (outs[3] > 0).any()


def test_noalloc_rejects_too_many_columns():
    n = MAX_COLS
    A = gp_ffi.DeviceCsr(2, n + 1, np.array([0, 1, 2], np.int32), np.array([0, n], np.int32),
                         np.array([1.0, 1.0]))
    dl, du = gp_ffi.DeviceArray(np.zeros(4)), gp_ffi.DeviceArray(np.ones(4))
    rc = _call(gp_ffi.lib().gp_row_activity_noalloc, A, dl, du, 1e-9, _outputs(2), None)
    assert rc == GP_ERR_INVALID_ARGUMENT
    assert b"2,097,152" in gp_ffi.lib().gp_last_error()
