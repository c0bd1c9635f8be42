"""Host side of the solver's row kernels (CPU): gpk_pdhg_row_blocks, the
partition of a CSR's rows into the workgroup blocks the device kernels
assume -- contiguous, covering every row once, a row of more than
GPK_PDHG_LONG_ROW entries alone in its block, and any block of several rows
within 256 rows and GPK_PDHG_ENTRY_BUDGET entries."""
import ctypes as C
import os

import numpy as np
import pytest

import lp_synth

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gpu-presolver_amd", "lib",
                   "libgpu_presolver_amd.so")
LONG_ROW, BUDGET = 256, 2048


def _blocks(ptr):
    L = C.CDLL(LIB)
    L.gpk_pdhg_row_blocks.restype = C.c_int32
    L.gpk_pdhg_row_blocks.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
    ptr = np.ascontiguousarray(ptr, np.int32)
    rows = len(ptr) - 1
    blk = np.full(rows + 1, -1, np.int32)
    nb = L.gpk_pdhg_row_blocks(ptr.ctypes.data, rows, blk.ctypes.data)
    assert L.gpk_pdhg_row_blocks(ptr.ctypes.data, rows, None) == nb
    return blk[:nb + 1]


def _check(ptr):
    b = _blocks(ptr)
    rows = len(ptr) - 1
    assert b[0] == 0 and b[-1] == rows
    nr = np.diff(b)
    assert np.all(nr >= 1)
    cnt = ptr[b[1:]] - ptr[b[:-1]]
    multi = nr > 1
    assert np.all(nr[multi] <= 256) and np.all(cnt[multi] <= BUDGET)
    lens = np.diff(ptr)
    single = set(b[:-1][nr == 1].tolist())
    assert all(r in single for r in np.nonzero(lens > LONG_ROW)[0].tolist())
    return b


@pytest.mark.parametrize("lens", [
    [], [0], [0] * 1000, [3] * 1000, [257, 1, 1, 256, 300, 0, 2048, 1], [1] * 300 + [5000] + [1] * 10,
    [200] * 30,  # 10 rows of 200 fill the entry budget
])
def test_partition(lens):
    ptr = np.concatenate([[0], np.cumsum(np.asarray(lens, np.int64))]).astype(np.int32)
    b = _check(ptr)
    if lens == [200] * 30:
        assert np.diff(b).tolist() == [10, 10, 10]
    if lens == [3] * 1000:
        assert np.diff(b).max() == 256


def test_zipf_columns():
    lp = lp_synth.generate(20000, 30000, seed=5, row_len=(2, 14), zipf=1.5)
    _check(lp["A_ptr"])
    b = _check(lp["AT_ptr"])
    assert len(b) > 1
