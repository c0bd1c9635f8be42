"""HIP primitives against the oracle: row activity, column max |a|,
compaction and transpose (bit-exact; integer/permutation work and dyadic
FP64 sums)."""
import numpy as np
import pytest

import gp_ffi
import lp_synth
import oracle_ffi as O
from helpers import assert_same_array

pytestmark = pytest.mark.gpu


def _random_csr(m, n, seed, max_len=12, long_rows=()):
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for i in range(m):
        k = int(rng.integers(0, max_len + 1))
        if i in long_rows:
            k = min(n, long_rows[i])
        cs = np.sort(rng.choice(n, size=min(k, n), replace=False)) if n else np.zeros(0, int)
        rows += [i] * len(cs)
        cols += list(cs)
    vals = np.round(rng.uniform(0.5, 2.0, len(rows)) * 8) / 8 * rng.choice([-1.0, 1.0], len(rows))
    ptr, idx, val = lp_synth.csr_from_coo(m, n, np.array(rows, np.int64), np.array(cols, np.int64), vals)
    return ptr, idx, val


def _bounds(n, seed):
    rng = np.random.default_rng(seed)
    l = np.round(rng.uniform(-4, 0, n) * 1024) / 1024
    u = np.round(rng.uniform(0, 4, n) * 1024) / 1024
    l[rng.random(n) < 0.15] = -np.inf
    u[rng.random(n) < 0.15] = np.inf
    return l, u


@pytest.mark.parametrize("m,n,seed,long_rows", [
    (0, 5, 1, {}), (7, 0, 2, {}), (1, 1, 3, {}), (257, 300, 4, {}),
    (5000, 4000, 5, {3: 3000, 4000: 2500}),  # rows longer than the LDS window
    (40000, 30000, 6, {}),
])
def test_row_activity(m, n, seed, long_rows):
    ptr, idx, val = _random_csr(m, n, seed, long_rows=long_rows)
    if len(val) > 3:
        val[::7] = 0.0  # explicit zeros are skipped by |a| > zero_tol
    l, u = _bounds(n, seed + 100)
    got = gp_ffi.row_activity(m, n, ptr, idx, val, l, u, 1e-10)
    want = O.row_activity(m, n, ptr, idx, val, l, u, 1e-10)
    for name, a, b in zip(("min_fin", "max_fin", "neg_inf", "pos_inf"), got, want):
        assert_same_array(name, a, b)


def test_row_activity_on_caller_stream():
    ptr, idx, val = _random_csr(3000, 2000, 9)
    l, u = _bounds(2000, 10)
    s = gp_ffi.Stream()
    got = gp_ffi.row_activity(3000, 2000, ptr, idx, val, l, u, 1e-10, stream=s.ptr)
    want = O.row_activity(3000, 2000, ptr, idx, val, l, u, 1e-10)
    for name, a, b in zip(("min_fin", "max_fin", "neg_inf", "pos_inf"), got, want):
        assert_same_array(name, a, b)


@pytest.mark.parametrize("m,n,seed", [(0, 3, 1), (4, 0, 2), (300, 200, 3), (20000, 30000, 4)])
def test_col_max_abs(m, n, seed):
    ptr, idx, val = _random_csr(m, n, seed)
    tp, ti, tv = lp_synth.transpose_csr(m, n, ptr, idx, val)
    assert_same_array("col_max_abs", gp_ffi.col_max_abs(n, m, tp, ti, tv), O.col_max_abs(n, m, tp, ti, tv))


# 65,536 rows are 256 tiles (offsets derived in the write pass), 65,537 rows
# 257 tiles (a scan of the tile counts).
@pytest.mark.parametrize("m,n,seed,frac", [(0, 4, 1, 0.5), (5, 0, 2, 0.5), (300, 400, 3, 0.3),
                                          (20000, 25000, 4, 0.15), (1000, 800, 5, 1.0), (1000, 800, 6, 0.0),
                                          (65536, 3000, 7, 0.2), (65537, 3000, 8, 0.2)])
def test_compact(m, n, seed, frac):
    ptr, idx, val = _random_csr(m, n, seed)
    rng = np.random.default_rng(seed)
    kr = (rng.random(m) >= frac).astype(np.uint8)
    kc = (rng.random(n) >= frac).astype(np.uint8)
    got = gp_ffi.compact_csr(m, n, ptr, idx, val, kr, kc)
    want = O.compact_csr(m, n, ptr, idx, val, kr, kc)
    assert got[:2] == want[:2]
    for name, a, b in zip(("ptr", "idx", "val"), got[2:], want[2:]):
        assert_same_array(name, a, b)


@pytest.mark.parametrize("m,n,seed,frac_r,frac_c", [(20000, 25000, 10, 0.3, 0.0), (20000, 25000, 11, 0.0, 0.3),
                                                   (65537, 3000, 12, 0.2, 0.0)])
def test_compact_one_side(m, n, seed, frac_r, frac_c):
    """Only rows or only columns dropped: with every column kept the entries
    keep their column indices without the rank words."""
    ptr, idx, val = _random_csr(m, n, seed)
    rng = np.random.default_rng(seed)
    kr = (rng.random(m) >= frac_r).astype(np.uint8)
    kc = (rng.random(n) >= frac_c).astype(np.uint8)
    got = gp_ffi.compact_csr(m, n, ptr, idx, val, kr, kc)
    want = O.compact_csr(m, n, ptr, idx, val, kr, kc)
    assert got[:2] == want[:2]
    for name, a, b in zip(("ptr", "idx", "val"), got[2:], want[2:]):
        assert_same_array(name, a, b)


def test_compact_many_tiles():
    """More than 16,384 tiles of 256 rows: the tile offsets come from the
    multi-workgroup look-back scan instead of the one-workgroup scan."""
    m, n = 4_200_000, 50_000
    rng = np.random.default_rng(9)
    lens = rng.integers(0, 4, size=m)
    rows = np.repeat(np.arange(m, dtype=np.int64), lens)
    cols = rng.integers(0, n, size=rows.size)
    vals = rng.choice([-1.5, 0.5, 2.0], size=rows.size)
    ptr, idx, val = lp_synth.csr_from_coo(m, n, rows, cols, vals)
    kr = (rng.random(m) >= 0.2).astype(np.uint8)
    kc = (rng.random(n) >= 0.2).astype(np.uint8)
    got = gp_ffi.compact_csr(m, n, ptr, idx, val, kr, kc)
    want = O.compact_csr(m, n, ptr, idx, val, kr, kc)
    assert got[:2] == want[:2]
    for name, a, b in zip(("ptr", "idx", "val"), got[2:], want[2:]):
        assert_same_array(name, a, b)


@pytest.mark.parametrize("m,n,seed", [(0, 3, 1), (3, 0, 2), (1, 1, 3), (300, 400, 4), (30000, 20000, 5),
                                      (2000, 70000, 6)])
def test_transpose(m, n, seed):
    ptr, idx, val = _random_csr(m, n, seed)
    got = gp_ffi.transpose_csr(m, n, ptr, idx, val)
    want = O.transpose_csr(m, n, ptr, idx, val)
    assert got[:2] == want[:2]
    for name, a, b in zip(("ptr", "idx", "val"), got[2:], want[2:]):
        assert_same_array(name, a, b)


# The sort has three launch shapes: one workgroup (n <= 4096), per-pass
# histogram + scatter that derives its offsets from the histogram (up to 64
# tiles of 2048 keys), and histogram + scan + scatter beyond that.
@pytest.mark.parametrize("n", [0, 1, 2, 3, 100, 4095, 4096, 4097, 20000, 131072, 131073, 300000])
@pytest.mark.parametrize("dtype,bits", [(np.uint32, 13), (np.uint32, 31), (np.uint64, 64)])
def test_radix_sort_stable(n, dtype, bits):
    rng = np.random.default_rng(n * 7 + bits)
    hi = min(2 ** bits, 2 ** 63)
    # Few distinct keys, so equal keys are common and stability is tested.
    pool = rng.integers(0, hi, size=max(1, n // 8), dtype=np.uint64).astype(dtype)
    keys = pool[rng.integers(0, pool.size, size=n)] if n else np.zeros(0, dtype)
    vals = rng.permutation(n).astype(np.int32)
    got_k, got_v = gp_ffi.radix_sort(keys, vals, bits)
    order = np.argsort(keys, kind="stable")
    assert_same_array("keys", got_k, keys[order])
    assert_same_array("values", got_v, vals[order])
