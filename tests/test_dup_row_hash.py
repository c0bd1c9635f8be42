"""The duplicate-row rule's row hash (rule_duplicate_rows.cu:78-178, chosen at
:484-497): sequential FNV over (length, column, a / a_first) per row, or,
when nnz > 32 m, the XOR of one FNV pair hash per entry mixed into the
length seed.  The reference's tests hold no hash values, so the oracle's two
forms are pinned by what tells them apart:

* both are invariant to scaling a row (every value enters as a / a_first);
* the XOR form is invariant to any reordering of a row's entries after the
  first (the pivot), the sequential form is not;
* the switch sits exactly at nnz = 32 m (32 m: sequential, 32 m + 1: XOR);
* rows dropped from the plan hash to their index, and rows with no entry
  or a pivot at most zero_tol to the FNV mix of their index, in both forms.

The GPU kernel (gpk_dup_row_hash, one thread per row or one wave per row)
must give the oracle's hash bit for bit in both regimes, and a presolve of
an LP with rows of about 40 entries and planted duplicate rows must give the
oracle's record."""
import ctypes as C

import numpy as np
import pytest

import lp_synth
import oracle_ffi as O

CTOL = 1e-9


def _rows_csr(rows, n):
    ptr = np.zeros(len(rows) + 1, np.int32)
    idx, val = [], []
    for i, r in enumerate(rows):
        ptr[i + 1] = ptr[i] + len(r)
        for j, v in r:
            idx.append(j)
            val.append(v)
    return ptr, np.asarray(idx, np.int32), np.asarray(val, np.float64)


def _random_rows(rng, m, n, length):
    rows = []
    for _ in range(m):
        cols = rng.choice(n, size=length, replace=False)
        vals = rng.integers(1, 16, size=length) * 2.0 ** rng.integers(-3, 4, size=length) * \
            rng.choice([-1.0, 1.0], size=length)
        rows.append(list(zip(cols.tolist(), vals.tolist())))
    return rows


def _hash(rows, n, keep=None):
    ptr, idx, val = _rows_csr(rows, n)
    keep = np.ones(len(rows), np.uint8) if keep is None else keep
    return O.dup_row_hashes(len(rows), n, ptr, idx, val, keep, CTOL)


@pytest.mark.parametrize("length,wide", [(8, False), (40, True)])
def test_forms_and_their_invariants(length, wide):
    rng = np.random.default_rng(length)
    n = 400
    rows = _random_rows(rng, 24, n, length)
    base = _hash(rows, n)
    # a row scaled by a power of two: same hash in both forms
    scaled = [list(r) for r in rows]
    scaled[3] = [(j, v * -4.0) for j, v in rows[3]]
    assert _hash(scaled, n)[3] == base[3]
    # entries after the pivot reversed: same hash only in the XOR form
    perm = [list(r) for r in rows]
    perm[5] = [rows[5][0]] + rows[5][1:][::-1]
    assert (_hash(perm, n)[5] == base[5]) == wide
    # distinct rows hash apart
    assert len(set(base.tolist())) == len(rows)


def test_switch_at_32_entries_per_row():
    rng = np.random.default_rng(7)
    n = 200
    rows = _random_rows(rng, 4, n, 32)  # nnz = 32 m: sequential
    perm = [list(r) for r in rows]
    perm[1] = [rows[1][0]] + rows[1][1:][::-1]
    assert _hash(perm, n)[1] != _hash(rows, n)[1]
    extra = [list(r) for r in rows]  # one more entry: nnz = 32 m + 1, XOR form
    used = {j for j, _ in rows[0]}
    extra[0] = rows[0] + [(next(j for j in range(n) if j not in used), 3.0)]
    perm2 = [list(r) for r in extra]
    perm2[1] = [extra[1][0]] + extra[1][1:][::-1]
    assert _hash(perm2, n)[1] == _hash(extra, n)[1]


@pytest.mark.parametrize("length", [8, 40])
def test_special_rows(length):
    rng = np.random.default_rng(3)
    n = 300
    rows = _random_rows(rng, 6, n, length)
    rows[2] = []  # empty row
    rows[4] = [(rows[4][0][0], 1e-12)] + rows[4][1:]  # pivot within zero_tol
    keep = np.ones(len(rows), np.uint8)
    keep[1] = 0
    h = _hash(rows, n, keep)
    assert h[1] == 1  # dropped row: its index
    # the empty row and the tiny-pivot row hash to the FNV mix of their index:
    # apart from each other, and unchanged when the other rows or their own
    # non-pivot entries change
    assert h[2] != h[4]
    rows2 = _random_rows(np.random.default_rng(4), 6, n, length)
    rows2[2] = []
    rows2[4] = [(rows[4][0][0], -1e-13)] + [(j, v * 3.0) for j, v in rows[4][1:]]
    h2 = _hash(rows2, n, keep)
    assert h2[2] == h[2] and h2[4] == h[4]


@pytest.mark.gpu
@pytest.mark.parametrize("length", [8, 40, 300])
def test_gpu_hash_matches(length):
    import gp_ffi
    import gp_types as T

    rng = np.random.default_rng(length + 1)
    n = 2000
    rows = _random_rows(rng, 333, n, length)
    rows[7] = []
    rows[9] = [(rows[9][0][0], 0.0)] + rows[9][1:]
    keep = np.ones(len(rows), np.uint8)
    keep[11] = 0
    ptr, idx, val = _rows_csr(rows, n)
    want = O.dup_row_hashes(len(rows), n, ptr, idx, val, keep, CTOL)
    L = gp_ffi.lib()
    L.gpk_dup_row_hash.argtypes = [T.GpCsr, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p]
    d = [gp_ffi.DeviceArray(a) for a in (ptr, idx, val, keep)]
    out = gp_ffi.DeviceArray(size=len(rows), dtype=np.uint64)
    g = T.GpCsr(len(rows), n, len(idx), d[0].ptr, d[1].ptr, d[2].ptr)
    gp_ffi.hip_check(L.gpk_dup_row_hash(g, d[3].ptr, CTOL, out.ptr, None), "gpk_dup_row_hash")
    gp_ffi.hip_check(gp_ffi.hip().hipDeviceSynchronize(), "sync")
    assert np.array_equal(out.numpy(), want)


@pytest.mark.gpu
def test_gpu_presolve_on_long_rows_with_duplicates():
    import gp_ffi
    from helpers import assert_same_lp, assert_same_record

    lp = lp_synth.generate(600, 3000, seed=21, mean_row_len=40.0, min_row_len=33, frac_dup_rows=0.05)
    assert len(lp["A_idx"]) > 32 * lp["m"]
    kw = dict(enable_structure_specific_rules=0)
    sg, rg, lg = gp_ffi.presolve(lp, gp_ffi.params(**kw), True)
    so, ro, lo = O.presolve(lp, O.params(**kw), True)
    assert (sg["reduced_rows"], sg["reduced_cols"]) == (so["reduced_rows"], so["reduced_cols"])
    assert_same_record(rg, ro)
    assert_same_lp(lg, lo)
