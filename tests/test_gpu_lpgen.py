"""The device-side LP generator (gp_generate_lp) and jobs on generated LPs.

The generator must give a well-formed LP (rows sorted without repeated
columns, AT the exact transpose of A, dyadic values, feasible by
construction through its point x0) with the planted fractions of its
description, and the same LP bit for bit for the same spec.  Served jobs
that generate their LP on the worker must equal the same job through the
plain entry points (gp_plain_job_spec), and the plain presolve and postsolve
of generated LPs must equal the oracle's."""
import numpy as np
import pytest

import gp_ffi
import oracle_ffi as O
import workloads

pytestmark = pytest.mark.gpu

SEED = 0x5EC
TOL = 1e-12


def _spec(seed, m, row_len=(4, 10)):
    return gp_ffi.lp_spec(seed, m, row_len=row_len)


def _gen(spec):
    g = gp_ffi.GeneratedLp(spec)
    try:
        return g.host()
    finally:
        g.free()


def _check_well_formed(lp):
    m, n = lp["m"], lp["n"]
    ptr, idx, val = lp["A_ptr"], lp["A_idx"], lp["A_val"]
    assert ptr[0] == 0 and np.all(np.diff(ptr) >= 0) and ptr[-1] == len(idx)
    assert idx.min() >= 0 and idx.max() < n
    rows = np.repeat(np.arange(m), np.diff(ptr))
    # strictly increasing columns inside each row
    same_row = rows[1:] == rows[:-1]
    assert np.all(idx[1:][same_row] > idx[:-1][same_row])
    # AT is the exact transpose, rows sorted
    order = np.lexsort((rows, idx))
    tptr = np.zeros(n + 1, np.int64)
    np.add.at(tptr, idx.astype(np.int64) + 1, 1)
    assert np.array_equal(lp["AT_ptr"], np.cumsum(tptr))
    assert np.array_equal(lp["AT_idx"], rows[order])
    assert np.array_equal(lp["AT_val"], val[order])
    # dyadic values: duplicate rows scale a 1/8-grid coefficient by 2^k, k >= -2
    assert np.all(np.abs(val) >= 0.125) and np.all(np.abs(val) <= 8.0)
    assert np.all(val * 32.0 == np.round(val * 32.0))
    for k in ("c", "l", "u"):
        f = lp[k][np.isfinite(lp[k])]
        assert np.all(f * 1024.0 == np.round(f * 1024.0)), k
    assert np.all(lp["l"] <= lp["u"]) and np.all(np.isfinite(lp["l"]))
    assert np.all(lp["c"][~np.isfinite(lp["u"])] >= 0.0)
    assert np.all(lp["AL"] <= lp["AU"])


def test_generator_is_deterministic_and_well_formed():
    a = _gen(_spec(11, 3000))
    b = _gen(_spec(11, 3000))
    c = _gen(_spec(12, 3000))
    for k in a:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k
    assert not (len(a["A_idx"]) == len(c["A_idx"]) and np.array_equal(a["A_idx"], c["A_idx"]))
    for lp in (a, c, _gen(gp_ffi.lp_spec(5, 700, mean_len=6.0, min_len=1))):
        _check_well_formed(lp)


def test_generator_fractions():
    m = 200_000
    lp = _gen(_spec(21, m))
    n = lp["n"]
    col_nnz = np.diff(lp["AT_ptr"])
    fixed = lp["l"] == lp["u"]
    assert 0.012 < np.mean(col_nnz == 0) < 0.03          # 1% planted empty, ~1% never drawn
    assert 0.015 < np.mean(fixed) < 0.025                  # 2% fixed
    assert 0.17 < np.mean(~np.isfinite(lp["u"])) < 0.23   # 20% u = +inf
    row_nnz = np.diff(lp["A_ptr"])
    assert 0.04 < np.mean(row_nnz == 1) < 0.06             # 5% singleton rows
    eq = lp["AL"] == lp["AU"]
    assert 0.30 < np.mean(eq) < 0.36                       # 30% equalities + doubletons
    one = ~np.isfinite(lp["AL"]) | ~np.isfinite(lp["AU"])
    assert 0.06 < np.mean(one) < 0.11                      # 10% of the non-singleton, non-doubleton rows
    general = row_nnz[(row_nnz >= 3)]
    assert general.max() <= 10 + 6                        # U[4,10] plus a few singleton-column entries
    assert n == int(round(1.5 * m))


def test_generated_lp_presolve_matches_oracle():
    kw = dict(enable_structure_specific_rules=0)
    for spec in (_spec(31, 2500), gp_ffi.lp_spec(32, 800, mean_len=6.0, min_len=1), _spec(33, 40_000)):
        g = gp_ffi.GeneratedLp(spec)
        lp = g.host()
        s_gpu, rec_gpu, red_gpu = gp_ffi.presolve(lp, gp_ffi.params(**kw), True)
        s_cpu, rec_cpu, red_cpu = O.presolve(lp, O.params(**kw), True)
        assert s_cpu["has_infeasible"] == 0 and s_cpu["has_unbounded"] == 0
        assert s_cpu["reduced_rows"] < lp["m"]
        for k in ("reduced_rows", "reduced_cols", "reduced_nnz", "iterations"):
            assert s_gpu[k] == s_cpu[k], k
        for k in ("row_red2org", "col_red2org", "fixed_idx", "fixed_val", "types", "indices", "vals"):
            assert np.array_equal(rec_gpu[k], rec_cpu[k]), k
        for k in ("A_ptr", "A_idx", "A_val", "c", "l", "u", "AL", "AU"):
            assert np.array_equal(red_gpu[k], red_cpu[k]), k
        # the plain job on the generated device LP equals the one on its host copy
        p = gp_ffi.plain_job(g, gp_ffi.params(**kw), SEED, 3, True)
        q = gp_ffi.plain_job_spec(spec, gp_ffi.params(**kw), SEED, 3, True)
        r = gp_ffi.plain_job(gp_ffi.DeviceLp(lp), gp_ffi.params(**kw), SEED, 3, True)
        assert p["checksum"] == q["checksum"] == r["checksum"] != 0
        g.free()


@pytest.mark.timeout(900)
def test_served_generated_requests_match_plain():
    """A Poisson(200/s) trace of requests whose LPs (m log-uniform in
    [1,000, 500,000], n = 1.5 m, 4-10 entries per general row) are generated
    on the serving workers, three of them above 100,000 rows."""
    rng = np.random.default_rng(44)
    arrival, _ = workloads.poisson_trace(200.0, 0.6, 1, seed=45)
    sizes = workloads.log_uniform_sizes(rng, len(arrival), 1_000, 500_000)
    sizes[:3] = [120_000, 260_000, 480_000]
    specs = [_spec(1000 + k, int(m)) for k, m in enumerate(sizes)]
    prm = gp_ffi.params(enable_structure_specific_rules=0)
    cfg = gp_ffi.serve_config(workers=12, postsolve=True, hbm_budget_bytes=200e9, seed=SEED)
    res = gp_ffi.serve_spec_trace(specs, prm, cfg, arrival)
    assert len({r["worker"] for r in res}) > 1
    for k, r in enumerate(res):
        assert r["status"] == 0, r
        assert r["t_arrival"] <= r["t_start"] <= r["t_ready"] <= r["t_end"]
        p = gp_ffi.plain_job_spec(specs[k], prm, SEED, k, True)
        got = (r["reduced_rows"], r["reduced_cols"], r["reduced_nnz"], r["iterations"], r["checksum"])
        want = (p["reduced_rows"], p["reduced_cols"], p["reduced_nnz"], p["iterations"], p["checksum"])
        assert got == want, f"request {k} ({sizes[k]} rows)"
    # the oracle on a large and a small request
    for k in (1, int(np.argmin(sizes))):
        g = gp_ffi.GeneratedLp(specs[k])
        lp = g.host()
        g.free()
        kw = dict(enable_structure_specific_rules=0)
        s_gpu, rec_gpu, _ = gp_ffi.presolve(lp, gp_ffi.params(**kw), False)
        s_cpu, rec_cpu, red_cpu = O.presolve(lp, O.params(**kw), True)
        for key in ("types", "indices", "vals", "row_red2org", "col_red2org", "fixed_idx"):
            assert np.array_equal(rec_gpu[key], rec_cpu[key]), key
        h = gp_ffi.Presolved(gp_ffi.DeviceLp(lp), gp_ffi.params(**kw), True)
        try:
            x, y, z = gp_ffi.synthetic_point(SEED ^ k, h.reduced_view())
        finally:
            h.free()
        got = gp_ffi.postsolve(rec_gpu, x, y, z, lp)
        want = O.postsolve(rec_cpu, x, y, z, lp)
        for name, a, b in zip("xyz", got, want):
            np.testing.assert_allclose(a, b, rtol=TOL, atol=TOL, err_msg=f"request {k} postsolve {name}")
        assert gp_ffi.checksum(*got) == res[k]["checksum"]
