"""Batched and served workloads at their stated sizes.

Batched: 4,096 independent LPs (m log-uniform in [100, 20,000], n = 1.5 m,
row lengths Poisson(6) with minimum 1) submitted at once to the serving
runtime with twelve workers.  Served: a Poisson(200/s) trace of presolve +
postsolve requests over LPs drawn from m log-uniform in [1,000, 500,000]
(n = 1.5 m, 4-10 nonzeros per row), at least three of them above 100,000
rows, under a 200 GB admission budget.

Every job's reduced dimensions and postsolve checksum must equal those of the
same job run alone through the plain synchronous entry points
(gp_plain_job); on a seeded sample of LPs the plain presolve record, reduced
LP and postsolve are compared with the oracle (record and reduced LP bit for
bit, postsolve within 1e-12)."""
import numpy as np
import pytest

import gp_ffi
import lp_synth
import oracle_ffi as O
import workloads

pytestmark = pytest.mark.gpu

SEED = 0xC3C4
TOL = 1e-12


def _prm():
    return gp_ffi.params(enable_structure_specific_rules=0)


def _oracle_check(lp, tag):
    """Plain GPU presolve + seeded point + postsolve of lp against the oracle;
    returns the GPU postsolve's checksum."""
    kw = dict(enable_structure_specific_rules=0)
    s_gpu, rec_gpu, red_gpu = gp_ffi.presolve(lp, gp_ffi.params(**kw), True)
    s_cpu, rec_cpu, red_cpu = O.presolve(lp, O.params(**kw), True)
    for k in ("reduced_rows", "reduced_cols", "reduced_nnz", "iterations", "has_infeasible", "has_unbounded"):
        assert s_gpu[k] == s_cpu[k], k
    for k in ("row_red2org", "col_red2org", "fixed_idx", "fixed_val", "types", "indices", "vals"):
        assert np.array_equal(rec_gpu[k], rec_cpu[k]), f"record {k}"
    for k in ("A_ptr", "A_idx", "A_val", "c", "l", "u", "AL", "AU"):
        assert np.array_equal(red_gpu[k], red_cpu[k]), f"reduced LP {k}"
    h = gp_ffi.Presolved(gp_ffi.DeviceLp(lp), gp_ffi.params(**kw), True)
    try:
        x, y, z = gp_ffi.synthetic_point(SEED ^ tag, h.reduced_view())
    finally:
        h.free()
    got = gp_ffi.postsolve(rec_gpu, x, y, z, lp)
    want = O.postsolve(rec_cpu, x, y, z, lp)
    for name, a, b in zip("xyz", got, want):
        np.testing.assert_allclose(a, b, rtol=TOL, atol=TOL, err_msg=f"postsolve {name}")
    return gp_ffi.checksum(*got)


def _compare_with_plain(dlps, lp_of_job, res, prm):
    assert [r["tag"] for r in res] == list(range(len(lp_of_job)))
    plain = {}
    for k, r in enumerate(res):
        assert r["status"] == 0, r
        p = gp_ffi.plain_job(dlps[lp_of_job[k]], prm, SEED, k, True)
        assert p["status"] == 0, p
        got = (r["reduced_rows"], r["reduced_cols"], r["reduced_nnz"], r["iterations"], r["checksum"])
        want = (p["reduced_rows"], p["reduced_cols"], p["reduced_nnz"], p["iterations"], p["checksum"])
        assert got == want, f"job {k} (LP {lp_of_job[k]}): served {got}, plain {want}"
        plain[k] = p
    return plain


@pytest.mark.timeout(900)
def test_batched_4096_lps_match_plain():
    lps = workloads.batched_pool(4096)
    ms = np.array([lp["m"] for lp in lps])
    assert len(lps) == 4096 and ms.min() >= 100 and ms.max() <= 20_000
    lens = np.concatenate([np.diff(lp["A_ptr"]) for lp in lps[:64]])
    assert lens.min() >= 1
    dlps = [gp_ffi.DeviceLp(lp) for lp in lps]
    prm = _prm()
    cfg = gp_ffi.serve_config(workers=12, postsolve=True, seed=SEED)
    jobs = np.arange(len(dlps), dtype=np.int32)
    res = gp_ffi.serve_trace(dlps, jobs, prm, cfg)
    assert len({r["worker"] for r in res}) > 1
    plain = _compare_with_plain(dlps, jobs, res, prm)
    sample = np.random.default_rng(31).choice(len(lps), size=12, replace=False)
    sample = list(sample) + [int(np.argmax(ms)), int(np.argmin(ms))]
    for i in sample:
        assert _oracle_check(lps[i], int(i)) == plain[int(i)]["checksum"], f"LP {i}"


def _serving_lps():
    """LPs of the served size distribution: m log-uniform in [1,000, 500,000]
    with at least three above 100,000 rows."""
    rng = np.random.default_rng(18)
    sizes = workloads.log_uniform_sizes(rng, 14, 1_000, 500_000)
    assert (sizes > 100_000).sum() >= 3, sizes
    return workloads.generate_pool(sizes, 18, row_len=(4, 10)), sizes


@pytest.mark.timeout(900)
def test_serving_trace_large_lps_match_plain():
    lps, sizes = _serving_lps()
    dlps = [gp_ffi.DeviceLp(lp) for lp in lps]
    prm = _prm()
    arrival, lp_of_job = workloads.poisson_trace(200.0, 0.75, len(dlps), seed=23)
    # every LP of the pool is requested at least once
    lp_of_job = np.concatenate([np.arange(len(dlps), dtype=np.int32), lp_of_job[len(dlps):]])
    cfg = gp_ffi.serve_config(workers=12, postsolve=True, hbm_budget_bytes=200e9, seed=SEED)
    res = gp_ffi.serve_trace(dlps, lp_of_job, prm, cfg, arrival)
    plain = _compare_with_plain(dlps, lp_of_job, res, prm)
    for r, t in zip(res, arrival):
        assert r["t_arrival"] == pytest.approx(t) and r["t_arrival"] <= r["t_start"] <= r["t_end"]
    # overlap: some job started before an earlier-started one had finished
    spans = sorted((r["t_start"], r["t_end"]) for r in res)
    assert any(s1 < e0 for (s0, e0), (s1, e1) in zip(spans, spans[1:]))
    big = [k for k in range(len(lps)) if sizes[k] > 100_000]
    small = [k for k in range(len(lps)) if sizes[k] <= 100_000]
    for i in big[:2] + small[:1]:
        assert _oracle_check(lps[i], i) == plain[i]["checksum"], f"LP {i} ({sizes[i]} rows)"


def test_plain_job_matches_python_path():
    """gp_plain_job equals the job assembled from the separate entry points."""
    lp = lp_synth.generate(900, 1350, seed=41, mean_row_len=6.0, min_row_len=1)
    d = gp_ffi.DeviceLp(lp)
    p = gp_ffi.plain_job(d, _prm(), SEED, 5, True)
    assert p["status"] == 0 and p["worker"] == -1
    assert _oracle_check(lp, 5) == p["checksum"]
    q = gp_ffi.plain_job(d, _prm(), SEED, 5, False)
    assert q["checksum"] == 0 and q["reduced_rows"] == p["reduced_rows"]
