"""Serving runtime (several jobs in flight, one stream each): every job's
result is bit-identical to the same job run alone through the plain entry
points, whose postsolve the oracle checks; FIFO admission by a memory
budget; the streaming submit/wait form; presolve-only jobs."""
import numpy as np
import pytest

import gp_ffi
import lp_synth
import oracle_ffi as O

pytestmark = pytest.mark.gpu

SEED = 0x5EED
SHAPES = ((300, 1), (1200, 2), (4000, 3), (800, 4))
TOL = 1e-12


@pytest.fixture(scope="module")
def lps():
    return [lp_synth.generate(m, int(1.5 * m), seed=s, mean_row_len=6.0) for m, s in SHAPES]


@pytest.fixture(scope="module")
def dlps(lps):
    return [gp_ffi.DeviceLp(lp) for lp in lps]


_plain_cache = {}


def _plain(lps, i, tag):
    """Job (LP i, tag) through gp_presolve, gp_synthetic_point and gp_postsolve;
    the postsolve is compared with the oracle's on the same point."""
    key = (i, tag)
    if key in _plain_cache:
        return _plain_cache[key]
    lp = lps[i]
    prm = gp_ffi.params(enable_structure_specific_rules=0)
    h = gp_ffi.Presolved(gp_ffi.DeviceLp(lp), prm, True)
    try:
        s, rec = h.summary(), h.record()
        red = h.reduced()
        x, y, z = gp_ffi.synthetic_point(SEED ^ tag, h.reduced_view())
    finally:
        h.free()
    assert np.all(x >= red["l"]) and np.all(x <= red["u"])
    assert np.all(np.abs(y) <= 1.0) and np.all(np.abs(z) <= 1.0)
    got = gp_ffi.postsolve(rec, x, y, z, lp)
    want = O.postsolve(rec, x, y, z, lp)
    for name, a, b in zip("xyz", got, want):
        np.testing.assert_allclose(a, b, rtol=TOL, atol=TOL, err_msg=f"postsolve {name}")
    _plain_cache[key] = (s, gp_ffi.checksum(*got))
    return _plain_cache[key]


def _check_results(lps, res, lp_of_job, postsolve=True):
    assert [r["tag"] for r in res] == list(range(len(lp_of_job)))
    for k, r in enumerate(res):
        assert r["status"] == 0, r
        s, ck = _plain(lps, lp_of_job[k], k)
        assert (r["reduced_rows"], r["reduced_cols"], r["reduced_nnz"]) == (
            s["reduced_rows"], s["reduced_cols"], s["reduced_nnz"])
        assert r["checksum"] == (ck if postsolve else 0), f"job {k}"
        assert r["t_arrival"] <= r["t_start"] <= r["t_end"]


def test_batch_matches_plain(lps, dlps):
    lp_of_job = [0, 1, 2, 3] * 3
    cfg = gp_ffi.serve_config(workers=4, seed=SEED)
    res = gp_ffi.serve_trace(dlps, lp_of_job, gp_ffi.params(enable_structure_specific_rules=0), cfg)
    _check_results(lps, res, lp_of_job)
    assert len({r["worker"] for r in res}) > 1


def test_trace_with_arrivals(lps, dlps):
    lp_of_job = [2, 0, 3, 1, 2, 0, 1, 3]
    arrival = np.linspace(0.0, 0.05, len(lp_of_job))
    cfg = gp_ffi.serve_config(workers=3, seed=SEED)
    res = gp_ffi.serve_trace(dlps, lp_of_job, gp_ffi.params(enable_structure_specific_rules=0), cfg, arrival)
    _check_results(lps, res, lp_of_job)
    for k, r in enumerate(res):
        assert r["t_arrival"] == pytest.approx(arrival[k])


def test_budget_admits_one_job_at_a_time(lps, dlps):
    lp_of_job = [1, 2, 0, 3, 2]
    cfg = gp_ffi.serve_config(workers=4, hbm_budget_bytes=1.0, seed=SEED)
    res = gp_ffi.serve_trace(dlps, lp_of_job, gp_ffi.params(enable_structure_specific_rules=0), cfg)
    _check_results(lps, res, lp_of_job)
    spans = sorted((r["t_start"], r["t_end"]) for r in res)
    for (s0, e0), (s1, e1) in zip(spans, spans[1:]):
        assert s1 >= e0, "two jobs overlapped under a budget that fits one"


def test_presolve_only(lps, dlps):
    lp_of_job = [3, 1]
    cfg = gp_ffi.serve_config(workers=2, postsolve=False, seed=SEED)
    res = gp_ffi.serve_trace(dlps, lp_of_job, gp_ffi.params(enable_structure_specific_rules=0), cfg)
    _check_results(lps, res, lp_of_job, postsolve=False)


def test_streaming_submit_wait(lps, dlps):
    srv = gp_ffi.Server(gp_ffi.params(enable_structure_specific_rules=0), gp_ffi.serve_config(workers=2, seed=SEED))
    try:
        lp_of_job = [0, 2, 1, 3, 0, 2]
        for k, i in enumerate(lp_of_job):
            assert srv.submit(dlps[i], tag=k) == k
        got = []
        while len(got) < len(lp_of_job):
            got += srv.wait(timeout_s=5.0)
        got.sort(key=lambda r: r["tag"])
        _check_results(lps, got, lp_of_job)
    finally:
        srv.close()
