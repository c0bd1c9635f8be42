"""Many-LP workloads for the serving runtime: seeded LP pools, Poisson
arrival traces, latency statistics and the worker-count autotuner.

  batched  4,096 independent LPs, m log-uniform in [100, 20,000], n = 1.5 m,
           row lengths Poisson(6) with minimum 1, presolved together (all
           arrive at once);
  serving  Poisson arrivals (200 requests/s) of presolve + postsolve jobs,
           m log-uniform in [1,000, 500,000], n = 1.5 m, 4-10 nonzeros per row.

Every LP uses the value, bound and planted-reduction distributions of
lp_synth.generate.  Generation runs in a process pool (the generator is
numpy code at a few microseconds per row).
"""
from __future__ import annotations

import multiprocessing as mp
import os

import numpy as np

import lp_synth


def log_uniform_sizes(rng, count, lo, hi):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size=count)).astype(np.int64)


def _gen_one(args):
    m, seed, kw = args
    return lp_synth.generate(int(m), int(round(1.5 * m)), seed=int(seed), **kw)


def generate_pool(sizes, seed, workers=None, **kw):
    """One LP per entry of sizes (rows), seeds seed*1_000_003 + k."""
    jobs = [(m, seed * 1_000_003 + k, kw) for k, m in enumerate(sizes)]
    workers = workers or min(16, os.cpu_count() or 1)
    if workers <= 1 or len(jobs) < 4:
        return [_gen_one(j) for j in jobs]
    with mp.get_context("fork").Pool(workers) as pool:
        return pool.map(_gen_one, jobs, chunksize=max(1, len(jobs) // (4 * workers)))


def batched_pool(count=4096, seed=3, workers=None):
    rng = np.random.default_rng(seed)
    sizes = log_uniform_sizes(rng, count, 100, 20_000)
    return generate_pool(sizes, seed, workers, mean_row_len=6.0, min_row_len=1)


def serving_pool(count=48, seed=4, workers=None):
    rng = np.random.default_rng(seed)
    sizes = log_uniform_sizes(rng, count, 1_000, 500_000)
    return generate_pool(sizes, seed, workers, row_len=(4, 10))


def poisson_trace(rate, duration, pool_size, seed=5):
    """Arrival times (s) of a Poisson process and a uniform pool index per job."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=int(rate * duration * 1.5) + 16)
    t = np.cumsum(gaps)
    t = t[t < duration]
    return t, rng.integers(0, pool_size, size=len(t)).astype(np.int32)


def latency_stats(results):
    lat = np.array([r["t_end"] - r["t_arrival"] for r in results])
    t0 = min(r["t_arrival"] for r in results)
    t1 = max(r["t_end"] for r in results)
    return {"jobs": len(results), "lp_per_s": len(results) / (t1 - t0) if t1 > t0 else 0.0,
            "p50_ms": 1e3 * float(np.percentile(lat, 50)), "p99_ms": 1e3 * float(np.percentile(lat, 99)),
            "max_ms": 1e3 * float(lat.max()), "makespan_s": t1 - t0}


def autotune_workers(run, candidates=(1, 2, 4, 8, 12, 16)):
    """Picks the worker count with the highest throughput: run(w) processes a
    fixed sample with w workers and returns its results."""
    table = {}
    for w in candidates:
        table[w] = latency_stats(run(w))["lp_per_s"]
    best = max(table, key=table.get)
    return best, table
