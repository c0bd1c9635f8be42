"""Writes long_rows_presolve.json: the oracle's presolve of the large LP with
row lengths U[2,64] (lp_synth.config1(seed=1, row_len=(2, 64)): 1M x 1.5M,
about 30.5M nonzeros) with the structure-specific rules off, as digests of
the record and reduced-LP arrays, the summary, the number of batched
singleton closure runs, and norms of the oracle's postsolve of a seeded
reduced point.  The oracle needs about two minutes on one core for this LP,
so the GPU test compares against this file instead of running it.

Run:  python tests/golden/make_long_rows_fixture.py
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "tests"), os.path.join(ROOT, "gpu-presolver_amd"), os.path.join(ROOT, "oracle")]

import numpy as np  # noqa: E402

import lp_synth  # noqa: E402
import oracle_ffi as O  # noqa: E402
from helpers import LP_KEYS, RECORD_KEYS, digests  # noqa: E402

SUMMARY_KEYS = ("reduced_rows", "reduced_cols", "reduced_nnz", "iterations", "has_infeasible", "has_unbounded")


def main():
    lp = lp_synth.config1(seed=1, row_len=(2, 64))
    runs = O.batched_closure_runs()
    t = time.time()
    s, rec, red = O.presolve(lp, O.params(enable_structure_specific_rules=0), True)
    secs = time.time() - t
    x, y, z = lp_synth.reduced_solution(6, red["m"], red["n"], red["l"], red["u"])
    post = {}
    for tag, original in (("original", lp), ("record", None)):
        xo, yo, zo = O.postsolve(rec, x, y, z, original)
        post[tag] = {name: [float(np.sum(v)), float(np.sum(v * v))] for name, v in zip("xyz", (xo, yo, zo))}
    out = {"lp": "lp_synth.config1(seed=1, row_len=(2, 64))", "nnz": int(len(lp["A_idx"])),
           "params": {"enable_structure_specific_rules": 0},
           "summary": {k: int(s[k]) for k in SUMMARY_KEYS},
           "obj_constant_delta": float(s["obj_constant_delta"]),
           "record_dims": [int(rec[k]) for k in ("m0", "n0", "m1", "n1")],
           "obj_constant_new": float(rec["obj_constant_new"]),
           "closure_runs": O.batched_closure_runs() - runs,
           "record": digests(rec, RECORD_KEYS), "reduced": digests(red, LP_KEYS),
           "reduced_solution_seed": 6, "postsolve_sum_sumsq": post, "oracle_seconds": round(secs, 1)}
    path = os.path.join(HERE, "long_rows_presolve.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {path} (oracle {secs:.0f} s, closure runs {out['closure_runs']})")


if __name__ == "__main__":
    main()
