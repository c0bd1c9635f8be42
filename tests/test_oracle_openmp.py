"""The OpenMP build of the oracle (oracle/build/liboracle_omp.so, the CPU
baseline's all-core leg) gives bit for bit the serial build's record and
reduced LP: both run in subprocesses on the same seeded LP."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "oracle", "build")

CODE = r'''
import hashlib, json, sys
import numpy as np
sys.path[:0] = [sys.argv[1] + "/gpu-presolver_amd", sys.argv[1] + "/oracle"]
import lp_synth, oracle_ffi
lp = lp_synth.generate(60000, 90000, seed=5, mean_row_len=6.0)
s, rec, red = oracle_ffi.presolve(lp, oracle_ffi.params(enable_structure_specific_rules=0), True)
h = hashlib.sha256()
for d in (rec, red):
    for k in sorted(d):
        h.update(k.encode()); h.update(np.asarray(d[k]).tobytes())
print(json.dumps({"hash": h.hexdigest(), "rows": s["reduced_rows"]}))
'''


def _run(lib, threads):
    env = dict(os.environ, ORACLE_LIB=os.path.join(BUILD, lib), OMP_NUM_THREADS=str(threads))
    r = subprocess.run([sys.executable, "-c", CODE, ROOT], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_openmp_build_matches_serial_build():
    if not os.path.exists(os.path.join(BUILD, "liboracle_omp.so")):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, stdout=subprocess.DEVNULL)
    serial = _run("liboracle.so", 1)
    omp = _run("liboracle_omp.so", 4)
    assert serial == omp and serial["rows"] > 0
