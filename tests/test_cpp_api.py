"""The source-compatible C++ API as a C++ caller sees it: tests/cpp/
cpp_api_test.cpp is compiled against include/gpu_presolver/presolve/*.hpp
and linked to the in-tree library (built by __graft_entry__.build()).  On
the CPU the program must build and link; on the GPU each of its cases must
print "ok"."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
BIN = os.path.join(CPP, "build", "cpp_api_test")
CASES = ("summary_ownership", "presolve_postsolve_fixed", "presolve_postsolve_tiered",
         "scheduler_empties_lp_tiered", "scheduler_empties_lp_fixed", "rules_on_partial_plans",
         "has_good_nnz_progress", "input_errors")


def _built():
    if not os.path.exists(BIN):
        subprocess.run(["make", "-C", CPP], check=True, stdout=subprocess.DEVNULL)
    return BIN


def test_cpp_caller_builds_and_links():
    exe = _built()
    out = subprocess.run(["ldd", exe], check=True, capture_output=True, text=True).stdout
    assert "libgpu_presolver_amd.so" in out and "not found" not in out


@pytest.mark.gpu
def test_cpp_caller_runs():
    exe = _built()
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    lines = r.stdout.splitlines()
    assert r.returncode == 0, r.stdout + r.stderr
    for case in CASES:
        assert f"ok {case}" in lines, r.stdout
