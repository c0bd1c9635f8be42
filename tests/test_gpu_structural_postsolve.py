"""Postsolve of records with the structure-specific reductions' primal
recovery steps (L1 splits, outer pairs, linked and max slacks, antipodal
components) interleaved with the tape by checkpoints: the C++ program
tests/cpp/structural_postsolve_test runs the reference's own cases
(cpp/tests/test_gpu_postsolve.cpp:868-1058) and hand-worked cases of each
recovery kind through postsolve_gpu and
postsolve_restore_structural_primal_gpu."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "build", "structural_postsolve_test")
CASES = ("dual_replay_after_structural_recovery", "tape_checkpoint_after_structural_primal",
         "timeline_in_exact_reverse_order", "rejects_invalid_timelines", "structural_kinds_in_order",
         "antipodal_component")


def _built():
    if not os.path.exists(EXE):
        subprocess.run(["make", "-C", os.path.dirname(os.path.dirname(EXE))], check=True, stdout=subprocess.DEVNULL)
    return EXE


def test_structural_postsolve_program_builds():
    out = subprocess.run(["ldd", _built()], check=True, capture_output=True, text=True).stdout
    assert "libgpu_presolver_amd.so" in out and "not found" not in out


@pytest.mark.gpu
def test_structural_postsolve_cases():
    r = subprocess.run([_built()], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    for case in CASES:
        assert f"ok {case}" in r.stdout.splitlines(), r.stdout
