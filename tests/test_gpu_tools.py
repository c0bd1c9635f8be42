"""Command-line tools on the GPU: the reference's CLI regression cases
(cpp/tests/test_tools.cmake:1-35, input files as in
cpp/tests/test_tool_common.cpp:73-110) run against gpu_presolver_mps,
gpu_presolver_export_reduced and gpu_presolver_check_postsolve_solution, and
an MPS model read by the library presolves exactly as the oracle presolves
the same host LP."""
import os
import subprocess

import numpy as np
import pytest

import gp_ffi
import oracle_ffi as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "gpu-presolver_amd", "bin")
CHECKER = os.path.join(BIN, "gpu_presolver_check_postsolve_solution")
EXPORTER = os.path.join(BIN, "gpu_presolver_export_reduced")
PRESOLVER = os.path.join(BIN, "gpu_presolver_mps")
TOOL_TEST = os.path.join(ROOT, "tests", "cpp", "build", "tool_common_test")

TOY = ("NAME toy\nROWS\n N OBJ\n G DEMAND\nCOLUMNS\n    X OBJ 1 DEMAND 1\n    Y OBJ 2 DEMAND 1\nRHS\n"
       "    RHS1 DEMAND 3\nBOUNDS\n UP BND1 X 2\n UP BND1 Y 4\nENDATA\n")
PRECISION = ("NAME precision\nROWS\n N OBJ\n G R\nCOLUMNS\n    X OBJ 1 R 1\n"
             "RHS\n    RHS1 R 0\nBOUNDS\n FX BND1 X 1.2345678901234567\nENDATA\n")
OVERFLOW = ("NAME overflow\nROWS\n N OBJ\n L R\nCOLUMNS\n    X OBJ 1 R 1e308\n"
            "RHS\n    RHS1 R 1\nBOUNDS\n FR BND1 X\nENDATA\n")
LARGE = ("NAME large\nROWS\n N OBJ\n L R\nCOLUMNS\n    X OBJ 1 R 1\n"
         "RHS\n    RHS1 R 1e200\nBOUNDS\n FR BND1 X\nENDATA\n")


def _solution(d, x, y=(0.0,)):
    os.makedirs(d, exist_ok=True)
    np.asarray(x, np.float64).tofile(os.path.join(d, "x_f64.bin"))
    np.asarray(y, np.float64).tofile(os.path.join(d, "y_f64.bin"))
    np.zeros(len(x)).tofile(os.path.join(d, "z_f64.bin"))


@pytest.fixture(scope="module")
def work(tmp_path_factory):
    w = tmp_path_factory.mktemp("tools")
    for name, text in (("toy.mps", TOY), ("precision.mps", PRECISION), ("activity-overflow.mps", OVERFLOW),
                       ("large-bounds.mps", LARGE)):
        (w / name).write_text(text)
    _solution(w / "valid", [2.0, 1.0])
    _solution(w / "infeasible", [0.0, 0.0])
    _solution(w / "nan", [np.nan, 1.0])
    _solution(w / "inf", [np.inf, 1.0])
    _solution(w / "long", [2.0, 1.0, 0.0])
    _solution(w / "short", [2.0])
    _solution(w / "empty", [], [])
    _solution(w / "overflow", [2.0])
    _solution(w / "large-infeasible", [2.0e200])
    _solution(w / "large-valid", [0.0])
    return w


def _run(args, max_iters):
    env = dict(os.environ, GPUPRESOLVER_MAX_ITERS=str(max_iters))
    return subprocess.run(args, capture_output=True, text=True, timeout=120, env=env)


@pytest.mark.parametrize("model,solution,code", [
    ("toy.mps", "valid", 0), ("toy.mps", "infeasible", 1),
    ("toy.mps", "nan", 2), ("toy.mps", "inf", 2), ("toy.mps", "long", 2), ("toy.mps", "short", 2),
    ("activity-overflow.mps", "overflow", 2),
    ("large-bounds.mps", "large-infeasible", 1), ("large-bounds.mps", "large-valid", 0),
])
def test_checker_exit_codes(work, model, solution, code):
    r = _run([CHECKER, str(work / model), str(work / solution)], 0)
    assert r.returncode == code, r.stdout + r.stderr
    if code == 0:
        keys = {line.split()[0] for line in r.stdout.splitlines()}
        assert {"primal_feas_l2_normalized", "record_hash", "x_org_hash", "y_org_hash", "z_org_hash"} <= keys


def test_export_keeps_full_precision(work):
    out = work / "precision-export"
    r = _run([EXPORTER, str(work / "precision.mps"), str(out)], 10)
    assert r.returncode == 0, r.stdout + r.stderr
    meta = dict(line.split(" ", 1) for line in (out / "meta.txt").read_text().splitlines())
    x = gp_ffi.MpsModel(PRECISION).lp["l"][0]
    assert float(meta["obj_constant"]) == x and float(meta["objective_shift"]) == x
    assert meta["format"] == "gpu_presolver_reduced_lp_v1" and meta["status"] == "ok"
    m, n = int(meta["reduced_rows"]), int(meta["reduced_cols"])
    assert np.fromfile(out / "A_rowPtr_i32.bin", np.int32).size == m + 1
    assert np.fromfile(out / "c_f64.bin", np.float64).size == n
    # everything presolves away: the empty reduced solution checks
    r = _run([CHECKER, str(work / "precision.mps"), str(work / "empty")], 10)
    assert r.returncode == 0, r.stdout + r.stderr


def test_export_reports_failed_writes(work):
    if not os.path.exists("/dev/full"):
        pytest.skip("no /dev/full")
    for name in ("c_f64.bin", "meta.txt"):
        d = work / f"full-{name}"
        d.mkdir()
        os.symlink("/dev/full", d / name)
        r = _run([EXPORTER, str(work / "toy.mps"), str(d)], 0)
        assert r.returncode == 2, r.stdout + r.stderr


def test_presolve_tool_summary(work):
    r = _run([PRESOLVER, str(work / "toy.mps")], 10)
    assert r.returncode == 0, r.stdout + r.stderr
    got = dict(line.split(" ", 1) for line in r.stdout.splitlines())
    assert got["model"] == "toy" and got["status"] == "ok" and got["original_rows"] == "1"
    assert _run([PRESOLVER, str(work / "missing.mps")], 10).returncode == 2


def test_tool_layer_owner_moves():
    r = subprocess.run([TOOL_TEST, "gpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "ok owner_moves" in r.stdout, r.stdout + r.stderr


def test_mps_model_presolves_like_the_oracle():
    """An MPS file of a synthetic LP, read and uploaded by the library,
    presolves to the oracle's record on the reader's host LP."""
    import lp_synth

    lp = lp_synth.generate(400, 600, seed=77, mean_row_len=5.0)
    lo_all, hi_all = lp["AL"], lp["AU"]
    kind = ["E" if lo_all[i] == hi_all[i] else ("L" if not np.isfinite(lo_all[i]) else "G") for i in range(lp["m"])]
    lines = ["NAME synth", "ROWS", " N OBJ"]
    lines += [f" {kind[i]} R{i}" for i in range(lp["m"])]
    lines.append("COLUMNS")
    for j in range(lp["n"]):
        lines.append(f"    C{j} OBJ {float(lp['c'][j])!r}")
        for p in range(lp["AT_ptr"][j], lp["AT_ptr"][j + 1]):
            lines.append(f"    C{j} R{lp['AT_idx'][p]} {float(lp['AT_val'][p])!r}")
    lines.append("RHS")
    lines += [f"    RHS R{i} {float(hi_all[i] if kind[i] == 'L' else lo_all[i])!r}" for i in range(lp["m"])]
    lines.append("RANGES")
    for i in range(lp["m"]):
        lo, hi = lp["AL"][i], lp["AU"][i]
        if lo != hi and np.isfinite(lo) and np.isfinite(hi):
            lines.append(f"    RNG R{i} {float(hi - lo)!r}")
    lines.append("BOUNDS")
    for j in range(lp["n"]):
        lines.append(f" LO BND C{j} {float(lp['l'][j])!r}")
        if np.isfinite(lp["u"][j]):
            lines.append(f" UP BND C{j} {float(lp['u'][j])!r}")
    lines.append("ENDATA")
    model = gp_ffi.MpsModel("\n".join(lines) + "\n")
    host = model.lp
    dev = model.upload()
    kw = dict(enable_structure_specific_rules=0)
    h = gp_ffi.Presolved(dev, gp_ffi.params(**kw), True)
    try:
        s_gpu, rec_gpu = h.summary(), h.record()
    finally:
        h.free()
    dev.free()
    s_cpu, rec_cpu, _ = O.presolve(host, O.params(**kw), True)
    for k in ("reduced_rows", "reduced_cols", "reduced_nnz", "iterations"):
        assert s_gpu[k] == s_cpu[k], k
    for k in ("row_red2org", "col_red2org", "fixed_idx", "fixed_val", "types", "indices", "vals"):
        assert np.array_equal(rec_gpu[k], rec_cpu[k]), k
    assert s_cpu["reduced_rows"] < host["m"]
