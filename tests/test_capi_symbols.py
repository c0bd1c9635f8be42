"""The built library loads without a GPU and exports every entry point the
public headers declare: the C-ABI (include/gp_presolve_capi.h) by name and
the source-compatible C++ API by demangled signature."""
import ctypes as C
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "gpu-presolver_amd", "lib", "libgpu_presolver_amd.so")
CAPI_H = os.path.join(ROOT, "include", "gp_presolve_capi.h")


@pytest.fixture(scope="module")
def built_lib():
    if not os.path.exists(LIB):
        subprocess.run(["make", "-C", os.path.join(ROOT, "gpu-presolver_amd"), "-j", "8"], check=True,
                       stdout=subprocess.DEVNULL)
    return LIB


def capi_functions():
    text = open(CAPI_H).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gp_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = capi_functions()
    for required in ("gp_presolve", "gp_postsolve", "gp_rule_apply", "gp_row_activity", "gp_col_max_abs",
                     "gp_compact_csr", "gp_transpose_csr", "gp_presolve_free", "gp_last_error"):
        assert required in names


def test_library_loads_and_exports_capi(built_lib):
    lib = C.CDLL(built_lib)
    missing = [name for name in capi_functions() if not hasattr(lib, name)]
    assert not missing, f"undefined C-ABI symbols: {missing}"


def test_params_default_matches_reference_defaults(built_lib):
    import gp_types as T

    lib = C.CDLL(built_lib)
    p = T.GpParams()
    lib.gp_params_default(C.byref(p))
    assert p.max_iters == 10 and p.record_postsolve_tape == 1 and p.record_postsolve_tape_cpu == 0
    assert p.feasibility_tol == 1e-6 and p.bound_tol == 1e-6 and p.zero_tol == 1e-10
    assert p.doubleton_equations_min_selected_per_batch == 256
    assert p.doubleton_equations_min_selected_ratio == 0.005
    assert p.dual_fix_max_nonzero_cost_col_degree == 128
    assert p.duplicate_columns_max_abs_ratio == 2.0
    assert p.duplicate_columns_scale_guard_min_groups == 128
    assert p.use_tiered_scheduler == 1 and p.enable_redundant_bounds == 0
    assert p.implied_variable_bounds_bound_only_nnz_round_budget == 8000000


CXX_API = [
    "gpu_presolver::presolve::run_gpu_presolve_with_reduced_lp(gpu_presolver::presolve::LPInfoGpu const&, "
    "gpu_presolver::presolve::PresolveParams const&)",
    "gpu_presolver::presolve::run_gpu_presolve_with_record(gpu_presolver::presolve::LPInfoGpu const&, "
    "gpu_presolver::presolve::PresolveParams const&)",
    "gpu_presolver::presolve::free_gpu_presolve_reduced_lp(gpu_presolver::presolve::GpuPresolveSummary&)",
    "gpu_presolver::presolve::_has_good_nnz_progress(int, int, double)",
    "gpu_presolver::presolve::postsolve_gpu(double*, double*, double*, "
    "gpu_presolver::presolve::PresolveRecordGpu const&, double const*)",
    "gpu_presolver::presolve::postsolve_gpu(double*, double*, double*, "
    "gpu_presolver::presolve::PresolveRecordGpu const&, gpu_presolver::presolve::LPInfoGpu const*, double const*)",
    "gpu_presolver::presolve::compute_row_activity_summary(",
    "gpu_presolver::presolve::compute_col_max_abs(",
    "gpu_presolver::presolve::append_compact_fixed_col_tape_from_device(",
    "gpu_presolver::presolve::apply_rule_empty_rows(",
    "gpu_presolver::presolve::apply_rule_singleton_rows(",
    "gpu_presolver::presolve::apply_rule_infeasible_redundant_rows(",
    "gpu_presolver::presolve::apply_rule_implied_variable_bounds(",
    "gpu_presolver::presolve::apply_rule_duplicate_rows(",
    "gpu_presolver::presolve::apply_rule_infeasible_fixed_variables(",
    "gpu_presolver::presolve::apply_rule_empty_cols(",
    "gpu_presolver::presolve::apply_rule_column_singletons_eq(",
    "gpu_presolver::presolve::apply_rule_column_singletons_dual_infer(",
    "gpu_presolver::presolve::apply_rule_doubleton_equations(",
    "gpu_presolver::presolve::apply_rule_dual_fix(",
    "gpu_presolver::presolve::apply_rule_duplicate_columns(",
    "gpu_presolver::presolve::apply_rule_redundant_bounds(",
    "gpu_presolver::presolve::release_column_singletons_workspace()",
    "gpu_presolver::presolve::release_implied_variable_bounds_workspace()",
    "gpu_presolver::presolve::postsolve_restore_structural_primal_gpu(",
    "gpu_presolver::presolve::GpuPresolveSummary::GpuPresolveSummary(gpu_presolver::presolve::GpuPresolveSummary&&)",
    "gpu_presolver::presolve::PostsolveTapeGpu::~PostsolveTapeGpu()",
    "gpu_presolver::presolve::PostsolveTapeGpu::reset()",
]


@pytest.mark.skipif(shutil.which("nm") is None, reason="nm not available")
def test_library_exports_cxx_api(built_lib):
    out = subprocess.run(["nm", "-D", "-C", "--defined-only", built_lib], check=True, capture_output=True,
                         text=True).stdout
    missing = [sig for sig in CXX_API if sig not in out]
    assert not missing, f"C++ API symbols not exported: {missing}"


def test_batched_closure_threshold_default(built_lib):
    """kMinNnzForBatchedClosure (reference gpu_presolve.cu:2506); settable
    without a GPU."""
    lib = C.CDLL(built_lib)
    lib.gp_batched_closure_min_nnz.restype = C.c_int32
    lib.gp_set_batched_closure_min_nnz.argtypes = [C.c_int32]
    assert lib.gp_batched_closure_min_nnz() == 10_000_000
    lib.gp_set_batched_closure_min_nnz(5)
    assert lib.gp_batched_closure_min_nnz() == 5
    lib.gp_set_batched_closure_min_nnz(10_000_000)
