"""The HIP library (through the C-ABI) against the known-answer cases of the
reference's own unit tests (tests/golden/reference_kats.json)."""
import pytest

import gp_ffi
from golden_runner import load_cases, run_case

pytestmark = pytest.mark.gpu
CASES = load_cases()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_gpu_matches_reference_kat(case):
    run_case(gp_ffi, case)
