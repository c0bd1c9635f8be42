"""The CPU oracle against the known-answer cases of the reference's own
unit tests (tests/golden/reference_kats.json)."""
import pytest

import oracle_ffi
from golden_runner import load_cases, run_case

CASES = load_cases()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_matches_reference_kat(case):
    run_case(oracle_ffi, case)
