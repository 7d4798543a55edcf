"""The oracle's pointwise stages and their composition in the step against
closed forms (CPU).  The checks and what each pins are in tests/pins.py;
tests/test_oracle_mutations.py shows that each fails on a planted mistake."""
import pytest

import pins


@pytest.mark.parametrize("check", pins.ALL_CHECKS, ids=lambda c: c.__name__)
def test_pin(oracle_mod, check):
    check(oracle_mod)
