"""The oracle against its pins (CPU), and every planted mistake caught by them."""
import pytest

import oracle
from oracle_pins import ALL_PINS, MUTANTS


@pytest.mark.parametrize("pin", ALL_PINS, ids=lambda f: f.__name__)
def test_pin(orc, pin):
    pin(orc)


@pytest.mark.parametrize("mutant", sorted(MUTANTS), ids=lambda m: f"mutant{m}")
def test_mutant_caught(mutant):
    m = oracle.Oracle(oracle.build(mutant))
    failed = []
    for pin in ALL_PINS:
        try:
            pin(m)
        except AssertionError:
            failed.append(pin.__name__)
    assert failed, f"mutant {mutant} ({MUTANTS[mutant]}) passes every pin"
