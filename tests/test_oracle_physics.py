"""The oracle's physics pins (closed forms and invariants, tests/oracle_pins.py) hold for the
oracle as built."""
import pytest

from oracle_pins import PINS


@pytest.mark.parametrize("pin", PINS, ids=lambda f: f.__name__[len("pin_"):])
def test_pin(pin):
    pin(None)
