"""Pins the oracle's Krylov recurrences: exact two-iteration termination on
matrices whose Jacobi-scaled operator has two eigenvalues, and iterate-by-
iterate agreement with scipy's cg / bicgstab (checks in krylov_pins.py)."""

import pytest

import krylov_pins as kp


@pytest.mark.parametrize("pin", kp.ALL_PINS, ids=[p.__name__ for p in kp.ALL_PINS])
def test_krylov_pin(orc, pin):
    pin(orc)
