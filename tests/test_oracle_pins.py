"""The oracle's routines against checks the mathematics fixes (tests/oracle_pins.py), on the
FP64 build: closed forms of the advection and pressure-gradient operators, the spectrum of
the implicit vertical acoustic operator against the continuum modes of an isothermal column,
the off-centred trapezoidal form of advance_w, hydrostatic equilibrium, the linearisation of
the equation of state, the mass budget of advance_mu_t, the coupling round trip of the
small step, constant-mixing-ratio preservation of the whole step, and rest over terrain.
tests/test_oracle_mutations.py shows that each of them rejects plausible mistakes."""
import pytest

import oracle_pins


@pytest.mark.parametrize("pin", oracle_pins.PINS, ids=lambda p: p.__name__)
def test_pin(oracle64, pin):
    pin(oracle64)
