"""Closed-form checks of the oracle that do not depend on the reference.

A vacuum box periodic in x, y, z holding one Fourier mode of Ez obeys the leapfrog
recurrence E^{n+1} + E^{n-1} = r E^n exactly, with r fixed by the discrete dispersion
relation sin^2(w dt/2)/dt^2 = sum_a sin^2(k_a d_a/2)/d_a^2 (coefficients as stored).  A
wrong curl sign, index offset, coefficient or ghost copy moves r by far more than the
tolerances below."""
import numpy as np
import pytest

from scenarios import dispersion

MODES = [(17, 3, 2), (17, 1, 0), (13, 4, 5), (21, 8, 1)]


@pytest.mark.parametrize("cells,m,p", MODES)
def test_dispersion_fp64(cells, m, p):
    r, expect = dispersion("oracle", cells, m, p, 120, fp64=True)
    assert len(r) > 20
    assert np.abs(r - expect).max() < 1e-11 * abs(expect) + 1e-12


@pytest.mark.parametrize("cells,m,p", MODES[:2])
def test_dispersion_fp32(cells, m, p):
    r, expect = dispersion("oracle", cells, m, p, 120, fp64=False)
    assert np.abs(r - expect).max() < 1e-4


def test_dispersion_detects_wrong_relation():
    """The check discriminates: the continuum relation (no sin) is rejected."""
    r, expect = dispersion("oracle", 17, 3, 2, 120, fp64=True)
    cont = 2 - (2 - expect) * 1.001
    assert np.abs(r - cont).max() > 1e-4
