"""Host logic of the serving loop (alis_amd/host/serve.c) on the CPU: the loop and the
request generator run in the host library over the CPU oracle engine (oracle/liboracle.so,
the same host code as libalis.so over a plain-C engine), on grids small enough to step in
a second.

* Continuous batching keeps each result: requests served together equal each served alone.
* Admission honours the slot count and the device-memory budget.
* Invalid request lists are refused.
* The seeded load: deterministic, within the stated ranges, Poisson inter-arrival times.
"""
import ctypes
import os

import numpy as np
import pytest

from scenarios import ROOT
from alis_amd import serve

ORACLE = ROOT / "oracle" / "liboracle.so"


@pytest.fixture(scope="module")
def lib():
    if not ORACLE.exists():
        pytest.skip("oracle/liboracle.so is not built")
    return ctypes.CDLL(str(ORACLE))


# (arrival s, grid edge, steps, radius cells, plasma eV)
SMALL = [(0.0, 40, 40, 2, 9.0), (0.0, 44, 24, 3, 8.5), (0.0, 42, 32, 2, 8.0), (0.0, 40, 16, 1, 9.5)]


def _spectra(res):
    return [list(r.spectrum) for r in res]


def test_mix_equals_alone(lib):
    os.environ["ALIS_PRECISION"] = "fp32"
    mix = serve.run(serve.make(SMALL), 1e12, 3, 8, lib)
    assert max(r.batchMax for r in mix) == 3
    for spec, m in zip(SMALL, mix):
        alone = serve.run(serve.make([spec]), 1e12, 3, 8, lib)[0]
        assert alone.batchMax == 1
        assert list(m.spectrum) == list(alone.spectrum)
        assert list(m.wavelength) == list(alone.wavelength)
        assert m.cellSteps == spec[1] ** 3 * spec[2]
    s = np.array(_spectra(mix))
    assert np.isfinite(s).all() and (s[0] > 0).all()
    wl = np.array(mix[0].wavelength)
    assert np.all(np.diff(wl) > 0) and abs(wl[0] - 385) < 1e-3 and abs(wl[-1] - 910) < 1e-3


def test_slots_and_budget(lib):
    os.environ["ALIS_PRECISION"] = "fp32"
    one = serve.run(serve.make(SMALL), 1e12, 1, 8, lib)
    assert all(r.batchMax == 1 for r in one)
    # requests finish in arrival order when they run one at a time
    fin = [r.finished for r in one]
    assert fin == sorted(fin)
    # a budget below two requests' bytes admits one at a time, whatever the slots
    tight = serve.run(serve.make(SMALL), 1.5 * min(r.bytes for r in one), 4, 8, lib)
    assert all(r.batchMax == 1 for r in tight)
    assert _spectra(tight) == _spectra(one)


def test_late_arrivals_join_a_running_batch(lib):
    os.environ["ALIS_PRECISION"] = "fp32"
    # the second request arrives while the first one (many ticks) is running
    specs = [(0.0, 40, 400, 2, 9.0), (0.05, 40, 16, 2, 8.0)]
    res = serve.run(serve.make(specs), 1e12, 4, 4, lib)
    assert res[1].admitted >= 0.05 and res[1].admitted < res[0].finished
    assert res[0].batchMax == 2 and res[1].batchMax == 2
    assert res[1].finished < res[0].finished
    alone = serve.run(serve.make([specs[1]]), 1e12, 4, 4, lib)[0]
    assert list(alone.spectrum) == list(res[1].spectrum)


def test_invalid_lists_are_refused(lib):
    with pytest.raises(ValueError):
        serve.run(serve.make([(0.1, 40, 8, 2, 9.0), (0.0, 40, 8, 2, 9.0)]), 1e12, 2, 4, lib)  # out of order
    with pytest.raises(ValueError):
        serve.run(serve.make([(0.0, 40, 8, 9, 9.0)]), 1e12, 2, 4, lib)                       # sphere too big
    with pytest.raises(ValueError):
        serve.run(serve.make([(0.0, 40, 0, 2, 9.0)]), 1e12, 2, 4, lib)                       # no steps
    with pytest.raises(ValueError):
        serve.run(serve.make([(0.0, 40, 8, 2, 9.0)]), 1e12, 0, 4, lib)                       # no slots
    assert len(serve.run(serve.make([]), 1e12, 2, 4, lib)) == 0


def test_max_radius(lib):
    serve._bind(lib)
    # grid edge n: domain of n - 18 nodes; sphere + 8 cells within half of it
    assert [lib.alisServeMaxRadius(n) for n in (64, 96, 128, 192)] == [14, 30, 46, 78]


def test_request_generator(lib):
    a = serve.requests(2024, 4000, 20.0, lib)
    b = serve.requests(2024, 4000, 20.0, lib)
    c = serve.requests(2025, 4000, 20.0, lib)
    key = lambda R: [(r.arrival, r.cells, r.steps, r.radius, r.plasma) for r in R]
    assert key(a) == key(b) and key(a) != key(c)
    t = np.array([r.arrival for r in a])
    gaps = np.diff(np.concatenate([[0], t]))
    assert (gaps > 0).all()
    # exponential gaps of mean 1/20 s: mean and coefficient of variation (both 1/20 and 1)
    assert abs(gaps.mean() * 20 - 1) < 0.05
    assert abs(gaps.std() / gaps.mean() - 1) < 0.06
    cells = np.array([r.cells for r in a])
    assert set(cells) == {64, 96, 128, 192}
    assert all(abs((cells == n).mean() - 0.25) < 0.03 for n in (64, 96, 128, 192))
    rad = np.array([r.radius for r in a])
    assert rad.min() == 10 and rad.max() == 40
    assert all(10 <= r.radius <= min(40, lib.alisServeMaxRadius(r.cells)) for r in a)
    steps = np.array([r.steps for r in a])
    assert steps.min() >= 1000 and steps.max() <= 6000 and steps.min() < 1100 and steps.max() > 5900
    plasma = np.array([r.plasma for r in a])
    assert plasma.min() >= 8 and plasma.max() <= 9.5
