"""The probe-spectrum sums behind writeSpectrum (alis_dev_spectrum_*: e^{i 2 pi tau (nTop - b)}
weighted sums of each probe's samples, FP64 with compensated summation) against closed
forms: a constant series sums to a geometric series in every bin, a cosine series to two of
them.  The CPU oracle's restatement is pinned this way on the CPU; the device kernel is held
to the same closed forms on the GPU (the GPU parity tests compare the two directly).  A
conjugated phase convention fails the check."""
import ctypes

import numpy as np
import pytest

from scenarios import ROOT

DELTA, NTOP, NBIN, SAMPLES, H0 = 0.0137, 40, 30, 3000, 31.5


def _abi(path):
    lib = ctypes.CDLL(str(path))
    lib.alis_dev_error.restype = ctypes.c_char_p
    lib.alis_dev_spectrum_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    lib.alis_dev_spectrum_add.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
    lib.alis_dev_spectrum_read.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
    lib.alis_dev_spectrum_release.argtypes = [ctypes.c_void_p]
    return lib


def _geometric(theta, n):
    """sum_{s=1..n} e^{i theta s}"""
    z = np.exp(1j * theta)
    return z * (1 - z ** n) / (1 - z)


def expected():
    b = np.arange(NBIN)
    theta = 2 * np.pi * DELTA * (NTOP - b)
    phi = 2 * np.pi * DELTA * H0
    const = _geometric(theta, SAMPLES)
    cosine = 0.5 * (_geometric(theta + phi, SAMPLES) + _geometric(theta - phi, SAMPLES))
    return np.stack([const, cosine])


def measured(lib):
    S = ctypes.c_void_p()
    assert lib.alis_dev_spectrum_create(2, NBIN, NTOP, ctypes.byref(S)) == 0, lib.alis_dev_error()
    v = (ctypes.c_double * 2)()
    for s in range(1, SAMPLES + 1):
        v[0], v[1] = 1.0, np.cos(2 * np.pi * DELTA * H0 * s)
        assert lib.alis_dev_spectrum_add(S, s * DELTA, v) == 0
    out = np.zeros(2 * 2 * NBIN)
    assert lib.alis_dev_spectrum_read(S, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
    lib.alis_dev_spectrum_release(S)
    return (out[0::2] + 1j * out[1::2]).reshape(2, NBIN)


def _check(got):
    want = expected()
    err = np.abs(got - want).max() / np.abs(want).max()
    # phases reach ~1e4 rad, so the rounding of tau = s * DELTA alone moves a term by ~1e-12
    assert err <= 1e-10, err
    # the conjugate convention e^{-i ...} is far off
    assert np.abs(got - np.conj(want)).max() / np.abs(want).max() > 1e-3


def test_oracle_spectrum_closed_form():
    _check(measured(_abi(ROOT / "oracle" / "liboracle.so")))


@pytest.mark.gpu
def test_device_spectrum_closed_form():
    print("[gpu] probe spectrum sums: closed forms, 3000 samples (12 staged launches)", flush=True)
    _check(measured(_abi(ROOT / "alis_amd" / "libalis.so")))
