"""The flux sums entry point that allocates nothing (alis_dev_phaser_flux_sums_noalloc, for
callers that cannot allocate device memory) against the allocating one
(alis_dev_phaser_flux_sums): the same inputs give the same bits, on two calls in a row whose
second is smaller, so nothing of the first call's larger output survives in the shared
scratch array; requests beyond the array are refused with an error, before any device work.
"""
import ctypes

import numpy as np
import pytest

import alis_amd
from scenarios import drive_lib

pytestmark = pytest.mark.gpu

SCRATCH_DOUBLES = 1 << 19


def _abi():
    lib = alis_amd.library()
    dp, ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    for name in ("alis_dev_phaser_flux_sums", "alis_dev_phaser_flux_sums_noalloc"):
        f = getattr(lib, name)
        f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, dp, ctypes.c_long, dp, ctypes.c_long]
        f.restype = ctypes.c_int
    lib.alis_dev_phaser_size.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.alis_dev_phaser_size.restype = ctypes.c_long
    return lib


def _request(rng, ncell, shapes):
    """Terms over the first cells of the phaser faces: (kind, rows, row length) per term, with
    one unused output slot between consecutive terms (zero in both results)."""
    terms, woff, ooff, first = [], 0, 0, 0
    for kind, na, nb in shapes:
        assert first + na * nb <= ncell
        x, y = rng.integers(0, 4, 2)
        terms.append([kind, first, na, nb, int(x), int(y), woff, ooff])
        woff += nb if kind == 0 else na
        ooff += (na * 64 if kind == 0 else nb) + 1
        first += 3
    w = rng.standard_normal(woff)
    return np.array(terms, dtype=np.int32).ravel(), w, ooff


def _call(f, dev, pid, q, terms, w, nout):
    out = np.full(nout, np.nan)
    rc = f(dev, pid, q, len(terms) // 8, terms.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
           w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(w), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), nout)
    return rc, out


@pytest.mark.parametrize("precision", ["fp32", "ref"])
def test_noalloc_flux_sums_equal_allocating(precision):
    import os
    os.environ["ALIS_PRECISION"] = precision
    lib = _abi()
    drive = drive_lib("hip")
    drive.alis_drive_hold.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p),
                                      ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    dev, pid, nf = ctypes.c_void_p(), ctypes.c_int(), ctypes.c_int()
    print(f"[gpu] no-alloc flux sums: nanoblock 30 steps, {precision}", flush=True)
    assert drive.alis_drive_hold(b"nanoblock", 30, ctypes.byref(dev), ctypes.byref(pid), ctypes.byref(nf)) == 0
    try:
        ncell = lib.alis_dev_phaser_size(dev, pid.value) // (nf.value * 8)
        rng = np.random.default_rng(7)
        big = _request(rng, ncell, [(0, 20, 90), (1, 30, 40), (0, 5, 150), (1, 55, 20), (0, 3, 64)])
        small = _request(rng, ncell, [(1, 12, 33), (0, 5, 17)])
        for label, (terms, w, nout), q in (("large", big, 0), ("smaller", small, nf.value - 1)):
            print(f"[gpu] {label} request: {len(terms) // 8} terms, {nout} outputs, frequency {q}", flush=True)
            rc_a, a = _call(lib.alis_dev_phaser_flux_sums, dev, pid.value, q, terms, w, nout)
            rc_b, b = _call(lib.alis_dev_phaser_flux_sums_noalloc, dev, pid.value, q, terms, w, nout)
            assert rc_a == 0 and rc_b == 0
            assert np.isfinite(a).all() and np.abs(a).max() > 0
            assert np.array_equal(a.view(np.int64), b.view(np.int64)), label
        # outputs between the terms' ranges are zero in both
        terms, w, nout = small
        assert b[33] == 0.0 and a[33] == 0.0
        # a request beyond the scratch array is refused before any device work
        t = np.array([0, 0, 1, 1, 0, 0, 0, 0], dtype=np.int32)
        rc, _ = _call(lib.alis_dev_phaser_flux_sums_noalloc, dev, pid.value, 0, t, np.ones(1), SCRATCH_DOUBLES)
        assert rc == -1
        assert b"scratch" in lib.alis_dev_error()
    finally:
        drive.alis_drive_release()
