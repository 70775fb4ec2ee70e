"""FP64 phaser accumulators add with Neumaier compensation (alis_dev_phaser_update; the sums
are folded before every read).  A fresh one-frequency phaser on the z = 2 plane of a stepped
world takes one sample with factor 1 (so it holds the face fields e exactly), then 20000
samples with factor 1e-16, each of which a plain double sum rounds away or up by a sizeable
fraction of itself.  The result must be the correctly rounded sum of the terms (math.fsum)
to within 2 ulp; the plain sum, simulated here, misses it by more than 1e-13 relative, so the
check tells the two apart.  Run on the oracle (CPU) and on the engine (GPU), FP32 and FP64
fields."""
import ctypes
import math
import os

import numpy as np
import pytest

from scenarios import drive_lib

N, TINY = 20000, 1e-16
BOX = (-5, 4, -5, 4, 2, 2)          # i, j ranges and the plane k


class PhaserDesc(ctypes.Structure):
    _fields_ = [("nf", ctypes.c_int), ("iMin", ctypes.c_int), ("iMax", ctypes.c_int), ("jMin", ctypes.c_int),
                ("jMax", ctypes.c_int), ("kMin", ctypes.c_int), ("kMax", ctypes.c_int),
                ("cf", ctypes.POINTER(ctypes.POINTER(ctypes.c_float)))]


def _stream(kind, precision):
    os.environ["ALIS_PRECISION"] = precision
    lib = drive_lib(kind)
    lib.alis_drive_hold.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p),
                                    ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    lib.alis_dev_error.restype = ctypes.c_char_p
    dp = ctypes.POINTER(ctypes.c_double)
    lib.alis_dev_phaser_add.argtypes = [ctypes.c_void_p, ctypes.POINTER(PhaserDesc), ctypes.POINTER(ctypes.c_int)]
    lib.alis_dev_phaser_update.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp]
    lib.alis_dev_phaser_read.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, ctypes.c_long]
    lib.alis_dev_phaser_size.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.alis_dev_phaser_size.restype = ctypes.c_long
    lib.alis_dev_phaser_remove.argtypes = [ctypes.c_void_p, ctypes.c_int]
    dev, pid, nf = ctypes.c_void_p(), ctypes.c_int(), ctypes.c_int()
    assert lib.alis_drive_hold(b"nanoblock", 200, ctypes.byref(dev), ctypes.byref(pid), ctypes.byref(nf)) == 0
    try:
        ncell = (BOX[1] - BOX[0] + 1) * (BOX[3] - BOX[2] + 1)
        zero = [(ctypes.c_float * (2 * ncell))() for _ in range(4)]
        cf = (ctypes.POINTER(ctypes.c_float) * 24)()
        for c in range(4):
            cf[4 * 4 + c] = ctypes.cast(zero[c], ctypes.POINTER(ctypes.c_float))   # the zMin face only
        desc = PhaserDesc(1, *BOX[:4], BOX[4], BOX[5], cf)
        q = ctypes.c_int()
        assert lib.alis_dev_phaser_add(dev, ctypes.byref(desc), ctypes.byref(q)) == 0, lib.alis_dev_error()
        n = lib.alis_dev_phaser_size(dev, q.value)
        assert n == 4 * ncell * 2
        out = np.zeros(n)
        h0 = (ctypes.c_double * 2)(0.0, 0.0)

        def update(f):
            e = (ctypes.c_double * 2)(f, 0.0)
            assert lib.alis_dev_phaser_update(dev, q.value, 0, 1, e, h0) == 0, lib.alis_dev_error()

        def read():
            assert lib.alis_dev_phaser_read(dev, q.value, out.ctypes.data_as(dp), n) == 0, lib.alis_dev_error()
            return out.copy()

        update(1.0)
        first = read()
        for _ in range(N):
            update(TINY)
        last = read()
        lib.alis_dev_phaser_remove(dev, q.value)
        return first, last
    finally:
        lib.alis_drive_release()


def _check(first, last):
    e = first[0:4 * 2 * 100:2][:200]           # real parts of E1 and E2 (the first two components)
    got = last[0:4 * 2 * 100:2][:200]
    assert np.count_nonzero(e) > 100
    y = np.float64(TINY) * e                   # the device's term, the same IEEE product
    want = np.array([math.fsum([a] + [b] * N) for a, b in zip(e, y)])
    plain = e.copy()
    for _ in range(N):
        plain = plain + y
    scale = np.abs(want)
    nz = scale > 0
    err = (np.abs(got - want)[nz] / scale[nz]).max()
    plain_err = (np.abs(plain - want)[nz] / scale[nz]).max()
    print(f"compensated {err:.2e}, plain {plain_err:.2e}")
    assert err <= 4.5e-16 and plain_err > 1e-13
    # the imaginary parts and the H components took only zero terms
    assert np.array_equal(last[1::2], first[1::2]) and np.array_equal(last[400:], first[400:])


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_oracle_compensated_phaser(precision):
    _check(*_stream("oracle", precision))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_device_compensated_phaser(precision):
    print(f"[gpu] compensated phaser sums, {precision}: {N} samples", flush=True)
    _check(*_stream("hip", precision))
