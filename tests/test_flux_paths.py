"""The net flux behind phaserPoyntingOut (src/phaser.c:228-245; phaserFluxD, in double
before its rounding to float) is the same whether the phaser's accumulators are on the
device (inner trapezoid sums from alis_dev_phaser_flux_sums: 64 lane sums per face row) or
were brought home by alisInvalidate (the same sums restated on the host in the same order,
alis_amd/host/phasers.c): equal to the bit, for a phaser whose x and y faces have rows of
181 cells, filled with field after 240 steps.  A host sum that adds each row in one sequence
gives other bits here (checked against the earlier host code).  Float accumulation, where
both paths read the same values (with FP64 accumulators the host copy holds floats)."""
import ctypes
import os

import numpy as np
import pytest

from scenarios import drive_lib


def _paths(kind):
    os.environ["ALIS_PRECISION"] = "ref"
    lib = drive_lib(kind)
    lib.alis_drive_flux_paths.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    out = np.zeros(2 * 16)
    n = lib.alis_drive_flux_paths(240, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 16)
    assert n == 5
    return out[0:2 * n:2], out[1:2 * n:2]


def test_flux_same_on_host_and_device_oracle():
    dev, host = _paths("oracle")
    assert np.all(np.isfinite(dev)) and np.all(dev != 0) and np.abs(dev).max() < 1e10 and np.array_equal(dev.view(np.int64), host.view(np.int64))


@pytest.mark.gpu
def test_flux_same_on_host_and_device_gpu():
    print("[gpu] flux with device and host accumulators", flush=True)
    dev, host = _paths("hip")
    assert np.all(np.isfinite(dev)) and np.all(dev != 0) and np.abs(dev).max() < 1e10 and np.array_equal(dev.view(np.int64), host.view(np.int64))
