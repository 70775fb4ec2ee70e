"""Near-to-far projection on the GPU (alis_dev_farfield: the FP64 MFMA complex GEMM k_ff_gemm over the
phaser faces plus the phase/profile reductions) against the CPU oracle's surface sums.

The far field of every frequency of a scenario's phaser array, in double precision, at
directions that cover normal and oblique incidence, the averaged 90-degree band, the lower
half space and azimuths that rotate the polarisation components:
* FP64 worlds: within 1e-12 of the FP64 oracle (relative to the largest value of the
  frequency);
* default precision (FP32 fields, FP64 DFT and far field): within 1e-5 of the FP64 oracle.
nanoblock has all six PML faces; symband a z symmetry plane with Bloch x/y; oblique an
oblique plane wave in a periodic cell (z faces only, array factor); pillars a periodic
pillar array under normal incidence (z faces only, no array factor).
"""
import numpy as np
import pytest

from scenarios import SCENARIOS, farfield

pytestmark = pytest.mark.gpu

CASES = ("nanoblock", "symband", "oblique", "pillars")


def _rel(a, b):
    scale = np.abs(b).max(axis=(1, 2), keepdims=True)
    scale[scale == 0] = 1
    return (np.abs(a - b) / scale).max()


@pytest.mark.parametrize("name", CASES)
def test_fp64_farfield_matches_oracle(name):
    print(f"[gpu] {name} far field fp64 vs oracle", flush=True)
    g = farfield("hip", name, SCENARIOS[name], "fp64")
    o = farfield("oracle", name, SCENARIOS[name], "fp64")
    assert g.shape == o.shape and np.abs(o).max() > 0
    err = _rel(g, o)
    print(f"[gpu] {name}: far field error {err:.2e}")
    assert err <= 1e-12


@pytest.mark.parametrize("name", CASES)
def test_default_precision_farfield(name):
    print(f"[gpu] {name} far field fp32 fields vs fp64 oracle", flush=True)
    g = farfield("hip", name, SCENARIOS[name], "fp32")
    o = farfield("oracle", name, SCENARIOS[name], "fp64")
    err = _rel(g, o)
    print(f"[gpu] {name}: far field error {err:.2e}")
    assert err <= 1e-5
