"""The serving loop on the GPU (alis_amd/libalis.so): continuous batching of scattering
requests of different grid sizes in shared launches.

* Each request's spectrum served in the mix -- admitted while others run, sharing curl and
  post launches with them -- is bit for bit its spectrum served alone.
* In the reference's precision (FP32 everywhere) a request served in a mix on the GPU equals
  the same request alone on the CPU oracle bit for bit.
* In the default precision (FP32 fields, FP64 DFT) the served spectrum is within 1e-5
  (relative, each wavelength) of the FP64 oracle.
"""
import ctypes
import os

import numpy as np
import pytest

from scenarios import ROOT
from alis_amd import serve

pytestmark = pytest.mark.gpu

ORACLE = ROOT / "oracle" / "liboracle.so"


def _gpu():
    return ctypes.CDLL(str(ROOT / "alis_amd" / "libalis.so"))


def test_mix_equals_alone():
    os.environ["ALIS_PRECISION"] = "fp32"
    # three at once fill the slots; the fourth and fifth are admitted as slots free
    specs = [(0.0, 64, 600, 12, 9.0), (0.0, 96, 400, 24, 8.5), (0.0, 72, 500, 14, 8.0),
             (0.01, 64, 300, 10, 9.5), (0.02, 80, 200, 18, 8.8)]
    print(f"[gpu] serving {len(specs)} requests, 3 slots", flush=True)
    mix = serve.run(serve.make(specs), 64 << 30, 3, 50, _gpu())
    assert max(r.batchMax for r in mix) == 3
    for spec, m in zip(specs, mix):
        print(f"[gpu] request {spec} alone", flush=True)
        alone = serve.run(serve.make([spec]), 64 << 30, 3, 50, _gpu())[0]
        assert alone.batchMax == 1
        assert list(m.spectrum) == list(alone.spectrum), spec
        sp = np.array(m.spectrum)
        assert np.isfinite(sp).all() and np.any(sp != 0)   # partial runs: net flux of either sign


def test_reference_precision_equals_oracle():
    os.environ["ALIS_PRECISION"] = "ref"
    specs = [(0.0, 48, 240, 6, 9.0), (0.0, 40, 160, 2, 8.0), (0.0, 44, 200, 4, 8.5)]
    print("[gpu] serving 3 requests in the reference's precision", flush=True)
    mix = serve.run(serve.make(specs), 64 << 30, 3, 40, _gpu())
    assert max(r.batchMax for r in mix) == 3
    oracle = ctypes.CDLL(str(ORACLE))
    for spec, m in zip(specs, mix):
        alone = serve.run(serve.make([spec]), 64 << 30, 1, 40, oracle)[0]
        assert list(m.spectrum) == list(alone.spectrum), spec


def test_default_precision_spectrum_vs_fp64_oracle():
    specs = [(0.0, 48, 1500, 6, 9.0), (0.0, 56, 1500, 8, 8.2)]
    os.environ["ALIS_PRECISION"] = "fp32"
    print("[gpu] serving 2 requests, fp32 fields / fp64 DFT", flush=True)
    mix = serve.run(serve.make(specs), 64 << 30, 2, 100, _gpu())
    os.environ["ALIS_PRECISION"] = "fp64"
    oracle = ctypes.CDLL(str(ORACLE))
    for spec, m in zip(specs, mix):
        ref = np.array(serve.run(serve.make([spec]), 64 << 30, 1, 100, oracle)[0].spectrum, dtype=np.float64)
        got = np.array(m.spectrum, dtype=np.float64)
        err = np.abs(got - ref) / np.abs(ref)
        print(f"[gpu] {spec}: spectrum error vs fp64 oracle {err.max():.2e}")
        assert err.max() <= 1e-5
