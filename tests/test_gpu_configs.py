"""Benchmark configurations at their full sizes on the GPU, against the FP64 oracle and the
reference's own C code (the workloads of alis_amd/workloads.c, built over the engine and,
as oracle/libwl_oracle.so, over the CPU oracle).

* The 64^3 FP64 dipole as configured (10-cell CPML, Ez point dipole, 8 probes reading Ez
  and Hy every step, 1000 steps): every probe sample within 1e-12 of the FP64 oracle,
  relative to the largest sample.
* The 64-world k-point sweep (65x65x50 Bloch worlds, 128 entries per batched launch, so the
  entry lookup takes its second 64-lane trip; packed-row launches with 1/eps): in FP64 after
  24 steps every world's E and H (both parts) and every world's 256-bin probe spectrum
  (8 probes) within 1e-12 of the FP64 oracle; with FP32 fields after 300 steps the sum of
  squares of every world's state plus every probe value equal to that of the reference's
  own C code (oracle/_ref/wl_ref) to the last bit.
"""
import ctypes
import subprocess

import numpy as np
import pytest

from scenarios import ROOT

pytestmark = pytest.mark.gpu

REF_WL = ROOT / "oracle" / "_ref" / "wl_ref"


def _lib(path):
    w = ctypes.CDLL(str(path))
    dp = ctypes.POINTER(ctypes.c_double)
    w.wl_create.restype = ctypes.c_void_p
    w.wl_create.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    w.wl_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
    w.wl_checksum.restype = ctypes.c_double
    w.wl_checksum.argtypes = [ctypes.c_void_p]
    w.wl_trace.restype = ctypes.c_long
    w.wl_trace.argtypes = [ctypes.c_void_p, dp, ctypes.c_long]
    w.wl_spectrum.restype = ctypes.c_long
    w.wl_spectrum.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, ctypes.c_long]
    w.wl_field.restype = ctypes.c_long
    w.wl_field.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, ctypes.c_long]
    w.wl_free.argtypes = [ctypes.c_void_p]
    return w


GPU = ROOT / "alis_amd" / "libalis_workloads.so"
ORACLE = ROOT / "oracle" / "libwl_oracle.so"


def _trace(w, c):
    n = w.wl_trace(c, None, 0)
    out = np.zeros(n)
    w.wl_trace(c, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n)
    return out


def _rel(a, b):
    scale = np.abs(b).max()
    return np.abs(a - b).max() / scale if scale else np.abs(a - b).max()


def _dipole_trace(path, steps):
    w = _lib(path)
    c = w.wl_create(b"dipole", 64, 64, 64)
    w.wl_step(c, steps)
    t = _trace(w, c)
    w.wl_free(c)
    return t


def test_dipole_64_fp64_probe_traces():
    print("[gpu] dipole 64^3 fp64, 1000 steps, 8 probes: GPU", flush=True)
    g = _dipole_trace(GPU, 1000)
    print("[gpu] dipole 64^3 fp64: oracle", flush=True)
    o = _dipole_trace(ORACLE, 1000)
    assert g.shape == o.shape == (1000 * 16,)
    err = _rel(g, o)
    print(f"[gpu] dipole probe traces: max error {err:.2e} of the largest sample {np.abs(o).max():.3e}")
    assert np.abs(o).max() > 0 and err <= 1e-12


def _kpoint_state(path, steps):
    """Every world's fields (E then H, real and imaginary parts) and probe spectra."""
    w = _lib(path)
    c = w.wl_create(b"kpoints", 64, 64, 64)
    w.wl_step(c, steps)
    dp = ctypes.POINTER(ctypes.c_double)
    fields, spectra = [], []
    per = 67 * 67 * 52
    buf = np.zeros(per)
    sbuf = np.zeros(2 * 9 * 256)
    for q in range(64):
        f = []
        for h in (0, 1):
            for m in (0, 1):
                for a in (0, 1, 2):
                    assert w.wl_field(c, q, h, m, a, buf.ctypes.data_as(dp), per) == per
                    f.append(buf.copy())
        fields.append(np.concatenate(f))
        assert w.wl_spectrum(c, q, sbuf.ctypes.data_as(dp), sbuf.size) == sbuf.size
        spectra.append(sbuf.copy())
    w.wl_free(c)
    return fields, spectra


def test_kpoints_64_worlds_fp64_vs_oracle():
    steps = 24
    print(f"[gpu] 64 k-point worlds fp64, {steps} steps: GPU", flush=True)
    gf, gs = _kpoint_state(GPU, steps)
    print("[gpu] 64 k-point worlds fp64: oracle", flush=True)
    of, os_ = _kpoint_state(ORACLE, steps)
    worst_f = max(_rel(a, b) for a, b in zip(gf, of))
    worst_s = max(_rel(a[512:], b[512:]) for a, b in zip(gs, os_))     # probe rows (row 0: the source)
    print(f"[gpu] k-point worlds: fields {worst_f:.2e}, spectra {worst_s:.2e}")
    assert all(np.abs(b).max() > 0 for b in of)
    assert all(np.abs(b[512:]).max() > 0 for b in os_)
    assert worst_f <= 1e-12 and worst_s <= 1e-12


@pytest.mark.skipif(not REF_WL.exists(), reason="reference build (oracle/_ref) not present")
def test_kpoints_64_worlds_full_size_matches_reference():
    print("[gpu] 64 k-point worlds, 300 steps: the reference's C code", flush=True)
    ref = subprocess.run([str(REF_WL), "kpoints", "64", "2", "298"], capture_output=True, text=True,
                         timeout=600, check=True).stdout.split()
    print("[gpu] 64 k-point worlds, 300 steps: GPU (FP32 fields)", flush=True)
    w = _lib(GPU)
    c = w.wl_create(b"kpoints", 64, 32, 64)
    w.wl_step(c, 300)
    gpu = w.wl_checksum(c)
    w.wl_free(c)
    print(f"[gpu] checksum gpu {gpu!r} reference {ref[3]}")
    assert float(ref[3]) > 0
    assert gpu == float(ref[3])
