"""Fused steps: k_step runs both curl halves of a step in one pass over the grid from one set
of arrays into the other, then the E half's end-of-step program runs on the new set and
k_fix recomputes H where that program changed E.  The results must be those of the separate
half steps bit for bit.

Fused steps are opt-in.  ALIS_FUSE=force fuses every world whose additive operations lie in
the update range however small it is, so the scenarios of the parity fixtures (TF/SF boxes,
dipoles, Drude-Lorentz objects, CPML, periodic, Bloch and hexagonal boundaries, phasers) run
through k_step here; ALIS_FUSE=1 fuses the worlds large enough to fill the GPU.
"""
import ctypes

import numpy as np
import pytest

from scenarios import ROOT, SCENARIOS, compare, dispersion, golden, run_raw, run_scenario

pytestmark = pytest.mark.gpu

# scenarios that step one world at a time (the k-point sweep steps in batches, never fused)
SINGLE = sorted(n for n in SCENARIOS if n != "kpoints")


@pytest.fixture
def force(monkeypatch):
    monkeypatch.setenv("ALIS_FUSE", "force")


@pytest.mark.parametrize("name", SINGLE)
def test_fused_reference_precision_bitwise(name, force):
    print(f"[gpu] {name} fused steps, float accumulation", flush=True)
    got = run_scenario("hip", name, SCENARIOS[name], "ref")
    bad = compare(golden(name), got)
    assert not bad, "\n".join(bad[:20])


@pytest.mark.parametrize("name", SINGLE)
def test_fused_fp64_matches_fp64_oracle(name, force):
    print(f"[gpu] {name} fused steps, fp64 vs oracle", flush=True)
    fg, pg = run_raw("hip", name, SCENARIOS[name], "fp64")
    fo, po = run_raw("oracle", name, SCENARIOS[name], "fp64")
    assert fg.shape == fo.shape and pg.shape == po.shape
    scale = np.abs(fo).max()
    assert np.abs(fg - fo).max() <= 1e-12 * scale
    if po.size:
        assert np.abs(pg - po).max() <= 1e-12 * np.abs(po).max()


@pytest.mark.parametrize("cells,m,p", [(17, 3, 2)])
def test_fused_reads_between_halves(cells, m, p, monkeypatch):
    """A read of E after the E half step (alis_dev_read_field) completes the deferred half the
    separate way: the recurrence of a periodic box is the same with and without fusion."""
    monkeypatch.delenv("ALIS_FUSE", raising=False)
    r0, e0 = dispersion("hip", cells, m, p, 120, fp64=True)
    monkeypatch.setenv("ALIS_FUSE", "force")
    r1, e1 = dispersion("hip", cells, m, p, 120, fp64=True)
    assert len(r0) > 20 and np.array_equal(r0, r1) and e0 == e1


def _libs():
    w = ctypes.CDLL(str(ROOT / "alis_amd" / "libalis_workloads.so"))
    w.wl_create.restype = ctypes.c_void_p
    w.wl_create.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    w.wl_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
    w.wl_free.argtypes = [ctypes.c_void_p]
    w.wl_device.restype = ctypes.c_void_p
    w.wl_device.argtypes = [ctypes.c_void_p]
    w.wl_field.restype = ctypes.c_long
    w.wl_field.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 4 + [ctypes.POINTER(ctypes.c_double), ctypes.c_long]
    e = ctypes.CDLL(str(ROOT / "alis_amd" / "libalis.so"))
    e.alis_dev_timing.argtypes = [ctypes.c_void_p, ctypes.c_int]
    e.alis_dev_timing_read.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_long)]
    return w, e


def _run(name, cells, steps, timed=4):
    """Final E and H of a workload after `steps` steps, and the launches of the last `timed`
    steps per timing slot (E curl, H curl, k_step)."""
    w, e = _libs()
    c = w.wl_create(name.encode(), cells, 32, 64)
    w.wl_step(c, steps - timed)
    dev = w.wl_device(c)
    assert e.alis_dev_timing(dev, 1) == 0
    w.wl_step(c, timed)
    launches = []
    for h in range(3):
        ms, n = ctypes.c_double(), ctypes.c_long()
        assert e.alis_dev_timing_read(dev, h, ctypes.byref(ms), ctypes.byref(n)) == 0
        launches.append(n.value)
    e.alis_dev_timing(dev, 0)
    per = (cells + 4) ** 3
    buf = (ctypes.c_double * per)()
    out = []
    for h in range(2):
        for a in range(3):
            n = w.wl_field(c, 0, h, 0, a, buf, per)
            assert n > 0
            out.append(np.frombuffer(buf, dtype=np.float64, count=n).copy())
    w.wl_free(c)
    return np.concatenate(out), launches


def test_fused_full_size_equals_separate(monkeypatch):
    """The 192^3 nanoblock (TF/SF box, gold block, CPML, phaser) after 60 steps: fused with
    ALIS_FUSE=1, identical to the separate half steps (the default)."""
    print("[gpu] nanoblock 192^3, fused vs separate half steps", flush=True)
    monkeypatch.setenv("ALIS_FUSE", "1")
    fused, n1 = _run("nanoblock", 192, 60)
    monkeypatch.delenv("ALIS_FUSE", raising=False)
    separate, n0 = _run("nanoblock", 192, 60)
    print(f"[gpu] launches fused {n1}, separate {n0}")
    assert n1 == [0, 0, 4] and n0 == [4, 4, 0]
    assert np.abs(fused).max() > 0
    assert np.array_equal(fused, separate)
