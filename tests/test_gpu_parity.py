"""Parity of the HIP engine (alis_amd/libalis.so on the GPU) with the reference and the
oracle.

* Every scenario with float accumulation, as the reference computes: identical to the
  reference's golden fixture (fields with ghosts, CPML psi slabs, ADE J/K, TF/SF
  auxiliary worlds, probe traces, phaser accumulators); flux and far field (double
  post-processing) to 1e-4.
* Default precision (FP32 fields, FP64 phaser accumulators): fields still identical; the
  spectra agree with the float fixture to 1e-5 and with the FP64 oracle to 1e-5.
* FP64 worlds: the final device state agrees with the FP64 oracle to 1e-12 relative.
* The closed-form dispersion recurrence of a periodic vacuum box, FP64 on the device.
* Phaser reads between deferred flushes (37 steps), alone and batched: bitwise the oracle.
* Full 256^3 nanoblock size: the GPU's state after 300 steps has the same checksum as the
  reference's own C code run on the host (oracle/_ref/wl_ref), and FP64 tracks FP32.
"""
import ctypes
import subprocess
from pathlib import Path

import numpy as np
import pytest

from scenarios import split_run, ROOT, SCENARIOS, WITH_PHASERS, compare, dispersion, golden, run_raw, run_scenario

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", sorted(SCENARIOS))
def test_reference_precision_bitwise(name):
    print(f"[gpu] {name} float accumulation", flush=True)
    got = run_scenario("hip", name, SCENARIOS[name], "ref")
    bad = compare(golden(name), got)
    assert not bad, "\n".join(bad[:20])


@pytest.mark.parametrize("name", sorted(SCENARIOS))
def test_default_precision(name):
    print(f"[gpu] {name} fp32 fields / fp64 DFT", flush=True)
    got = run_scenario("hip", name, SCENARIOS[name], "fp32")
    bad = compare(golden(name), got, phaser_exact=False)
    assert not bad, "\n".join(bad[:20])


def _rel(a, b):
    scale = np.abs(b).max()
    return np.abs(a - b).max() / scale if scale else np.abs(a - b).max()


@pytest.mark.parametrize("name", sorted(SCENARIOS))
def test_fp64_matches_fp64_oracle(name):
    """FP64 worlds: every world's final state, the phaser accumulators and (k-point worlds)
    the probe spectra of writeSpectrum within 1e-12 of the FP64 oracle."""
    print(f"[gpu] {name} fp64 vs oracle", flush=True)
    fg, pg, sg = run_raw("hip", name, SCENARIOS[name], "fp64", spectra=True)
    fo, po, so = run_raw("oracle", name, SCENARIOS[name], "fp64", spectra=True)
    assert fg.shape == fo.shape and pg.shape == po.shape and sg.shape == so.shape
    assert _rel(fg, fo) <= 1e-12, _rel(fg, fo)
    if po.size:
        assert _rel(pg, po) <= 1e-12, _rel(pg, po)
    if name == "kpoints":
        assert so.size and np.abs(so).max() > 0
        assert _rel(sg, so) <= 1e-12, _rel(sg, so)


@pytest.mark.parametrize("name", WITH_PHASERS + ("kpoints",))
def test_fp32_spectra_vs_fp64_oracle(name):
    """The production precision: DFT spectra (phaser accumulators; writeSpectrum sums of the
    k-point worlds) within 1e-5 of the FP64 oracle."""
    print(f"[gpu] {name} fp32 spectra vs fp64 oracle", flush=True)
    fg, pg, sg = run_raw("hip", name, SCENARIOS[name], "fp32", spectra=True)
    fo, po, so = run_raw("oracle", name, SCENARIOS[name], "fp64", spectra=True)
    err_f = _rel(fg, fo)
    err_p = _rel(pg, po) if po.size else _rel(sg, so)
    print(f"[gpu] {name}: fp32 fields err {err_f:.2e}, spectra err {err_p:.2e} (vs fp64 oracle)")
    assert (po.size or so.size) and err_p <= 1e-5


@pytest.mark.parametrize("cells,m,p", [(17, 3, 2), (21, 8, 1)])
def test_dispersion_fp64_on_device(cells, m, p):
    r, expect = dispersion("hip", cells, m, p, 120, fp64=True)
    assert len(r) > 20
    assert np.abs(r - expect).max() < 1e-11 * abs(expect) + 1e-12


def _workloads():
    w = ctypes.CDLL(str(ROOT / "alis_amd" / "libalis_workloads.so"))
    w.wl_create.restype = ctypes.c_void_p
    w.wl_create.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    w.wl_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
    w.wl_checksum.restype = ctypes.c_double
    w.wl_checksum.argtypes = [ctypes.c_void_p]
    w.wl_free.argtypes = [ctypes.c_void_p]
    return w


def _gpu_checksum(name, cells, fbits, steps):
    w = _workloads()
    c = w.wl_create(name.encode(), cells, fbits, 64)
    w.wl_step(c, steps)
    s = w.wl_checksum(c)
    w.wl_free(c)
    return s


REF_WL = ROOT / "oracle" / "_ref" / "wl_ref"


@pytest.mark.skipif(not REF_WL.exists(), reason="reference build (oracle/_ref) not present")
def test_full_size_nanoblock_matches_reference():
    print("[gpu] nanoblock 256^3, 300 steps vs the reference's C code", flush=True)
    ref = subprocess.run([str(REF_WL), "nanoblock", "256", "2", "298"], capture_output=True, text=True,
                         timeout=300, check=True).stdout.split()
    gpu = _gpu_checksum("nanoblock", 256, 32, 300)
    print(f"[gpu] checksum gpu {gpu!r} reference {ref[3]}")
    assert float(ref[3]) > 0
    assert gpu == float(ref[3])


def test_full_size_fp64_tracks_fp32():
    print("[gpu] nanoblock 256^3 fp64 vs fp32, 300 steps", flush=True)
    s32 = _gpu_checksum("nanoblock", 256, 32, 300)
    s64 = _gpu_checksum("nanoblock", 256, 64, 300)
    assert s32 > 0 and abs(s64 - s32) <= 1e-4 * s32




@pytest.mark.parametrize("precision", ["ref", "fp32"])
def test_batch_equals_alone(precision):
    """Continuous batching's invariant: every scenario stepped in one batch with all the
    others (shared curl and post launches, TF/SF auxiliary worlds included) gives exactly
    its result alone -- the reference's fixture in float accumulation, and the same bits as
    its solo GPU run in the default precision."""
    names = sorted(SCENARIOS)
    steps = SCENARIOS[names[0]]
    print(f"[gpu] batch of {len(names)} scenarios, {precision}", flush=True)
    mix = split_run(run_scenario("hip", ",".join(names), steps, precision), names)
    for n in names:
        alone = run_scenario("hip", n, steps, precision)
        assert mix[n] == alone, n
        if precision == "ref":
            bad = compare(golden(n), mix[n])
            assert not bad, n + ": " + "\n".join(bad[:10])


@pytest.mark.parametrize("prec", ["ref", "fp32"])
def test_flux_device_sums_bitwise(prec):
    """The net flux (phaserPoyntingOut, src/phaser.c:228-245) takes its inner trapezoid sums
    from k_flux_sums, which adds the cells in the order the oracle's restatement does: every
    phaser's flux of every phaser scenario equals the oracle's to the bit (the float the API
    returns, printed to round-trip), after 38 steps (one flush done, six samples pending in the
    ring)."""
    for name in sorted(WITH_PHASERS):
        print(f"[gpu] {name} {prec}: flux sums", flush=True)
        g, o = run_scenario("hip", name, 38, prec), run_scenario("oracle", name, 38, prec)
        keys = [k for k in o if k.endswith(".flux")]
        assert keys and all(k in g for k in keys)
        assert all(g[k] == o[k] for k in keys), [(k, g[k], o[k]) for k in keys if g[k] != o[k]][:4]
        assert any(o[k] != 0 for k in keys)


@pytest.mark.parametrize("prec", ["ref", "fp32"])
def test_observers_device_sums_bitwise(prec):
    """poyntingX/Y/Z/Out and objectAbsorption (src/output.c:233-435) take their sums from the
    device (k_plane_sums, k_absorption) without copying the world to the host, added in the
    order of the oracle's restatement: the values at the last step and the running sum of
    poyntingOut every 25 steps equal the oracle's to the bit in every scenario."""
    for name in sorted(n for n in SCENARIOS if n != "kpoints"):
        print(f"[gpu] {name} {prec}: observers", flush=True)
        g, o = run_scenario("hip", name, 60, prec), run_scenario("oracle", name, 60, prec)
        keys = [k for k in o if k.endswith(".poynting") or ".absorb" in k]
        assert keys and all(k in g for k in keys)
        assert all(g[k] == o[k] for k in keys), [(k, g[k], o[k]) for k in keys if g[k] != o[k]][:4]


@pytest.mark.parametrize("name", WITH_PHASERS)
def test_phaser_read_between_flushes(name):
    """Phaser updates reach the accumulators 32 samples at a time; a read after 37 steps
    (one flush done, five samples pending) returns the running sums of every update: in float
    accumulation the same bits as the oracle's one update per step, alone and in a batch
    (whose phasers are updated together)."""
    print(f"[gpu] {name}: phaser read after 37 steps", flush=True)
    fg, pg = run_raw("hip", name, 37, "ref")
    fo, po = run_raw("oracle", name, 37, "ref")
    assert pg.size and np.array_equal(pg, po) and np.array_equal(fg, fo)
    names = sorted(WITH_PHASERS)
    mix = split_run(run_scenario("hip", ",".join(names), 37, "ref"), names)
    bad = compare(run_scenario("oracle", name, 37, "ref"), mix[name])
    assert not bad, "\n".join(bad[:10])
