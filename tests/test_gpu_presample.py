"""Phaser samples of z faces stored by the H curl (k_curl's PhsZ faces) instead of read by
k_phs_sample: the same values, so every scenario with a phaser stays identical to the
reference's fixture.  ALIS_PRESAMPLE=force takes every eligible face (by default only faces
that cover at least half of a z plane, where it pays), so the small scenarios run through it.
"""
import pytest

from scenarios import SCENARIOS, WITH_PHASERS, compare, golden, run_raw, run_scenario

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", WITH_PHASERS)
def test_presampled_phasers_bitwise(name, monkeypatch):
    monkeypatch.setenv("ALIS_PRESAMPLE", "force")
    print(f"[gpu] {name} phaser z faces sampled by the H curl, float accumulation", flush=True)
    got = run_scenario("hip", name, SCENARIOS[name], "ref")
    bad = compare(golden(name), got)
    assert not bad, "\n".join(bad[:20])


@pytest.mark.parametrize("name", WITH_PHASERS)
def test_presampled_equals_read(name, monkeypatch):
    """FP64: the accumulators are the same with and without the stored samples."""
    monkeypatch.setenv("ALIS_PRESAMPLE", "force")
    f1, p1 = run_raw("hip", name, SCENARIOS[name], "fp64")
    monkeypatch.setenv("ALIS_PRESAMPLE", "0")
    f0, p0 = run_raw("hip", name, SCENARIOS[name], "fp64")
    assert p0.size and (f1 == f0).all() and (p1 == p0).all()
