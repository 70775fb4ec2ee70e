"""The product's host library (alis_amd/libalis.so, through tests/golden/libdrive.so) builds
the same setup as the reference: grid extents, Bloch phases, clock, dt/d, CPML profiles,
inverse permittivity, ADE tables and fill factors, source amplitudes, TF/SF boxes and
auxiliary worlds.  Zero steps: nothing is committed to a GPU."""
import ctypes

import pytest

from scenarios import SCENARIOS, drive_lib, golden, run_scenario, setup_keys


@pytest.mark.parametrize("name", sorted(SCENARIOS))
def test_setup_matches_reference(name):
    ref = setup_keys(golden(name))
    got = setup_keys(run_scenario("hip", name, 0, "ref"))
    assert len(ref) > 5
    assert set(ref) == set(got)
    bad = [f"{k}: {ref[k]} vs {got[k]}" for k in ref if ref[k] != got[k]]
    assert not bad, "\n".join(bad[:20])


def test_nearest_node_cache():
    """get() and the setup calls find grid nodes through a per-world cache of coordinate ->
    index: over 3000 coordinates (random, on nodes, between nodes; 64 slots per axis, so
    evictions and collisions) the cached indices equal those a fresh world computes."""
    lib = drive_lib("hip")
    lib.alis_drive_nodes.argtypes = [ctypes.c_int]
    assert lib.alis_drive_nodes(1000) == 0
