"""The CPU oracle (oracle/engine.c under the host library) against the reference's own
outputs: every scenario of tests/golden/drive.c, run with float accumulators like the
reference, must reproduce the fixture bit for bit (state, CPML, ADE currents, TF/SF
auxiliary worlds, probe traces, phaser accumulators); flux and far field, evaluated in
double here, agree to 1e-4 of their scale."""
import pytest

from scenarios import split_run, SCENARIOS, compare, golden, run_scenario


@pytest.mark.parametrize("name", sorted(SCENARIOS))
def test_oracle_matches_reference(name):
    got = run_scenario("oracle", name, SCENARIOS[name], "ref")
    bad = compare(golden(name), got)
    assert not bad, "\n".join(bad[:20])


@pytest.mark.parametrize("name", ["nanoblock", "oblique"])
def test_oracle_default_precision(name):
    """FP32 fields with FP64 phaser accumulators: fields stay bitwise, the spectra agree
    to 1e-5 in their sums of squares."""
    got = run_scenario("oracle", name, SCENARIOS[name], "fp32")
    bad = compare(golden(name), got, phaser_exact=False)
    assert not bad, "\n".join(bad[:20])


def test_fixture_catches_planted_mistake(tmp_path):
    """The comparison is sharp: one changed last digit of one state value fails it."""
    ref = golden("dipole")
    got = dict(ref)
    key = next(k for k in ref if k.endswith(".E0.z"))
    h, s = got[key]
    got[key] = [h, s[:-1] + ("1" if s[-1] != "1" else "2")]
    assert compare(ref, got)




def test_oracle_batch_matches_reference():
    """All scenarios stepped together (updateWorldsE/updateWorldsH over the batch ABI) give
    each scenario's own reference output."""
    names = sorted(SCENARIOS)
    got = split_run(run_scenario("oracle", ",".join(names), SCENARIOS[names[0]], "ref"), names)
    for n in names:
        bad = compare(golden(n), got[n])
        assert not bad, n + ": " + "\n".join(bad[:10])
