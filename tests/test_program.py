"""A whole alis.h program end to end (tests/golden/dipole_power.c: a y-oriented dipole behind
three symmetry planes, phasers over 500-900 nm, stepped under timer() with updateE, updateH
and updatePhaser), built against the CPU oracle and against the engine: the radiated power
from the phaser surface (phaserPoyntingOut), the far-field intensity along z times 8 pi / 3
(farFieldI) and twice the upper-hemisphere power (farFieldFlux) are the dipole's known
answers, 1, to within 1 % (the normalised source radiates unit power with a sin^2 pattern);
the GPU run gives the oracle's numbers to 1e-6."""
import subprocess

import numpy as np
import pytest

from scenarios import ROOT

SRC = ROOT / "tests" / "golden" / "dipole_power.c"


def _run(tmp_path, libdir, libname, env=None):
    exe = tmp_path / f"dipole_power_{libname}"
    r = subprocess.run(["gcc", "-std=c99", "-fopenmp", "-O2", f"-I{ROOT/'include'}", str(SRC), f"-L{libdir}", f"-l{libname}",
                        f"-Wl,-rpath,{libdir}", "-lm", "-o", str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600, cwd=tmp_path, env=env)
    assert r.returncode == 0, r.stdout[-1000:] + r.stderr[-1000:]
    return np.array([float(v) for v in r.stdout.strip().splitlines()[-1].split()])


def test_dipole_known_answers_oracle(tmp_path):
    v = _run(tmp_path, ROOT / "oracle", "oracle")
    print("oracle: power, 8 pi/3 I(0,0), 2 x hemisphere:", v)
    assert v.shape == (3,) and np.abs(v - 1).max() < 0.01


@pytest.mark.gpu
def test_dipole_known_answers_gpu(tmp_path):
    print("[gpu] dipole_power.c against libalis.so, and against the oracle", flush=True)
    g = _run(tmp_path, ROOT / "alis_amd", "alis")
    o = _run(tmp_path, ROOT / "oracle", "oracle")
    print("[gpu] power, 8 pi/3 I(0,0), 2 x hemisphere:", g, "oracle:", o)
    assert g.shape == (3,) and np.abs(g - 1).max() < 0.01
    assert np.abs(g - o).max() < 1e-6
