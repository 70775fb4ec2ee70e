"""Helpers of the parity tests: run the scenarios of tests/golden/drive.c over one alis.h
implementation (the HIP library or the CPU oracle), parse their "key value..." lines and
compare them with the reference's golden fixtures."""
import ctypes
import json
import os
import re
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = ROOT / "tests" / "golden"
MANIFEST = json.loads((GOLDEN / "manifest.json").read_text())
SCENARIOS = MANIFEST["scenarios"]
WITH_PHASERS = ("symband", "nanoblock", "oblique", "pillars", "particles")
# "kpoints" runs the 40 worlds kp0..kp39 of the k-point sweep together; their output keys
# start with their own names
KPOINTS = tuple(f"kp{q}" for q in range(40))


def members(name):
    """The world names whose output lines a scenario name produces."""
    return KPOINTS if name == "kpoints" else (name,)


def split_run(got, names):
    """The lines of a run of several scenarios, per scenario name."""
    owner = {m: n for n in names for m in members(n)}
    out = {n: {} for n in names}
    for k, v in got.items():
        n = owner.get(k.split(".")[0].split("#")[0])
        if n is not None:
            out[n][k] = v
    return out

LIBS = {"hip": GOLDEN / "libdrive.so", "oracle": GOLDEN / "libdrive_oracle.so"}
_loaded = {}


def drive_lib(kind):
    """The scenario library over the HIP engine ("hip") or the CPU oracle ("oracle")."""
    if kind not in _loaded:
        path = LIBS[kind]
        if not path.exists():
            raise RuntimeError(f"{path} is not built: run __graft_entry__.build()")
        lib = ctypes.CDLL(str(path))
        lib.alis_drive.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p]
        lib.alis_drive_raw.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p]
        lib.alis_physics_dispersion.argtypes = [ctypes.c_int] * 5 + [ctypes.POINTER(ctypes.c_double)] * 2
        dp = ctypes.POINTER(ctypes.c_double)
        lib.alis_drive_farfield.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, dp, dp, ctypes.c_int, dp]
        _loaded[kind] = lib
    return _loaded[kind]


def run_scenario(kind, name, steps, precision):
    """Run a scenario; returns its output lines as {key: [tokens]}."""
    os.environ["ALIS_PRECISION"] = precision
    with tempfile.TemporaryDirectory() as d:
        out = Path(d) / "out.txt"
        rc = drive_lib(kind).alis_drive(name.encode(), steps, str(out).encode())
        assert rc == 0, f"scenario {name} returned {rc}"
        return parse(out.read_text())


def run_raw(kind, name, steps, precision, spectra=False):
    """Final device state of a scenario as doubles: (fields of every world, the first world's
    phaser accumulators[, the raw sums of every k-point world's probe spectrum])."""
    os.environ["ALIS_PRECISION"] = precision
    with tempfile.TemporaryDirectory() as d:
        out = Path(d) / "raw.bin"
        rc = drive_lib(kind).alis_drive_raw(name.encode(), steps, str(out).encode())
        assert rc == 0, f"raw scenario {name} returned {rc}"
        b = out.read_bytes()
    head = np.frombuffer(b[:64], dtype=np.int64)
    data = np.frombuffer(b[64:], dtype=np.float64)
    fields, phaser = data[:head[4]].copy(), data[head[4]:head[4] + head[5]].copy()
    if spectra:
        return fields, phaser, data[head[4] + head[5]:head[4] + head[5] + head[6]].copy()
    return fields, phaser


# directions of the far-field checks (degrees): normal, oblique, both sides of the 90-degree
# average, the lower half space, and azimuths that mix the polarisation components
FF_THETA = (0.0, 20.0, 45.0, 89.5, 90.0, 120.0, 135.0, 180.0, 271.0, 330.0)
FF_PHI = (0.0, 30.0, 0.0, 10.0, 45.0, 200.0, 75.0, 0.0, 300.0, 135.0)


def farfield(kind, name, steps, precision, theta=FF_THETA, phi=FF_PHI, maxq=64):
    """Double-precision far field of a scenario's phasers: array [frequency][direction][P re,
    P im, S re, S im]."""
    os.environ["ALIS_PRECISION"] = precision
    n = len(theta)
    th = (ctypes.c_double * n)(*theta)
    ph = (ctypes.c_double * n)(*phi)
    out = (ctypes.c_double * (maxq * n * 4))()
    nq = drive_lib(kind).alis_drive_farfield(name.encode(), steps, n, th, ph, maxq, out)
    assert nq > 0, f"far field of {name} returned {nq}"
    return np.array(out[:nq * n * 4]).reshape(nq, n, 4)


def dispersion(kind, cells, m, p, steps, fp64):
    r = (ctypes.c_double * steps)()
    e = ctypes.c_double()
    n = drive_lib(kind).alis_physics_dispersion(cells, m, p, steps, int(fp64), r, ctypes.byref(e))
    return np.array(r[:n]), e.value


def parse(text):
    """{key: [tokens]}; a key repeated on later lines gets #1, #2, ... appended."""
    out = {}
    for line in text.splitlines():
        t = line.split()
        if not t:
            continue
        key, n = t[0], 0
        while key in out:
            n += 1
            key = f"{t[0]}#{n}"
        out[key] = t[1:]
    return out


def golden(name):
    return parse((GOLDEN / f"{name}.txt").read_text())


_FLOAT_POST = re.compile(r"\.(flux|ff\d+|poynting|absorb\d+)$")   # double-precision post-processing
_PHASER_DATA = re.compile(r"\.ph\d+\.(xMin|xMax|yMin|yMax|zMin|zMax)\.")


def compare(ref, got, phaser_exact=True, post_tol=1e-4, phaser_tol=1e-5):
    """Differences between a fixture and a run.  State, setup, probe traces and (with
    phaser_exact) phaser accumulators must be identical; flux and far-field values are
    computed in double here and in float by the reference, so they agree to post_tol
    relative to the largest magnitude on their line.  Without phaser_exact (FP64
    accumulators) the accumulators' sums of squares agree to phaser_tol, and the numbers of a
    spectrum file (writeSpectrum) sum to within 1e-4."""
    bad = []
    if set(ref) != set(got):
        bad.append(f"keys differ: only in fixture {sorted(set(ref) - set(got))[:5]}, only in run {sorted(set(got) - set(ref))[:5]}")
    for key in sorted(set(ref) & set(got)):
        a, b = ref[key], got[key]
        if _FLOAT_POST.search(key):
            x, y = np.array(a, dtype=float), np.array(b, dtype=float)
            scale = max(np.abs(x).max(), np.abs(y).max(), 1e-300)
            if x.shape != y.shape or np.abs(x - y).max() > post_tol * scale:
                bad.append(f"{key}: {a} vs {b}")
        elif not phaser_exact and key.endswith(".spectrum"):
            # the spectrum file's numbers (six digits) from double instead of float sums
            x, y = float(a[1]), float(b[1])
            if not abs(x - y) <= 1e-4 * max(abs(x), 1e-300):
                bad.append(f"{key}: sum of the file's numbers {x} vs {y}")
        elif not phaser_exact and _PHASER_DATA.search(key):
            x, y = float(a[1]), float(b[1])
            if abs(x - y) > phaser_tol * max(abs(x), 1e-300):
                bad.append(f"{key}: sum of squares {x} vs {y}")
        elif a != b:
            bad.append(f"{key}: {' '.join(a)} vs {' '.join(b)}")
    return bad


# Keys written before the first time step (setup arrays, extents, coefficients).
_SETUP = re.compile(r"^[a-z]+\d*(\.aux\d+)*\.(grid|phase|clock|dt[EH]\.[xyz]|cpml[EH]\.[xyz](#\d)?|eps\.[xyz]|obj\d+.*|src\d+.*|tfsf\d+)$")


def setup_keys(d):
    return {k: v for k, v in d.items() if _SETUP.match(k)}
