"""Text outputs against the reference's own files: writeRow, writeTxt, writePoyntingSpectrum,
writeFarFieldSpectrum, farFieldTheta and farFieldPhi of two scenarios with phaser arrays
(symband: symmetry planes; nanoblock: a phaser box around a TF/SF-lit gold block), written by
tests/golden/drive.c (alis_drive_text) over this library and, for the fixtures under
tests/golden/text/, over the reference library (tests/golden/make_golden.py).

The file set, the row and column counts and the first column (wavelengths and angles, %g of
the same floats) match exactly, as do the rows of given numbers and the formatted text; the
flux and far-field columns, which this library computes in double and the reference in
float, agree to 1e-4 of the file's largest value (a column of one polarisation that
symmetry makes zero holds only rounding noise, in both).  Beyond the reference: a text with a
'%' is written as formatted, one longer than 1 KiB whole, and a row with no numbers creates
no file."""
import ctypes
from pathlib import Path

import numpy as np
import pytest

from scenarios import drive_lib

GOLDEN = Path(__file__).resolve().parent / "golden" / "text"
CASES = ("symband", "nanoblock")
EXACT = ("rows.txt", "notes-3.txt")
TOL = 1e-4


def _write(kind, name, tmp_path):
    lib = drive_lib(kind)
    lib.alis_drive_text.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p]
    d = tmp_path / kind
    d.mkdir()
    assert lib.alis_drive_text(name.encode(), 200, str(d).encode()) == 0
    return d


def _table(path):
    return [line.split("\t") for line in path.read_text().splitlines()]


def _check(d, name):
    ref = GOLDEN / name
    want = sorted(p.name for p in ref.iterdir())
    got = sorted(p.name for p in d.iterdir())
    assert got == sorted(want + ["long.txt", "percent.txt"])
    assert (d / "percent.txt").read_text() == "100%\n"
    assert (d / "long.txt").read_text() == "x" * 3000 + "\n"
    for f in want:
        if f in EXACT:
            assert (d / f).read_text() == (ref / f).read_text(), f
            continue
        a, b = _table(d / f), _table(ref / f)
        assert len(a) == len(b) > 0, f
        assert [len(r) for r in a] == [len(r) for r in b], f
        assert [r[0] for r in a] == [r[0] for r in b], f
        x = np.array([[float(v) for v in r[1:]] for r in a])
        y = np.array([[float(v) for v in r[1:]] for r in b])
        err = np.abs(x - y).max() / max(np.abs(y).max(), 1e-30)
        assert err <= TOL, (f, err)


@pytest.mark.parametrize("name", CASES)
def test_text_oracle_matches_reference(name, tmp_path):
    _check(_write("oracle", name, tmp_path), name)


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_text_gpu_matches_reference(name, tmp_path):
    print(f"[gpu] {name}: text outputs", flush=True)
    _check(_write("hip", name, tmp_path), name)
