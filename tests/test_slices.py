"""Slices (sliceSnap, sliceTimeAvg, sliceFreqDom with the txt and png writers): the host
library samples a cut from mirrors refreshed over the cut's own box in the arrays the
accessor reads.  On the CPU oracle (whose copies cover every array) the files are complete
and well formed; on the GPU every file is byte for byte the oracle's, so the partial refresh
misses nothing.  symband has symmetry planes (mirrored negative indices) and a graded grid
(resampling onto the uniform grid); nanoblock has objects (the index contour).  The reference
writes its slices with libpng, which this image lacks, so its own files are not available:
the file contents are pinned to the oracle's, not to the reference's."""
import ctypes
import struct
import zlib

import numpy as np
import pytest

from scenarios import drive_lib

CASES = ("nanoblock", "symband")


def _slices(kind, name, tmp_path):
    lib = drive_lib(kind)
    lib.alis_drive_slices.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p]
    d = tmp_path / kind
    d.mkdir()
    assert lib.alis_drive_slices(name.encode(), 60, str(d).encode()) == 0
    return {p.name: p.read_bytes() for p in sorted(d.iterdir())}


@pytest.mark.parametrize("name", CASES)
def test_slices_oracle(name, tmp_path):
    files = _slices("oracle", name, tmp_path)
    want = {"avg-Hy-XY00001.txt", "avg-Hy-XY00002.txt", "image-Ez-XY.png", "line-Ex-X.txt",
            "snap-Ez-XY.txt", "snap-Hy-XY.txt"} | {f"phasor-Ez-XY{p:02d}.txt" for p in range(1, 13)}
    assert set(files) == want
    snap = np.loadtxt(bytes(files["snap-Ez-XY.txt"]).decode().splitlines())
    line = np.loadtxt(bytes(files["line-Ex-X.txt"]).decode().splitlines())
    assert snap.ndim == 2 and np.isfinite(snap).all() and np.abs(snap).max() > 0
    assert line.shape[1] == 2 and np.all(np.diff(line[:, 0]) > 0)
    png = files["image-Ez-XY.png"]
    assert png[:8] == b"\x89PNG\r\n\x1a\n"
    w, h = struct.unpack(">II", png[16:24])
    assert (h, w) == snap.shape
    n = struct.unpack(">I", png[33:37])[0]
    assert png[37:41] == b"IDAT" and len(zlib.decompress(png[41:41 + n])) == h * (3 * w + 1)


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_slices_gpu_equal_oracle(name, tmp_path):
    print(f"[gpu] {name}: slices", flush=True)
    g, o = _slices("hip", name, tmp_path), _slices("oracle", name, tmp_path)
    assert set(g) == set(o)
    bad = [f for f in o if g[f] != o[f]]
    assert not bad, bad
