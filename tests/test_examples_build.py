"""Drop-in check: the reference's own example programs compile against include/alis.h and
link against alis_amd/libalis.so without a change (they are read where the reference's
sources lie; skipped where those are absent)."""
import subprocess
from pathlib import Path

import pytest

from scenarios import ROOT

EXAMPLES = Path("/root/reference/examples")
PROGRAMS = sorted(EXAMPLES.glob("*.c")) if EXAMPLES.exists() else []


@pytest.mark.skipif(not PROGRAMS, reason="reference examples not present")
@pytest.mark.parametrize("src", PROGRAMS, ids=lambda p: p.stem)
def test_example_builds(src, tmp_path):
    r = subprocess.run(["gcc", "-std=c99", "-fopenmp", "-O0", f"-I{ROOT/'include'}", str(src),
                        f"-L{ROOT/'alis_amd'}", "-lalis", "-lm", "-o", str(tmp_path / src.stem)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
