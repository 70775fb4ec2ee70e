"""timer() (src/update.c:14-45) prints the reference's progress line: a short loop run
through this library's host code (over the CPU oracle engine) and through the reference
library compiled from its sources (oracle/_ref, where present) prints the same bytes (the
loop takes well under a second, so both show zero times)."""
import subprocess

import pytest

from scenarios import ROOT

PROGRAM = r'''
#include "alis.h"
#include <stdio.h>
int main(void)
{
	for (long n = 0; timer(n, 5, "steps"); n++) ;
	for (long n = 0; timer(n, 3, ""); n++) ;
	timer(1, 2, "again");
	timer(3, 2, "again");
	return 0;
}
'''


def _run(tmp_path, name, link):
    src = tmp_path / "t.c"
    src.write_text(PROGRAM)
    exe = tmp_path / name
    subprocess.run(["gcc", "-std=gnu99", "-fopenmp", f"-I{ROOT / 'include'}", str(src), "-o", str(exe), *link, "-lm"],
                   check=True, capture_output=True)
    return subprocess.run([str(exe)], check=True, capture_output=True, timeout=60).stdout


def test_timer_line(tmp_path):
    out = _run(tmp_path, "ours", [f"-L{ROOT / 'oracle'}", "-loracle", f"-Wl,-rpath,{ROOT / 'oracle'}"])
    assert out == b"\rsteps [00:00:00/00:00:00] [5/5: 100.0%]\n\r[00:00:00/00:00:00] [3/3: 100.0%]\n" \
                  b"\ragain [00:00:00/00:00:00] [2/2: 100.0%]\n"


@pytest.mark.skipif(not (ROOT / "oracle" / "_ref" / "libref.a").exists(), reason="reference library not built here")
def test_timer_line_matches_reference(tmp_path):
    ours = _run(tmp_path, "ours", [f"-L{ROOT / 'oracle'}", "-loracle", f"-Wl,-rpath,{ROOT / 'oracle'}"])
    ref = _run(tmp_path, "ref", [str(ROOT / "oracle" / "_ref" / "libref.a")])
    assert ours == ref
