"""The drop-in boundary: the built libraries load and export every function that
include/alis.h (the reference's public API, src/alis_c.h:132-334), include/alis_dev.h
(the engine ABI) and include/alis_serve.h (the serving loop) declare; the product library depends on no oracle code; without a GPU the
product stops with an error instead of computing on the CPU."""
import ctypes
import os
import re
import subprocess
import sys

import pytest

from scenarios import ROOT

LIB = ROOT / "alis_amd" / "libalis.so"
ORACLE = ROOT / "oracle" / "liboracle.so"


def declared(header):
    text = (ROOT / "include" / header).read_text()
    text = text.split("#ifndef ALIS_LIBRARY")[0]          # materials: defined in the user's file
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"ALIS_CPLX\((\w+)\)", r"\1_complex", text)
    text = re.sub(r'#ifdef __cplusplus\s*\n(extern "C" \{|\})\s*\n#endif', "", text)
    while re.search(r"\{[^{}]*\}", text):                 # drop struct/union bodies
        text = re.sub(r"\{[^{}]*\}", " ", text)
    names = set()
    for stmt in re.findall(r"^[^#\n][^;]*;", text, flags=re.M):
        s = " ".join(stmt.split())
        if s.startswith(("typedef", "}")) or "=" in s or "(*" in s:
            continue
        m = re.match(r"^(shape|field|waveform|projmap)\s+([\w\s,]+);$", s)
        if m:
            names.update(n.strip() for n in m.group(2).split(","))
            continue
        m = re.search(r"\((\w+)\)\s*\(", s) or re.search(r"(\w+)\s*\(", s)
        if m:
            names.add(m.group(1))
    return names


def test_headers_parse():
    api = declared("alis.h")
    dev = declared("alis_dev.h")
    for n in ("createWorld", "updateE", "updateH", "createPhasers", "updatePhaser", "farField", "get",
              "writeSpectrum", "timer", "planewave", "pointDipole", "Ex", "Box", "Sine", "Ppol"):
        assert n in api, n
    assert {"alis_dev_commit", "alis_dev_core", "alis_dev_strips", "alis_dev_ghosts", "alis_dev_phaser_update"} <= dev
    assert len(api) > 150 and len(dev) >= 20
    assert declared("alis_serve.h") == {"alisServe", "alisServeLastStats", "alisServeMaxRadius", "alisServeRequests"}


@pytest.mark.parametrize("header", ["alis.h", "alis_dev.h", "alis_serve.h"])
def test_library_exports_declared_symbols(header):
    lib = ctypes.CDLL(str(LIB))
    missing = sorted(n for n in declared(header) if not hasattr(lib, n))
    assert not missing, missing


def test_oracle_implements_engine_abi():
    lib = ctypes.CDLL(str(ORACLE))
    missing = sorted(n for n in declared("alis_dev.h") if not hasattr(lib, n))
    assert not missing, missing


def test_product_links_no_oracle():
    out = subprocess.run(["readelf", "-d", str(LIB)], capture_output=True, text=True, check=True).stdout
    needed = re.findall(r"\[(.*?)\]", "\n".join(l for l in out.splitlines() if "NEEDED" in l))
    assert "libamdhip64.so.7" in needed or any(n.startswith("libamdhip64") for n in needed)
    assert not any("oracle" in n or "ref" in n for n in needed), needed


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is visible here")
def test_no_gpu_fails_loudly(tmp_path):
    code = ("import ctypes,sys; l=ctypes.CDLL(sys.argv[1]); "
            "l.alis_drive.argtypes=[ctypes.c_char_p,ctypes.c_int,ctypes.c_char_p]; "
            "l.alis_drive(b'dipole', 3, sys.argv[2].encode())")
    r = subprocess.run([sys.executable, "-c", code, str(ROOT / "tests/golden/libdrive.so"), str(tmp_path / "o.txt")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "no GPU" in r.stderr
