"""HBM traffic per launch of the curl kernels from two rocprofv3 counter runs.

    rocprofv3 --pmc FETCH_SIZE -d DIR_F -o run --output-format csv -- python3 bench.py \\
        --steps 3 --warmup 1 --no-fp64 --cpu-steps 0 --copy-mib 256
    rocprofv3 --pmc WRITE_SIZE -d DIR_W -o run --output-format csv -- python3 bench.py \\
        --steps 3 --warmup 1 --no-fp64 --cpu-steps 0 --copy-mib 256
    python3 alis_amd/profile_traffic.py DIR_F DIR_W 256 > profiles/k_curl_nanoblock256_traffic.json

FETCH_SIZE and WRITE_SIZE (KiB per dispatch) are scaled by what they read for the
device-to-device copy of known size that bench.py makes first (the largest copy dispatch).
"""
import csv
import hashlib
import json
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def kernels_sha256():
    """SHA-256 of the device kernels of the engine (the "Kernels" section of engine.hip and
    the cell update it includes, alis_amd/csrc/curl_cell.inc): a traffic measurement stays
    valid while these are unchanged."""
    src = (ROOT / "alis_amd/csrc/engine.hip").read_text()
    a = src.index("// Kernels\n")
    b = src.index("// Device world\n", a)
    cell = (ROOT / "alis_amd/csrc/curl_cell.inc").read_text()
    return hashlib.sha256((src[a:b] + cell).encode()).hexdigest()


def rows(d):
    f = next(Path(d).rglob("*counter_collection.csv"))
    return list(csv.DictReader(f.open()))


def main():
    dF, dW, mib = sys.argv[1], sys.argv[2], int(sys.argv[3])
    F, W = rows(dF), rows(dW)
    nbytes = mib << 20
    cF = max(float(r["Counter_Value"]) for r in F if "copyBuffer" in r["Kernel_Name"])
    cW = max(float(r["Counter_Value"]) for r in W if "copyBuffer" in r["Kernel_Name"])
    fs, ws = nbytes / 1024 / cF, nbytes / 1024 / cW
    cells = max(int(r["Grid_Size"]) for r in F if "k_curl" in r["Kernel_Name"])
    out = {"kernels_sha256": kernels_sha256(),
           "cells": cells, "fields": "fp32",
           "calibration": {"copy_bytes": nbytes, "FETCH_SIZE_kib": cF, "WRITE_SIZE_kib": cW,
                           "fetch_scale": round(fs, 5), "write_scale": round(ws, 5)}}
    # k_curl<float, CM, ...>: CM 0 is the H launch (no 1/eps), 1 or 2 the E launch
    # (the largest launches of each kind: the main world, not its TF/SF auxiliary world)
    for name, pats in (("E", ("k_curl<float, 1", "k_curl<float, 2")), ("H", ("k_curl<float, 0",))):
        big = max(int(r["Grid_Size"]) for r in F if any(p in r["Kernel_Name"] for p in pats))

        def mine(r):
            return any(p in r["Kernel_Name"] for p in pats) and int(r["Grid_Size"]) == big
        f = [float(r["Counter_Value"]) for r in F if mine(r)]
        w = [float(r["Counter_Value"]) for r in W if mine(r)]
        fa, wa = sum(f) / len(f), sum(w) / len(w)
        out[name] = {"FETCH_SIZE_kib": round(fa, 1), "WRITE_SIZE_kib": round(wa, 1), "dispatches": len(f),
                     "traffic_bytes": round((fa * fs + wa * ws) * 1024)}
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
