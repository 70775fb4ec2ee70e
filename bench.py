#!/usr/bin/env python3
"""Throughput of the FDTD hot path on one MI355X.

Workload: the 256^3 nanoblock scattering run (grid edge 256 including a 16-cell CPML on
every face, dx = 2 nm, Courant 0.5; a 60x60x40 nm gold block with a Drude pole and two
Lorentz pairs on an n = 1.5 substrate; TF/SF x-polarised plane wave with a 400-900 nm band
waveform; 32-frequency phaser box), built by alis_amd/workloads.c through the public
alis.h API.  One step = updateE + updateH + updatePhaser.  Fields in FP32 with FP64 phaser
accumulators (the headline), then the same run with FP64 fields for comparison.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cells 256] [--cpu-steps S]

Prints one JSON line.  `value` is Mcell-updates/s of the FP32 run (cells including CPML x
steps / wall time, inputs resident in HBM, device synchronised on both sides of the timed
region).  `roofline` is the dominant kernel (the E-field curl + CPML update): algorithmic
bytes per launch / mean launch time from HIP events recorded on the world's stream around
every curl launch of the timed region.  `cpu_baseline` times the reference's own C code
(compiled from its sources into oracle/_ref) on the same workload on this host's cores.
"""
import argparse
import ctypes
import json
import math
import os
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
PEAK_HBM = 8.0e12        # MI355X HBM3E, bytes/s


class quiet:
    """Send the library's progress lines (C stdout) to /dev/null; the JSON line is the only
    thing this program prints."""
    def __enter__(self):
        sys.stdout.flush()
        self.saved = os.dup(1)
        null = os.open(os.devnull, os.O_WRONLY)
        os.dup2(null, 1)
        os.close(null)

    def __exit__(self, *exc):
        os.dup2(self.saved, 1)
        os.close(self.saved)


def load():
    lib = ROOT / "alis_amd" / "libalis.so"
    wl = ROOT / "alis_amd" / "libalis_workloads.so"
    if not lib.exists() or not wl.exists():
        sys.exit("bench: alis_amd/libalis.so is not built (run __graft_entry__.build())")
    alis = ctypes.CDLL(str(lib))
    w = ctypes.CDLL(str(wl))
    w.wl_create.restype = ctypes.c_void_p
    w.wl_create.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    w.wl_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
    w.wl_cells.restype = ctypes.c_long
    w.wl_cells.argtypes = [ctypes.c_void_p]
    w.wl_phaser_info.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_long)]
    w.wl_geometry.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    w.wl_device.restype = ctypes.c_void_p
    w.wl_flux.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    w.wl_device.argtypes = [ctypes.c_void_p]
    w.wl_free.argtypes = [ctypes.c_void_p]
    w.wl_checksum.restype = ctypes.c_double
    w.wl_checksum.argtypes = [ctypes.c_void_p]
    w.wl_worlds.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
    w.wl_farfield.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p]
    alis.alis_dev_timing.argtypes = [ctypes.c_void_p, ctypes.c_int]
    alis.alis_dev_timing_read.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_long)]
    alis.alis_dev_synchronize.argtypes = [ctypes.c_void_p]
    alis.alis_dev_bytes.restype = ctypes.c_long
    alis.alis_dev_bytes.argtypes = [ctypes.c_void_p]
    alis.alis_dev_eps_bytes.restype = ctypes.c_double
    alis.alis_dev_eps_bytes.argtypes = [ctypes.c_void_p]
    hip = ctypes.CDLL("libamdhip64.so")
    return alis, w, hip


def curl_bytes(num, pml, s, eps_bytes):
    """Algorithmic HBM bytes of one curl launch (the compulsory traffic: every state element
    read and written once, every coefficient read once): E from H reads E, H and writes E,
    3 components each, plus `eps_bytes` per cell of 1/eps; H from E reads H, E and writes H.
    CPML: each slab cell reads and writes its two psi values.  Slabs are Num = pml-1 cells
    thick over the full cross section (src/update.c:119-124).  The single-world E curl reads
    1/eps from a per-segment table where 64 cells along z hold one material (the engine's
    figure, alis_dev_eps_bytes: 0.34 bytes per cell on the nanoblock); batched launches read
    the three 1/eps arrays (3 s per cell)."""
    n = num[0] * num[1] * num[2]
    e = n * (9 * s + eps_bytes)
    h = n * 9 * s
    slab = 0
    for f in range(6):
        if pml[f]:
            ax = f // 2
            area = num[0] * num[1] * num[2] // num[ax]
            slab += (pml[f] - 1) * area
    return e + slab * 4 * s, h + slab * 4 * s


def run(alis, w, hip, name, cells, fbits, dbits, warmup, steps, farfield=None, flux=False):
    ctx = w.wl_create(name.encode(), cells, fbits, dbits)
    if not ctx:
        sys.exit(f"bench: unknown workload {name}")
    w.wl_step(ctx, max(warmup, 1))                  # the first step commits the world to HBM
    dev = w.wl_device(ctx)
    hip.hipDeviceSynchronize()
    alis.alis_dev_timing(dev, 1)
    hip.hipDeviceSynchronize()
    t0 = time.perf_counter()
    w.wl_step(ctx, steps)
    hip.hipDeviceSynchronize()
    t1 = time.perf_counter()
    ms, n = ctypes.c_double(), ctypes.c_long()
    kern = []
    for h in (0, 1):
        alis.alis_dev_timing_read(dev, h, ctypes.byref(ms), ctypes.byref(n))
        kern.append((ms.value, n.value))
    alis.alis_dev_timing(dev, 0)
    num, pml = (ctypes.c_int * 3)(), (ctypes.c_int * 6)()
    eps = w.wl_geometry(ctx, num, pml)
    surf = ctypes.c_long()
    nf = w.wl_phaser_info(ctx, ctypes.byref(surf))
    cplx = ctypes.c_int()
    nw = w.wl_worlds(ctx, ctypes.byref(cplx))
    out = {"cells": w.wl_cells(ctx), "sec": t1 - t0, "kern": kern, "num": list(num), "pml": list(pml),
           "eps": eps, "eps_bytes": alis.alis_dev_eps_bytes(dev), "nf": nf, "surf": surf.value,
           "hbm_bytes": alis.alis_dev_bytes(dev),
           "worlds": nw, "parts": 1 + cplx.value}
    if flux:
        # net flux of every frequency (phaserPoyntingOut) after one more step: the first call
        # adds the pending phaser samples and sums every frequency's faces in one k_flux_sums
        # launch; the host keeps the face fluxes until the next update, so the other calls
        # only add them up
        vals = (ctypes.c_double * 64)()
        w.wl_flux(ctx, vals, 1)                          # sizes the device workspace
        w.wl_step(ctx, 1)
        hip.hipDeviceSynchronize()
        t0 = time.perf_counter()
        n = w.wl_flux(ctx, vals, 64)
        out["flux"] = {"frequencies": n, "ms_per_frequency": round((time.perf_counter() - t0) / max(n, 1) * 1e3, 4),
                       "accumulator_bytes_per_frequency": surf.value * 4 * 2 * (8 if dbits == 64 else 4),
                       "finite": all(math.isfinite(vals[q]) for q in range(n))}
    if farfield:
        import tempfile
        X, Y = farfield
        w.wl_farfield(ctx, 1, 1, b"")                    # first call: far-field set-up
        hip.hipDeviceSynchronize()
        t0 = time.perf_counter()
        w.wl_farfield(ctx, X, Y, b"")                    # the map without its image files
        hip.hipDeviceSynchronize()
        out["farfield_sec"] = time.perf_counter() - t0
        with tempfile.TemporaryDirectory() as d:
            t0 = time.perf_counter()
            w.wl_farfield(ctx, X, Y, d.encode())         # and with its seven text images
            out["farfield_files_sec"] = time.perf_counter() - t0
    w.wl_free(ctx)
    return out


def roofline(r, s, dft_s, steps):
    be, bh = curl_bytes(r["num"], r["pml"], s, r["eps_bytes"])
    be_ref, _ = curl_bytes(r["num"], r["pml"], s, 3 * s if r["eps"] else 0)
    be, bh = be * r["worlds"] * r["parts"], bh * r["worlds"] * r["parts"]    # one launch covers all
    (mse, ne), (msh, nh) = r["kern"]
    te, th = mse / max(ne, 1) * 1e-3, msh / max(nh, 1) * 1e-3
    dom = ("E", be, te, ne) if mse >= msh else ("H", bh, th, nh)
    ach = dom[1] / dom[2] if dom[2] else 0.0
    # whole step: both curls + phaser accumulators (read+write, re/im, 4 components)
    phaser = r["surf"] * 4 * r["nf"] * 2 * 2 * dft_s
    step_bytes = be + bh + phaser
    return {
        "bound": "hbm", "achieved": round(ach / 1e9, 1), "peak": PEAK_HBM / 1e9, "unit": "GB/s",
        "frac": round(ach / PEAK_HBM, 4), "traffic": None,
        "kernel": f"k_curl ({dom[0]} half step, {'with' if dom[0] == 'E' and r['eps'] else 'no'} 1/eps)",
        "bytes_per_launch": dom[1], "launch_ms": round(dom[2] * 1e3, 5), "launches": dom[3],
        "other_curl": {"bytes_per_launch": bh if dom[0] == "E" else be,
                       "launch_ms": round((th if dom[0] == "E" else te) * 1e3, 5)},
        "eps_read_bytes_per_cell": round(r["eps_bytes"], 3),
        # the same launch priced with the reference's representation (three 1/eps values per
        # cell): an effective rate, not the bytes this kernel moves
        "effective_gbs_three_eps_arrays": round(be_ref * r["worlds"] * r["parts"] / te / 1e9, 1) if te and dom[0] == "E" else None,
        "step_bytes": step_bytes,
        "step_frac": round(step_bytes / (r["sec"] / steps) / PEAK_HBM, 4),
    }


def calibration_copy(hip, nbytes):
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(nbytes))
    hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(nbytes))
    hip.hipMemset(src, 0, ctypes.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3)      # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)


def measured_traffic(cells, fields):
    """HBM bytes per E-curl launch measured with rocprofv3 counters on these device kernels
    (profiles/k_curl_nanoblock256_traffic.json, made by alis_amd/profile_traffic.py), or
    None when that measurement was taken on other kernels or another workload."""
    f = ROOT / "profiles" / "k_curl_nanoblock256_traffic.json"
    if not f.exists():
        return None
    sys.path.insert(0, str(ROOT))
    from alis_amd.profile_traffic import kernels_sha256
    m = json.loads(f.read_text())
    if m.get("kernels_sha256") != kernels_sha256() or m.get("cells") != cells or m.get("fields") != fields:
        return None
    return m["E"]["traffic_bytes"]


def host_cpu():
    """CPU model and the cores this process may run on."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return model, cores


def cpu_baseline(cells, steps, runs=5):
    """The reference's own C code (oracle/_ref/wl_ref: its sources compiled with its flags)
    on the same workload, with OpenMP on every core of this process's affinity set: median
    Mcell-updates/s of `runs` runs of `steps` steps each (after 2 warm-up steps)."""
    exe = ROOT / "oracle" / "_ref" / "wl_ref"
    if not exe.exists():
        return None
    model, cores = host_cpu()
    env = dict(os.environ, OMP_NUM_THREADS=str(cores))
    vals, threads = [], cores
    for _ in range(runs):
        r = subprocess.run([str(exe), "nanoblock", str(cells), "2", str(steps)], capture_output=True, text=True,
                           env=env, timeout=600)
        if r.returncode:
            return None
        sec, ncell, threads = r.stdout.split()[:3]
        vals.append(int(ncell) / float(sec) / 1e6)
    vals.sort()
    return {"value": round(vals[len(vals) // 2], 2), "unit": "Mcell-updates/s", "cores": int(threads),
            "kind": "reference", "runs": [round(v, 1) for v in vals], "cpu": model, "nproc": os.cpu_count(),
            "sample": f"reference C (gcc -O3 -fopenmp, its own sources) on the same {cells}^3 nanoblock workload, "
                      f"median of {runs} runs of {steps} steps after 2 warm-up steps, OMP_NUM_THREADS={threads}"}


def farfield_cpu_baseline(X, Y, cells=160):
    """The reference's own far-field code (oracle/_ref/wl_ref) per direction on a reduced
    metasurface (cells x cells z faces), scaled to the full faces by their cell count: the
    cost of a direction is one pass over the face cells per component."""
    exe = ROOT / "oracle" / "_ref" / "wl_ref"
    if not exe.exists():
        return None
    import tempfile
    _, cores = host_cpu()
    env = dict(os.environ, OMP_NUM_THREADS=str(cores))
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run([str(exe), "metasurface", str(cells), "0", "0", str(X), str(Y), d],
                           capture_output=True, text=True, env=env, timeout=600)
    if r.returncode or len(r.stdout.split()) < 5:
        print(f"bench: far-field CPU baseline failed (rc {r.returncode}): {r.stdout[-200:]} {r.stderr[-400:]}",
              file=sys.stderr)
        return None
    per = float(r.stdout.split()[4])
    return {"sec_per_direction": per, "cores": cores, "kind": "reference", "face_cells": (cells + 1) ** 2,
            "sample": f"reference farField (P and S) at {X}x{Y} directions on the {cells}-cell metasurface"}


def serving(alis, args):
    """The serving load: seeded Poisson arrivals of scattering requests (grid edges 64-192
    with 8-cell CPML, a Drude-Lorentz sphere of 10-40 cells, 1000-6000 steps, 16-wavelength
    phaser), served with continuous batching; latency = finish - arrival."""
    sys.path.insert(0, str(ROOT))
    from alis_amd import serve
    os.environ["ALIS_PRECISION"] = "fp32"
    reqs = serve.requests(args.serve_seed, args.serve, args.serve_rate, alis)
    with quiet():
        res = serve.run(reqs, args.serve_budget_gib * 2**30, args.serve_slots, args.serve_tick, alis)
    s = serve.summary(reqs, res, serve.last_stats(alis))
    # the accuracy the latencies are quoted at: the request with the fewest cell-steps served
    # again alone in FP64 (the FP64 path agrees with the FP64 oracle to 1e-12), against its
    # spectrum from the mixed FP32-field run
    q = min(range(len(reqs)), key=lambda a: res[a].cellSteps)
    os.environ["ALIS_PRECISION"] = "fp64"
    with quiet():
        ref = serve.run(serve.make([(0.0, reqs[q].cells, reqs[q].steps, reqs[q].radius, reqs[q].plasma)]),
                        args.serve_budget_gib * 2**30, 1, args.serve_tick, alis)[0]
    os.environ["ALIS_PRECISION"] = "fp32"
    a, b = list(res[q].spectrum), list(ref.spectrum)
    spec_err = max(abs(x - y) for x, y in zip(a, b)) / max(abs(y) for y in b)
    # the same requests solved one at a time (all there at the start, one slot): the rate of
    # this load as single worlds, beside which the served rate stands
    solo_rate = None
    if args.serve_solo:
        solo = serve.make([(0.0, r.cells, r.steps, r.radius, r.plasma) for r in reqs])
        with quiet():
            solo_res = serve.run(solo, args.serve_budget_gib * 2**30, 1, args.serve_tick, alis)
        solo_rate = serve.summary(solo, solo_res)["mcell_updates_per_s"]
    return {
        "workload": f"{args.serve} requests, Poisson arrivals {args.serve_rate:g}/s (seed {args.serve_seed}), grid edge "
                    "{64,96,128,192} incl. 8-cell CPML, Drude-Lorentz sphere r 10-40 cells, TF/SF plane wave "
                    "400-900 nm, 1000-6000 steps, 16-wavelength phaser; FP32 fields / FP64 DFT",
        "slots": args.serve_slots, "steps_per_tick": args.serve_tick, "budget_gib": args.serve_budget_gib,
        "p50_latency_s": round(s["p50_latency_s"], 3), "p99_latency_s": round(s["p99_latency_s"], 3),
        "p50_queue_s": round(s["p50_queue_s"], 3), "p50_build_s": round(s["p50_build_s"], 3),
        "makespan_s": round(s["makespan_s"], 3),
        "stepping_fraction": round(s["stepping_fraction"], 3), "setup_thread_s": round(s["setup_s"], 2),
        "spectrum_rel_err_vs_fp64": spec_err,
        "accuracy_request": f"{reqs[q].cells}^3, {reqs[q].steps} steps, r {reqs[q].radius:g} cells",
        "value": round(s["mcell_updates_per_s"], 1), "unit": "Mcell-updates/s",
        "one_at_a_time": None if solo_rate is None else {
            "value": round(solo_rate, 1), "unit": "Mcell-updates/s",
            "served_over_one_at_a_time": round(s["mcell_updates_per_s"] / solo_rate, 2),
            "sample": "the same requests, all arriving at the start, served with one slot"},
        "max_batch": s["max_batch"], "max_request_gib": round(s["peak_request_bytes"] / 2**30, 3),
        "offered_gcell_updates": round(sum(r.cellSteps for r in res) / 1e9, 1),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cells", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=300, help="steps per CPU baseline run (median of 5 runs)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--kpoints", type=int, default=64, help="worlds of the batched k-point sweep (0: skip)")
    ap.add_argument("--serve", type=int, default=24, help="requests of the serving run (0: skip)")
    ap.add_argument("--metasurface", type=int, default=640, help="x/y period of the metasurface run in cells (0: skip)")
    ap.add_argument("--ff-angles", type=int, default=181, help="far-field map of N x N directions")
    ap.add_argument("--only", choices=["kpoints", "serving", "metasurface", "single"],
                    help="run this section alone and print its object (for profiling)")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="mean arrivals per second (Poisson)")
    ap.add_argument("--serve-seed", type=int, default=2024)
    ap.add_argument("--serve-slots", type=int, default=8, help="most requests on the device at once")
    ap.add_argument("--serve-tick", type=int, default=50, help="most steps between admissions")
    ap.add_argument("--serve-budget-gib", type=float, default=64.0, help="device memory for admitted requests")
    ap.add_argument("--serve-solo", type=int, default=1, help="also solve the requests one at a time (0: skip)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first copy this many MiB device to device (a copy of known size that "
                         "calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(independent worlds would shard with no exchange; see the design notes)", file=sys.stderr)
        sys.exit(2)
    alis, w, hip = load()
    if args.copy_mib:
        calibration_copy(hip, args.copy_mib << 20)
    if args.only:
        # one section alone (for profiling its kernels): its object as the JSON line
        out = {"kpoints": lambda: batched(alis, w, hip, args), "serving": lambda: serving(alis, args),
               "metasurface": lambda: metasurface(alis, w, hip, args),
               "single": lambda: single_fp64(alis, w, hip, args)}[args.only]()
        print(json.dumps({"section": args.only, **out}), flush=True)
        return
    with quiet():
        r32 = run(alis, w, hip, "nanoblock", args.cells, 32, 64, args.warmup, args.steps, flux=True)
    value = r32["cells"] * args.steps / r32["sec"] / 1e6
    rl = roofline(r32, 4, 8, args.steps)
    rl["traffic"] = measured_traffic(r32["cells"], "fp32")
    line = {
        "metric": "Mcell-updates/s & %HBM peak (FP32 vs FP64 on GPU); p50/p99 solve latency",
        "value": round(value, 1), "unit": "Mcell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(r32["sec"] / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (nanoblock scattering set up through alis.h; no external data)",
        "config": {"workload": f"nanoblock {args.cells}^3 (incl. 16-cell CPML), dx 2 nm, Au Drude+2 Lorentz "
                               f"60x60x40 nm block on n=1.5 substrate, TF/SF 400-900 nm band, 32-frequency phaser",
                   "cells": r32["cells"], "fields": "fp32", "dft": "fp64", "hbm_resident_bytes": r32["hbm_bytes"]},
        "roofline": rl,
        "flux": r32["flux"],
    }
    if not args.no_fp64:
        with quiet():
            r64 = run(alis, w, hip, "nanoblock", args.cells, 64, 64, args.warmup, max(args.steps // 2, 1))
        v64 = r64["cells"] * max(args.steps // 2, 1) / r64["sec"] / 1e6
        line["fp64"] = {"value": round(v64, 1), "ms_per_step": round(r64["sec"] / max(args.steps // 2, 1) * 1e3, 4),
                        "roofline_frac": roofline(r64, 8, 8, max(args.steps // 2, 1))["frac"], "fp32_over_fp64": round(value / v64, 2)}
    line["single_fp64"] = single_fp64(alis, w, hip, args)
    if args.kpoints > 0:
        line["batched"] = batched(alis, w, hip, args)
        line["batched"]["batched_fp32_over_single_fp64"] = round(line["batched"]["value"] / line["single_fp64"]["value"], 2)
    if args.serve > 0:
        line["serving"] = serving(alis, args)
    if args.metasurface > 0:
        line["metasurface"] = metasurface(alis, w, hip, args)
    line["cpu_baseline"] = cpu_baseline(args.cells, args.cpu_steps) if args.cpu_steps > 0 else None
    print(json.dumps(line), flush=True)


def single_fp64(alis, w, hip, args):
    """The plain FP64 single solve: the 64^3 dipole in vacuum (10-cell CPML, Ez point dipole,
    8 probes reading Ez and Hy every step), FP64 fields and DFT, one world per launch."""
    steps = max(args.steps, 200)
    with quiet():
        r = run(alis, w, hip, "dipole", 64, 64, 64, args.warmup, steps)
    return {"workload": "dipole 64^3 (incl. 10-cell CPML), dx 10 nm, Courant 0.5, Ez dipole 600 nm pulse, 8 probes (Ez, Hy) per step",
            "value": round(r["cells"] * steps / r["sec"] / 1e6, 1), "unit": "Mcell-updates/s", "dtype": "f64",
            "ms_per_step": round(r["sec"] / steps * 1e3, 4), "steps": steps}


def batched(alis, w, hip, args):
    """The k-point band sweep: all worlds in shared launches, with 8 probes per world read
    every step into a 256-bin FP64 probe spectrum (writeSpectrum); at least 512 steps, so the
    device adds the staged spectrum samples twice in the timed region."""
    ksteps = max(args.steps // 2, 512)
    with quiet():
        rb = run(alis, w, hip, "kpoints", args.kpoints, 32, 64, args.warmup, ksteps)
    # batched launches read the 1/eps arrays: once for both parts of a Bloch world when a
    # thread updates both (pair entries, the default; ALIS_BATCH_PAIRS=0 separates them)
    pairs = os.environ.get("ALIS_BATCH_PAIRS") != "0"
    rb["eps_bytes"] = 3 * 4 / (rb["parts"] if pairs else 1) if rb["eps"] else 0
    vb = rb["cells"] * ksteps / rb["sec"] / 1e6
    rlb = roofline(rb, 4, 8, ksteps)
    return {
        "workload": f"{args.kpoints} k-point worlds along Gamma-X-M, each 65x65x50 (64x64x31 cells + 8-cell CPML in z), "
                    "Bloch-periodic x/y, complex fields, eps=12 PhC slab with air holes, 8 random-phase dipoles, "
                    "8 probes per world read every step into a 256-bin FP64 writeSpectrum",
        "value": round(vb, 1), "unit": "Mcell-updates/s", "dtype": "f32", "ms_per_step": round(rb["sec"] / ksteps * 1e3, 4),
        "steps": ksteps,
        "roofline": {"bound": "hbm", "achieved": rlb["achieved"], "peak": rlb["peak"], "unit": "GB/s", "frac": rlb["frac"],
                     "kernel": "k_curl_batch (E half step of all worlds, both parts)", "bytes_per_launch": rlb["bytes_per_launch"],
                     "launch_ms": rlb["launch_ms"]},
    }


def metasurface(alis, w, hip, args):
    """The large periodic metasurface: step throughput and the far-field map of its phaser."""
    msteps = max(args.steps // 5, 1)
    X, Y = args.ff_angles, args.ff_angles
    with quiet():
        rm = run(alis, w, hip, "metasurface", args.metasurface, 32, 64, args.warmup, msteps, farfield=(X, Y))
    vm = rm["cells"] * msteps / rm["sec"] / 1e6
    rlm = roofline(rm, 4, 8, msteps)
    ms = {
        "workload": f"metasurface: {args.metasurface // 80}x{args.metasurface // 80} TiO2 pillars (eps 6.25, r 60-120 nm "
                    "seeded, 600 nm) on glass, dx 5 nm, periodic x/y, 16-cell CPML z, TF/SF plane wave 400-900 nm, "
                    f"{rm['num'][0]}x{rm['num'][1]}x{rm['num'][2]} nodes; one 550 nm phaser on the z faces",
        "value": round(vm, 1), "unit": "Mcell-updates/s", "dtype": "f32", "ms_per_step": round(rm["sec"] / msteps * 1e3, 4),
        "steps": msteps, "roofline": {k: rlm[k] for k in ("bound", "achieved", "peak", "unit", "frac", "kernel",
                                                        "bytes_per_launch", "launch_ms")},
        "farfield": {"directions": X * Y, "map": f"{X}x{Y} cylindrical (theta, phi), P and S", "dtype": "f64",
                     "sec": round(rm["farfield_sec"], 4), "directions_per_s": round(X * Y / rm["farfield_sec"], 1),
                     "sec_with_text_images": round(rm["farfield_files_sec"], 4)},
    }
    fc = farfield_cpu_baseline(X, 8) if args.cpu_steps > 0 else None
    if fc:
        scale = rm["num"][0] * rm["num"][1] / fc["face_cells"]
        fc["extrapolated_sec_full"] = round(fc["sec_per_direction"] * scale * X * Y, 1)
        fc["extrapolation"] = (f"measured per direction on {fc['face_cells']} face cells, scaled by face cells to "
                               f"{rm['num'][0]}x{rm['num'][1]} and by {X}x{Y} directions (not measured at full size)")
        ms["farfield"]["cpu_baseline"] = fc
    return ms


if __name__ == "__main__":
    main()
