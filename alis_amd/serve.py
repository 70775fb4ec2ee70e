"""Python side of the serving loop (include/alis_serve.h): ctypes mirrors of its structs and
thin calls into the library, plus the latency summary.

    from alis_amd import serve
    reqs = serve.requests(seed=2024, n=32, rate=20.0)
    res = serve.run(reqs, budget=64 << 30, slots=8, steps_per_tick=50)
    serve.summary(reqs, res)     # p50/p99 latency, Mcell-updates/s, ...
"""
import ctypes

import numpy as np

from . import library


class Request(ctypes.Structure):
    _fields_ = [("arrival", ctypes.c_double), ("cells", ctypes.c_int), ("steps", ctypes.c_int),
                ("radius", ctypes.c_float), ("plasma", ctypes.c_float)]


class Result(ctypes.Structure):
    _fields_ = [("ready", ctypes.c_double), ("admitted", ctypes.c_double), ("finished", ctypes.c_double),
                ("latency", ctypes.c_double),
                ("cellSteps", ctypes.c_double), ("bytes", ctypes.c_double), ("batchMax", ctypes.c_int),
                ("wavelength", ctypes.c_float * 16), ("spectrum", ctypes.c_float * 16)]


class Config(ctypes.Structure):
    _fields_ = [("budget", ctypes.c_double), ("slots", ctypes.c_int), ("stepsPerTick", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("makespan", ctypes.c_double), ("stepping", ctypes.c_double), ("setup", ctypes.c_double),
                ("steps", ctypes.c_long)]


def _bind(lib):
    lib.alisServe.argtypes = [ctypes.POINTER(Request), ctypes.c_int, ctypes.POINTER(Config), ctypes.POINTER(Result)]
    lib.alisServe.restype = ctypes.c_int
    lib.alisServeRequests.argtypes = [ctypes.c_ulonglong, ctypes.c_int, ctypes.c_double, ctypes.POINTER(Request)]
    lib.alisServeMaxRadius.argtypes = [ctypes.c_int]
    lib.alisServeLastStats.argtypes = [ctypes.POINTER(Stats)]
    lib.alisServeLastStats.restype = None
    return lib


def requests(seed, n, rate, lib=None):
    """The seeded load of alisServeRequests as an array of Request."""
    lib = _bind(lib or library())
    R = (Request * n)()
    lib.alisServeRequests(seed, n, rate, R)
    return R


def make(specs):
    """Requests from (arrival, cells, steps, radius, plasma) tuples."""
    R = (Request * len(specs))()
    for r, s in zip(R, specs):
        r.arrival, r.cells, r.steps, r.radius, r.plasma = s
    return R


def run(reqs, budget, slots, steps_per_tick, lib=None):
    """Serve `reqs`; returns the array of Result."""
    lib = _bind(lib or library())
    out = (Result * len(reqs))()
    cfg = Config(float(budget), int(slots), int(steps_per_tick))
    rc = lib.alisServe(reqs, len(reqs), ctypes.byref(cfg), out)
    if rc:
        raise ValueError("alisServe rejected the request list or configuration")
    return out


def last_stats(lib=None):
    """Statistics of the last run: makespan, seconds stepping and building, steps."""
    lib = _bind(lib or library())
    s = Stats()
    lib.alisServeLastStats(ctypes.byref(s))
    return {"makespan_s": s.makespan, "stepping_s": s.stepping, "setup_s": s.setup, "steps": s.steps}


def summary(reqs, res, stats=None):
    lat = np.array([r.latency for r in res])
    wait = np.array([r.admitted - q.arrival for q, r in zip(reqs, res)])
    build = np.array([r.ready - q.arrival for q, r in zip(reqs, res)])
    span = max(r.finished for r in res) - min(q.arrival for q in reqs)
    cells = sum(r.cellSteps for r in res)
    extra = {}
    if stats:
        # share of the run the serving thread spent in shared steps, whose launches keep the
        # GPU busy (the rest: waiting for a request to be built, admissions, retirements)
        extra = {"stepping_fraction": stats["stepping_s"] / stats["makespan_s"], "setup_s": stats["setup_s"],
                 "shared_steps": stats["steps"]}
    return {
        **extra,
        "p50_build_s": float(np.percentile(build, 50)),
        "requests": len(res),
        "p50_latency_s": float(np.percentile(lat, 50)),
        "p99_latency_s": float(np.percentile(lat, 99)),
        "max_latency_s": float(lat.max()),
        "p50_queue_s": float(np.percentile(wait, 50)),
        "makespan_s": span,
        "mcell_updates_per_s": cells / span / 1e6,
        "max_batch": max(r.batchMax for r in res),
        "peak_request_bytes": max(r.bytes for r in res),
    }
