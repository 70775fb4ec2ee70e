"""In-tree build of the engine library, the CPU oracle and (when its sources are present) the
reference's own C code for the oracle checks.

Products (all git-ignored, shipped to GPU boxes with the tree):
  alis_amd/libalis.so              host library (alis.h API, C99) + HIP engine for gfx950
  alis_amd/libalis_workloads.so    benchmark workloads (alis_amd/workloads.c) over libalis.so
  tests/golden/libdrive.so         parity scenarios (tests/golden/drive.c) over libalis.so
  oracle/liboracle.so              CPU restatement of the engine ABI under the host library
  tests/golden/libdrive_oracle.so  the same scenarios over liboracle.so
  oracle/libwl_oracle.so           the benchmark workloads over liboracle.so (parity at full size)
  oracle/_ref/                     the reference built from its sources (oracle/Makefile)
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "alis_amd"
BUILD = ROOT / "_build"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
ARCH = os.environ.get("ALIS_OFFLOAD_ARCH", "gfx950")

# Host C: the reference's semantics need IEEE float without contraction or fast-math and
# the x86-64 baseline ISA (no FMA), exactly like its own gcc -O3 -std=c99 build.
CFLAGS = ["-O3", "-std=gnu99", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
          "-Wall", "-Wpointer-arith", "-Wno-parentheses", "-Wno-unused-function", f"-I{ROOT/'include'}"]
# HIP: no contraction (bitwise FP32 parity), IEEE denormals, correctly rounded divide/sqrt,
# packed-FP32 instructions disabled.
HIPFLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
            "-fno-gpu-flush-denormals-to-zero", "-fhip-fp32-correctly-rounded-divide-sqrt",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-result", f"-I{ROOT/'include'}"]


def run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode:
        sys.stderr.write(" ".join(map(str, cmd)) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} {cmd[-1]}")
    return r


def build_library():
    """Compile the host library and the HIP engine and link alis_amd/libalis.so."""
    BUILD.mkdir(exist_ok=True)
    jobs = []
    for src in sorted((PKG / "host").glob("*.c")):
        obj = BUILD / (src.stem + ".o")
        jobs.append(["gcc", *CFLAGS, "-c", str(src), "-o", str(obj)])
    for src in sorted((PKG / "csrc").glob("*.hip")):
        obj = BUILD / (src.stem + ".hip.o")
        jobs.append([str(ROCM / "bin" / "hipcc"), *HIPFLAGS, "-c", str(src), "-o", str(obj)])
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as pool:
        list(pool.map(run, jobs))
    objs = [j[-1] for j in jobs]
    out = PKG / "libalis.so"
    tmp = PKG / "libalis.so.tmp"
    # -static-libgcc: complex division (__divsc3) from the compiler's own libgcc.a, the one
    # a program built from the reference's sources with gcc links; the system libgcc_s may
    # be newer and round float complex quotients differently.  -Bsymbolic: the library's
    # own calls bind inside it, also when another alis.h implementation is loaded.
    run(["g++", "-shared", "-static-libgcc", "-Wl,-Bsymbolic", "-o", str(tmp), *objs, "-fopenmp", "-lz", f"-L{ROCM/'lib'}", "-lamdhip64",
         f"-Wl,-rpath,{ROCM/'lib'}", "-Wl,--no-undefined"])
    os.replace(tmp, out)
    clients(PKG, "libalis.so", "alis")
    return out


def client(src, out, libdir, libname, defines=()):
    """A C file over one alis.h implementation, as a shared library."""
    tmp = out.with_suffix(".so.tmp")
    run(["gcc", *CFLAGS, *defines, "-shared", "-o", str(tmp), str(src), f"-L{libdir}", f"-l{libname}",
         f"-Wl,-rpath,$ORIGIN/{os.path.relpath(libdir, out.parent)}", "-lm"])
    os.replace(tmp, out)


def clients(libdir, soname, libname):
    if libname == "alis":
        client(PKG / "workloads.c", PKG / "libalis_workloads.so", libdir, libname)
        client(ROOT / "tests" / "golden" / "drive.c", ROOT / "tests" / "golden" / "libdrive.so", libdir, libname)
    else:
        client(ROOT / "tests" / "golden" / "drive.c", ROOT / "tests" / "golden" / "libdrive_oracle.so", libdir, libname)
        client(PKG / "workloads.c", ROOT / "oracle" / "libwl_oracle.so", libdir, libname)


def build_oracle():
    """CPU oracle (oracle/Makefile) and, when /root/reference exists, the reference build."""
    run(["make", "-s", "-C", str(ROOT / "oracle"), "liboracle.so"])
    clients(ROOT / "oracle", "liboracle.so", "oracle")
    if Path("/root/reference/src/update.c").exists():
        run(["make", "-s", "-C", str(ROOT / "oracle"), "ref", "ref-programs"])


def build(verbose=False):
    lib = build_library()
    build_oracle()
    if verbose:
        print("built", lib)
    return lib


if __name__ == "__main__":
    build(verbose=True)
