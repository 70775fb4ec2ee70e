"""alis_amd: an MI355X FDTD engine behind the reference's C API (include/alis.h).

The product is the shared library alis_amd/libalis.so: the host library (alis_amd/host,
C99) that keeps the reference's API and setup semantics, over the HIP engine
(alis_amd/csrc/engine.hip) reached through the C ABI of include/alis_dev.h.  C programs
link it like the reference's libalis.a; Python reaches it with ctypes.

    import alis_amd
    lib = alis_amd.library()          # ctypes handle, raises if it is not built
"""
import ctypes
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIBRARY = PKG / "libalis.so"
WORKLOADS = PKG / "libalis_workloads.so"

_lib = None


def library():
    """The engine library (ctypes.CDLL).  There is no CPU fallback: without the built HIP
    library this raises, and without a GPU every time step stops with an error."""
    global _lib
    if _lib is None:
        if not LIBRARY.exists():
            raise RuntimeError(f"{LIBRARY} is not built; run __graft_entry__.build()")
        _lib = ctypes.CDLL(str(LIBRARY))
        _lib.alis_dev_error.restype = ctypes.c_char_p
    return _lib
