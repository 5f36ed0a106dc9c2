"""invdft_amd -- MI355X-native inverse-DFT hot path (eigensolve, adjoint MINRES, force).

``load()`` returns the bound HIP library (invdft_amd/_lib/libidft.so).  There
is no fallback: if the library is missing or cannot be loaded the call
raises, so nothing silently runs on a CPU path.
"""
import ctypes
import os

from . import abi

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libidft.so")

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_LOCAL)
        abi.bind(lib)
        lib.idft_hip_init.restype = ctypes.c_int
        lib.idft_hip_init.argtypes = []
        rc = lib.idft_hip_init()
        if rc != 0:
            raise RuntimeError(f"libidft: no usable gfx950 device (code {rc})")
        _lib = lib
    return _lib
