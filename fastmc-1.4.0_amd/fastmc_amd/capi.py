"""ctypes binding of include/fastmc_amd.h (filled in with the HIP library)."""
from __future__ import annotations

import ctypes as C
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libfastmc_amd.so")


class FmcError(RuntimeError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FmcError(f"HIP library not built: {LIB_PATH}; run "
                           "__graft_entry__.build() (no CPU fallback)")
        _lib = C.CDLL(LIB_PATH)
    return _lib
