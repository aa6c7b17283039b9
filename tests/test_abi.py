"""The C-ABI library loads on the CPU host and exports every symbol that
include/fastmc_amd.h declares (no compute calls without a GPU)."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT


def _declared():
    src = open(os.path.join(ROOT, "include", "fastmc_amd.h")).read()
    return sorted(set(re.findall(r"^int\s+(fmc_\w+)\s*\(", src, re.M)))


def test_header_declares_entry_points():
    names = _declared()
    for n in ("fmc_ctx_create", "fmc_trial", "fmc_commit", "fmc_recip_full",
              "fmc_energy_full", "fmc_batch_run", "fmc_tail_delta"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from fastmc_amd import capi
    if not os.path.exists(capi.LIB_PATH):
        pytest.fail(f"{capi.LIB_PATH} missing: build first")
    L = C.CDLL(capi.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding covers the same set
    assert set(capi.EXPORTED) == set(_declared())


def test_version_without_gpu():
    from fastmc_amd import capi
    L = capi.lib()
    a, b = C.c_int(), C.c_int()
    assert L.fmc_version(C.byref(a), C.byref(b)) == 0
    assert (a.value, b.value) == (1, 0)
