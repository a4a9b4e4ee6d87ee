"""The C-ABI library loads and exports every entry point include/psoci.h
declares (no compute calls: runs without a GPU)."""
import ctypes
import os
import subprocess

import pytest

from psoci_amd import api


def test_header_declares_expected_entry_points():
    names = api.header_functions()
    for n in ["psoci_ctx_create", "psoci_set_integrals", "psoci_set_configs", "psoci_space_info", "psoci_block",
              "psoci_build_h", "psoci_sigma", "psoci_davidson", "psoci_rdm", "psoci_last_error"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(api.LIB_PATH):
        pytest.skip("libpsoci_amd.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", api.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in api.header_functions() if n not in exported]
    assert not missing, missing
    L = ctypes.CDLL(api.LIB_PATH)
    for n in api.header_functions():
        assert getattr(L, n) is not None


def test_null_context_is_rejected_without_device():
    if not os.path.exists(api.LIB_PATH):
        pytest.skip("libpsoci_amd.so not built")
    L = api.load()
    assert L.psoci_build_h(None, None) == api.ERR_ARG
    assert L.psoci_sync(None) == api.ERR_ARG
    assert L.psoci_last_error(None) == b"null context"
