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


def test_struct_layout_matches_ctypes_mirror():
    """psoci_dav_params / psoci_dav_stats as a C99 compiler lays them out
    (psoci_amd/cli/abi_layout, built against include/psoci.h) equal the
    ctypes mirror's sizes and field offsets."""
    probe = os.path.join(os.path.dirname(api.LIB_PATH), "cli", "abi_layout")
    if not os.path.exists(probe):
        pytest.skip("psoci_amd/cli/abi_layout not built")
    out = subprocess.run([probe], capture_output=True, text=True, check=True).stdout
    c = dict(line.rsplit(" ", 1) for line in out.splitlines())
    for struct, mirror in (("psoci_dav_params", api.DavParams), ("psoci_dav_stats", api.DavStats)):
        assert int(c[struct]) == ctypes.sizeof(mirror), struct
        names = [f for f, _ in mirror._fields_]
        assert sorted(names) == sorted(k.split(".", 1)[1] for k in c if k.startswith(struct + "."))
        for f in names:
            assert int(c[f"{struct}.{f}"]) == getattr(mirror, f).offset, (struct, f)


def test_cli_driver_builds_and_answers_help():
    """The C99 command-line driver links against the library and parses its
    options without touching a device."""
    exe = os.path.join(os.path.dirname(api.LIB_PATH), "cli", "psoci_run")
    if not os.path.exists(exe):
        pytest.skip("psoci_amd/cli/psoci_run not built")
    r = subprocess.run([exe, "-h"], capture_output=True, text=True)
    assert r.returncode == 0 and "usage: psoci_run" in r.stderr
    r = subprocess.run([exe, "-n", "2"], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr
