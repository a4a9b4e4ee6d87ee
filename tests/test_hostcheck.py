"""The product's own space builder and bit-mask element arithmetic
(psoci_amd/csrc/space.cpp, elem.hpp), compiled for the host CPU, against
the oracle: same determinants, same pair screening and bit-identical blocks.
The GPU runs exactly this arithmetic (the same inline functions)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle
from psoci_amd import inputs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "hostcheck", "host_elem.cpp")
OUT = os.path.join(ROOT, "tests", "hostcheck", "_build", "libhostcheck.so")


@pytest.fixture(scope="module")
def hc():
    deps = [SRC, os.path.join(ROOT, "psoci_amd", "csrc", "space.cpp"), os.path.join(ROOT, "psoci_amd", "csrc", "elem.hpp"),
            os.path.join(ROOT, "psoci_amd", "csrc", "common.hpp")]
    if not os.path.exists(OUT) or os.path.getmtime(OUT) < max(os.path.getmtime(d) for d in deps):
        os.makedirs(os.path.dirname(OUT), exist_ok=True)
        if not os.path.exists("/opt/rocm/bin/hipcc"):
            pytest.skip("hipcc not available")
        subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O2", "-std=c++17", "-fPIC", "-shared",
                        "-ffp-contract=off", "-x", "hip", SRC, os.path.join(ROOT, "psoci_amd", "csrc", "space.cpp"),
                        "-o", OUT], check=True)
    L = ctypes.CDLL(OUT)
    P = ctypes.c_void_p
    L.hc_space_new.restype = P
    L.hc_space_new.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64, P, P]
    L.hc_space_free.argtypes = [P]
    L.hc_block.argtypes = [P, P, P, ctypes.c_int64, ctypes.c_int64, P]
    L.hc_block_x.argtypes = [P, P, P, ctypes.c_int64, ctypes.c_int64, P]
    L.hc_connected.argtypes = [P, ctypes.c_int64, ctypes.c_int64]
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _compare(L, cfg, ints, rows):
    sp = oracle.OracleSpace(cfg, ints)
    h = L.hc_space_new(cfg.nbf, cfg.nelec, cfg.ksym, cfg.ncfg, _p(cfg.occ), _p(cfg.jsym))
    assert h
    nblocks = 0
    try:
        for I in rows:
            I = int(I)
            jl = set(sp.jlist(I).tolist())
            for J in range(I + 1):
                con = L.hc_connected(h, I, J)
                assert bool(con) == (J in jl), (I, J)
                if not con:
                    continue
                st, ref = sp.block(I, J)
                out = np.zeros_like(ref)
                st2 = L.hc_block(h, _p(ints.valone), _p(ints.valtwo), I, J, _p(out))
                assert st == st2, (I, J)
                assert np.array_equal(ref, out), (I, J, np.abs(ref - out).max())
                if I != J:
                    # the load-free path used by the direct sigma
                    out_x = np.zeros_like(ref)
                    st3 = L.hc_block_x(h, _p(ints.valone), _p(ints.valtwo), I, J, _p(out_x))
                    assert st3 == st, (I, J)
                    assert np.array_equal(ref, out_x), (I, J, np.abs(ref - out_x).max())
                nblocks += 1
    finally:
        L.hc_space_free(h)
    return nblocks


def test_fci6_all_blocks_bit_identical(hc, fci6):
    cfg, ints, _ = fci6
    assert _compare(hc, cfg, ints, range(cfg.ncfg)) == 4656


def test_fci10_sampled_blocks_bit_identical(hc, fci10):
    cfg, ints, _ = fci10
    rng = np.random.default_rng(11)
    rows = list(rng.choice(cfg.ncfg, 40, replace=False)) + [cfg.ncfg - 1]
    assert _compare(hc, cfg, ints, rows) > 1000


def test_mrci_sampled_blocks_bit_identical(hc):
    cfg = inputs.mrcisd_configs(4, 6, 4)
    ints = inputs.make_integrals(10, 99, n_internal=4)
    rng = np.random.default_rng(3)
    rows = rng.choice(cfg.ncfg, 25, replace=False)
    assert _compare(hc, cfg, ints, rows) > 100
