"""The string-driven direct sigma's plan and index arithmetic
(psoci_amd/csrc/ras.cpp, ras.hpp), compiled for the host CPU and evaluated
with plain loops, against the oracle's stored sigma: the determinant tables,
signs, sectors and SEF map that the GPU kernels use."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle
from psoci_amd import inputs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "psoci_amd", "csrc")
SRC = os.path.join(ROOT, "tests", "hostcheck", "host_ras.cpp")
OUT = os.path.join(ROOT, "tests", "hostcheck", "_build", "libhostras.so")


@pytest.fixture(scope="module")
def hr():
    deps = [SRC] + [os.path.join(CSRC, f) for f in ("ras.cpp", "ras.hpp", "space.cpp", "elem.hpp", "common.hpp")]
    if not os.path.exists(OUT) or os.path.getmtime(OUT) < max(os.path.getmtime(d) for d in deps):
        os.makedirs(os.path.dirname(OUT), exist_ok=True)
        if not os.path.exists("/opt/rocm/bin/hipcc"):
            pytest.skip("hipcc not available")
        subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O2", "-std=c++17", "-fPIC", "-shared",
                        "-ffp-contract=off", "-x", "hip", SRC, os.path.join(CSRC, "space.cpp"),
                        os.path.join(CSRC, "ras.cpp"), "-o", OUT], check=True)
    L = ctypes.CDLL(OUT)
    P = ctypes.c_void_p
    L.hr_plan.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64, P, P, P, P, P, ctypes.c_char_p]
    L.hr_sigma.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64, P, P, P, P, ctypes.c_int, P, P]
    L.hr_sigma_merged.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64, P, P, P, P, ctypes.c_int,
                                  P, P, P]
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _args(cfg, ints):
    occ = np.ascontiguousarray(cfg.occ, np.uint8)
    js = np.ascontiguousarray(cfg.jsym, np.int32)
    one = np.ascontiguousarray(ints.valone, np.float64)
    two = np.ascontiguousarray(ints.valtwo, np.float64)
    return occ, js, one, two


def plan(L, cfg, ints):
    occ, js, one, two = _args(cfg, ints)
    info = np.zeros(6, np.int64)
    why = ctypes.create_string_buffer(256)
    rc = L.hr_plan(cfg.nbf, cfg.nelec, cfg.ksym, cfg.ncfg, _p(occ), _p(js), _p(one), _p(two), _p(info), why)
    return rc, info, why.value.decode()


def sigma(L, cfg, ints, X):
    occ, js, one, two = _args(cfg, ints)
    X = np.ascontiguousarray(X, np.float64)
    Y = np.zeros_like(X)
    rc = L.hr_sigma(cfg.nbf, cfg.nelec, cfg.ksym, cfg.ncfg, _p(occ), _p(js), _p(one), _p(two), X.shape[1], _p(X),
                    _p(Y))
    assert rc == 0, f"host sigma failed ({rc})"
    return Y


CASES = [
    ("fci6", lambda: inputs.workload("fci6")[:2], 0, 0),
    ("fci6 mult -1", lambda: (inputs.fci_configs(6, 6, jsym=1, ksym=2), inputs.make_integrals(6, 6)), 0, 0),
    ("fci 8 orbitals 8 electrons (Ms 0, 2, 4 sectors)",
     lambda: (inputs.fci_configs(8, 8), inputs.make_integrals(8, 5)), 0, 0),
    ("fci 7 orbitals 2 electrons", lambda: (inputs.fci_configs(7, 2), inputs.make_integrals(7, 9)), 0, 0),
    ("mrcisd 4+5 / 4", lambda: (inputs.mrcisd_configs(4, 5, 4), inputs.make_integrals(9, 11, n_internal=4)), 2, 5),
    ("mrcisd 3+4 / 6", lambda: (inputs.mrcisd_configs(3, 4, 6), inputs.make_integrals(7, 12, n_internal=3)), 2, 4),
    ("mrcis 4+3 / 4", lambda: (inputs.mrcisd_configs(4, 3, 4, max_external=1),
                               inputs.make_integrals(7, 8, n_internal=4)), 1, 3),
    ("mrcisd 5+4 / 6 mult -1", lambda: (inputs.mrcisd_configs(5, 4, 6, jsym=1, ksym=2),
                                        inputs.make_integrals(9, 13, n_internal=5)), 2, 4),
]


@pytest.mark.parametrize("name,make,m,next_", CASES, ids=[c[0] for c in CASES])
def test_host_ras_sigma_matches_oracle(hr, name, make, m, next_):
    cfg, ints = make()
    rc, info, why = plan(hr, cfg, ints)
    assert rc == 1, why
    assert info[0] == m and info[1] == next_
    sp = oracle.OracleSpace(cfg, ints)
    sp.build_h()
    X = np.random.default_rng(3).standard_normal((sp.n, 3))
    ref = sp.sigma(X)
    Y = sigma(hr, cfg, ints, X)
    np.testing.assert_allclose(Y, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("name,make,m,next_", CASES, ids=[c[0] for c in CASES])
def test_host_ras_merged_terms_match_oracle(hr, name, make, m, next_):
    """The gather's term lists (ras_sef_terms): in spin-flip symmetric sectors
    the mirrored terms of an SEF's two determinants are one W element each,
    merged into one term (or cancelled); with the product's skipping of
    chunks whose pairs all add an orbital the tile's row string holds, the
    sigma is unchanged and, for mult = +1, the list is shorter than the
    determinants' terms."""
    cfg, ints = make()
    sp = oracle.OracleSpace(cfg, ints)
    sp.build_h()
    X = np.random.default_rng(4).standard_normal((sp.n, 2))
    ref = sp.sigma(X)
    occ, js, one, two = _args(cfg, ints)
    Y = np.zeros_like(X)
    nt = np.zeros(3, np.int64)
    rc = hr.hr_sigma_merged(cfg.nbf, cfg.nelec, cfg.ksym, cfg.ncfg, _p(occ), _p(js), _p(one), _p(two), X.shape[1],
                            _p(X), _p(Y), _p(nt))
    assert rc == 0
    np.testing.assert_allclose(Y, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    merged, full, skipped = int(nt[0]), int(nt[1]), int(nt[2])
    assert 0 < merged <= full
    # the product's chunk skipping (same result) is exercised where pair
    # runs are long enough (8- and 9-orbital spaces)
    if cfg.nbf >= 8 and not name.startswith("mrcisd 4+5"):
        assert skipped > 0
    if "mult -1" in name:
        assert merged == full  # no spin-flip symmetric sector: nothing merges
    elif (cfg.nopen() > 0).any():
        assert merged < full


def test_host_ras_refusals(hr):
    cfg, ints, _ = inputs.workload("fci6")
    # ODD determinant class
    odd = inputs.fci_configs(6, 6, jsym=1, ksym=3)
    rc, _, why = plan(hr, odd, ints)
    assert rc == 0 and "ODD" in why
    # odd electron count
    frac = inputs.fci_configs(6, 5)
    rc, _, why = plan(hr, frac, ints)
    assert rc == 0 and "odd electron" in why
    # a list that is not a complete RAS space: one configuration from the
    # middle dropped (dropping the last one, (0,0,0,2,2,2), leaves the RAS
    # space with orbitals 3-5 external and at most 5 electrons there)
    keep = np.ones(cfg.ncfg, bool)
    keep[cfg.ncfg // 2] = False
    part = inputs.Configs(nbf=cfg.nbf, nelec=cfg.nelec, ksym=cfg.ksym, occ=cfg.occ[keep].copy(),
                          jsym=cfg.jsym[keep].copy(), title="fci6 minus one")
    rc, _, why = plan(hr, part, ints)
    assert rc == 0 and "RAS" in why
    # a repeated configuration
    rep = inputs.Configs(nbf=cfg.nbf, nelec=cfg.nelec, ksym=cfg.ksym, occ=np.vstack([cfg.occ, cfg.occ[:1]]),
                         jsym=np.concatenate([cfg.jsym, cfg.jsym[:1]]), title="fci6 plus a repeat")
    rc, _, why = plan(hr, rep, ints)
    assert rc == 0 and "RAS" in why
    # two jsym values (spin-orbit blocks)
    js = cfg.jsym.copy()
    js[::2] = 2
    mix = inputs.Configs(nbf=cfg.nbf, nelec=cfg.nelec, ksym=cfg.ksym, occ=cfg.occ.copy(), jsym=js, title="mixed")
    rc, _, why = plan(hr, mix, ints)
    assert rc == 0 and "jsym" in why
    # the last configuration dropped: a RAS space with a different split
    last = inputs.Configs(nbf=cfg.nbf, nelec=cfg.nelec, ksym=cfg.ksym, occ=cfg.occ[:-1].copy(),
                          jsym=cfg.jsym[:-1].copy(), title="fci6 minus the last")
    rc, info, why = plan(hr, last, ints)
    assert rc == 1 and info[0] == 5 and info[1] == 3
    sp = oracle.OracleSpace(last, ints)
    sp.build_h()
    X = np.random.default_rng(5).standard_normal((sp.n, 2))
    ref = sp.sigma(X)
    np.testing.assert_allclose(sigma(hr, last, ints, X), ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def test_host_ras_plan_of_the_benchmark_space(hr):
    """mrci8x24: the plan's external set and counts (no sigma on the host)."""
    cfg, ints, _ = inputs.workload("mrci8x24")
    rc, info, why = plan(hr, cfg, ints)
    assert rc == 1, why
    assert info[0] == 2 and info[1] == 24
    # determinants of the evaluated sectors: Ms = 0 (3 alpha, 3 beta) and
    # n_alpha - n_beta = 4 (5 alpha, 1 beta)
    from math import comb
    def strings(n):
        return sum(comb(8, n - a) * comb(24, a) for a in range(0, 3) if n - a >= 0)
    s3 = [comb(8, 3 - a) * comb(24, a) for a in range(3)]
    s5 = [comb(8, 5 - a) * comb(24, a) for a in range(3)]
    s1 = [comb(8, 1 - a) * comb(24, a) if a <= 1 else 0 for a in range(3)]
    ms0 = sum(s3[a] * s3[b] for a in range(3) for b in range(3) if a + b <= 2)
    ms2 = sum(s5[a] * s1[b] for a in range(3) for b in range(3) if a + b <= 2)
    assert info[4] == ms0 + ms2
    del strings
