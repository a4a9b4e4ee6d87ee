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
            os.path.join(ROOT, "psoci_amd", "csrc", "common.hpp"), os.path.join(ROOT, "psoci_amd", "csrc", "detdrive.hpp"),
            os.path.join(ROOT, "psoci_amd", "csrc", "smalleig.hpp")]
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
    L.hc_kind.argtypes = [P, P, P, ctypes.c_int64, ctypes.c_int64]
    L.hc_block_x.argtypes = [P, P, P, ctypes.c_int64, ctypes.c_int64, P]
    L.hc_connected.argtypes = [P, ctypes.c_int64, ctypes.c_int64]
    L.hc_row.argtypes = [P, P, P, ctypes.c_int64, ctypes.c_int, P, P, ctypes.c_int]
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _compare(L, cfg, ints, rows, both=False):
    """Blocks (I, J) for J <= I (and J > I too with both=True) bit for bit."""
    sp = oracle.OracleSpace(cfg, ints)
    mixed = len(set(cfg.jsym.tolist())) > 1
    h = L.hc_space_new(cfg.nbf, cfg.nelec, cfg.ksym, cfg.ncfg, _p(cfg.occ), _p(cfg.jsym))
    assert h
    nblocks = 0
    try:
        for I in rows:
            I = int(I)
            jl = set(sp.jlist(I).tolist())
            for J in range(cfg.ncfg if both else I + 1):
                con = L.hc_connected(h, I, J)
                if J <= I:
                    assert bool(con) == (J in jl), (I, J)
                if not con:
                    continue
                st, ref = sp.block(I, J)
                out = np.zeros_like(ref)
                st2 = L.hc_block(h, _p(ints.valone), _p(ints.valtwo), I, J, _p(out))
                assert st == st2, (I, J)
                assert np.array_equal(ref, out), (I, J, np.abs(ref - out).max())
                if I != J and not mixed:
                    # the three-index-table path used by the direct sigma
                    # (one jsym: spaces with spin-orbit blocks run stored)
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


def _rows_agree(L, cfg, ints, rows, k=3):
    """Determinant-driven sigma rows (direct kernel's enumeration) against SEF
    pair-by-pair rows built from the bit-identical elements above."""
    h = L.hc_space_new(cfg.nbf, cfg.nelec, cfg.ksym, cfg.ncfg, _p(cfg.occ), _p(cfg.jsym))
    assert h
    try:
        sp = oracle.OracleSpace(cfg, ints)
        X = np.random.default_rng(5).standard_normal((sp.n, k))
        for I in rows:
            I = int(I)
            ni = int(sp.nsef[I])
            y0 = np.zeros((ni, k))
            y1 = np.zeros((ni, k))
            assert L.hc_row(h, _p(ints.valone), _p(ints.valtwo), I, k, _p(X), _p(y0), 0) == 0
            assert L.hc_row(h, _p(ints.valone), _p(ints.valtwo), I, k, _p(X), _p(y1), 1) == 0
            scale = max(1.0, np.abs(y0).max())
            assert np.abs(y1 - y0).max() <= 1e-13 * scale, (I, np.abs(y1 - y0).max())
    finally:
        L.hc_space_free(h)


def test_determinant_driven_rows_fci6(hc, fci6):
    cfg, ints, _ = fci6
    _rows_agree(hc, cfg, ints, range(cfg.ncfg))


def test_determinant_driven_rows_opposite_sef_sign(hc, fci6):
    """SEFs (D_2k - D_2k+1)/sqrt(2): configuration and state symmetry of
    opposite parity (ksym 2 over jsym 1)."""
    cfg, ints, _ = fci6
    cfg = inputs.fci_configs(cfg.nbf, cfg.nelec, jsym=1, ksym=2)
    _rows_agree(hc, cfg, ints, range(cfg.ncfg))


@pytest.mark.parametrize("ksym", [3, 4])
def test_odd_class_blocks_and_rows(hc, fci6, ksym):
    """ODD-class spaces (state symmetry class differs from the
    configurations'): odd-Ms determinants, SEF sign + (ksym 3) or - (ksym 4);
    blocks bit-identical to the oracle, determinant-driven rows equal to
    SEF-pair rows."""
    cfg, ints, _ = fci6
    cfg = inputs.fci_configs(cfg.nbf, cfg.nelec, jsym=1, ksym=ksym)
    _compare(hc, cfg, ints, range(cfg.ncfg))
    _rows_agree(hc, cfg, ints, range(cfg.ncfg))


@pytest.mark.parametrize("jsym", [1, 3])
def test_fractional_blocks_and_rows(hc, jsym):
    """Odd electron count (one determinant per SEF; jsym 3 swaps alpha and
    beta): blocks bit-identical to the oracle, determinant-driven rows equal
    to SEF-pair rows."""
    cfg = inputs.fci_configs(6, 5, jsym=jsym, ksym=jsym)
    ints = inputs.make_integrals(6, 424242)
    _compare(hc, cfg, ints, range(cfg.ncfg))
    _rows_agree(hc, cfg, ints, range(cfg.ncfg))
    cfg = inputs.mrcisd_configs(4, 6, 5, jsym=jsym, ksym=jsym)
    ints = inputs.make_integrals(10, 77, n_internal=4)
    rows = np.random.default_rng(jsym).choice(cfg.ncfg, 30, replace=False)
    _compare(hc, cfg, ints, rows)
    _rows_agree(hc, cfg, ints, rows)


def test_determinant_driven_rows_fci10_mrci(hc, fci10):
    cfg, ints, _ = fci10
    nopen = (cfg.occ == 1).sum(1)
    rng = np.random.default_rng(12)
    rows = list(rng.choice(cfg.ncfg, 25, replace=False)) + list(np.where(nopen == 8)[0][:3]) + [0, cfg.ncfg - 1]
    _rows_agree(hc, cfg, ints, rows)
    cfg = inputs.mrcisd_configs(4, 6, 4)
    ints = inputs.make_integrals(10, 99, n_internal=4)
    _rows_agree(hc, cfg, ints, np.random.default_rng(4).choice(cfg.ncfg, 30, replace=False))


def test_subspace_eigensolver_vs_lapack(hc):
    """The Davidson subspace eigensolver (Householder + implicit QR) against
    LAPACK, including degenerate and clustered spectra."""
    hc.hc_sym_eig.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    rng = np.random.default_rng(8)
    cases = []
    for m in (1, 2, 3, 7, 16, 48, 80, 160):
        A = rng.standard_normal((m, m))
        cases.append((A + A.T) / 2)
    # degenerate pairs and a cluster
    Q, _ = np.linalg.qr(rng.standard_normal((40, 40)))
    d = np.concatenate([np.repeat([1.0, 2.0, 3.0], 2), 5 + 1e-9 * np.arange(6), np.linspace(6, 9, 28)])
    cases.append(Q @ np.diag(d) @ Q.T)
    cases.append(np.diag(rng.standard_normal(12)))
    for A in cases:
        m = A.shape[0]
        a = np.ascontiguousarray(A.copy())
        w = np.zeros(m)
        assert hc.hc_sym_eig(m, _p(a), _p(w)) == 1
        ref = np.linalg.eigvalsh(A)
        scale = max(1.0, np.abs(ref).max())
        assert np.abs(w - ref).max() <= 1e-13 * scale * m, m
        V = a
        assert np.abs(V.T @ V - np.eye(m)).max() < 1e-12 * m
        assert np.abs(A @ V - V * w).max() < 1e-12 * scale * m


def test_top_orbitals_nbf64_equal_the_small_problem(hc, fci6):
    """The fci6 problem moved to orbitals 0, 1, 2, 61, 62, 63 of a 64-orbital
    basis (the top bits of the occupation masks; the other integrals random
    and unreachable): every block bit-identical between product and oracle,
    and bit-identical to the 6-orbital blocks, since orbital order, and with
    it every phase and summation order, is unchanged."""
    cfg, ints, _ = fci6
    cfg64, ints64 = inputs.embed(cfg, ints, [0, 1, 2, 61, 62, 63])
    _compare(hc, cfg64, ints64, range(cfg64.ncfg))
    _rows_agree(hc, cfg64, ints64, range(0, cfg64.ncfg, 7))
    small = oracle.OracleSpace(cfg, ints)
    big = oracle.OracleSpace(cfg64, ints64)
    for I in range(0, cfg.ncfg, 5):
        for J in range(I + 1):
            st, a = small.block(I, J)
            st64, b = big.block(I, J)
            assert st == st64 and np.array_equal(a, b), (I, J)


@pytest.mark.parametrize("nelec,ksym", [(10, 1), (10, 3), (9, 1)])
def test_nine_and_ten_open_shells(hc, nelec, ksym):
    """Configurations of 10 orbitals with 7 to 10 open shells (64 to 256
    SEFs, the sizes the direct sigma splits over wavefronts), EVEN, ODD and
    one-determinant-per-SEF classes: sampled blocks bit-identical to the
    oracle, determinant-driven rows equal to SEF-pair rows."""
    full = inputs.fci_configs(10, nelec, jsym=1, ksym=ksym)
    keep = (full.occ == 1).sum(1) >= 7
    cfg = inputs.Configs(nbf=10, nelec=nelec, ksym=ksym, occ=full.occ[keep], jsym=full.jsym[keep])
    ints = inputs.make_integrals(10, 910 + nelec + ksym)
    nopen = (cfg.occ == 1).sum(1)
    rows = list(np.where(nopen == nopen.max())[0][:2]) + list(np.random.default_rng(ksym).choice(cfg.ncfg, 4))
    _compare(hc, cfg, ints, rows)
    _rows_agree(hc, cfg, ints, rows)


def _mixed_space(nbf, nelec, jsyms, seed=5):
    """Every FCI configuration once per jsym value (distinct configurations
    of the reference's list), so configurations of different jsym are one
    electron apart: spin-orbit blocks."""
    base = inputs.fci_configs(nbf, nelec)
    occ = np.concatenate([base.occ] * len(jsyms))
    jsym = np.concatenate([np.full(base.ncfg, j, np.int32) for j in jsyms])
    cfg = inputs.Configs(nbf=nbf, nelec=nelec, ksym=1, occ=occ, jsym=jsym, title="mixed jsym")
    return cfg, inputs.make_integrals(nbf, seed)


@pytest.mark.parametrize("nbf,nelec,jsyms", [(4, 4, (1, 2)), (4, 4, (1, 3)), (4, 4, (2, 3, 4)), (5, 4, (1, 3)),
                                             (4, 3, (1, 3)), (5, 5, (2, 3))])
def test_spin_orbit_blocks_bit_identical(hc, nbf, nelec, jsyms):
    """Spaces with configurations of several jsym (EVEN and ODD classes,
    both SEF signs; an odd electron count: both fractional orientations):
    every block, in both orientations, spin-orbit singles included, equals
    the oracle's restatement of generateSOCIblocks/socig bit for bit."""
    cfg, ints = _mixed_space(nbf, nelec, jsyms)
    n = _compare(hc, cfg, ints, range(cfg.ncfg), both=True)
    assert n > 100
