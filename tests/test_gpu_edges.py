"""Edge cases of the configuration space on the GPU through the C ABI: the
top bits of the 64-orbital occupation masks, the largest open-shell counts
(9 and 10: configurations of 256 SEFs, which the direct sigma splits over
wavefronts), one-configuration spaces, and an empty list."""
import numpy as np
import pytest

import oracle
from psoci_amd import api, inputs

pytestmark = pytest.mark.gpu


def test_top_orbitals_equal_the_small_problem(fci6):
    """fci6 moved to the top orbitals of a 64-orbital basis (0, 1, 2, 61, 62,
    63; other integrals random and unreachable): the stored H is
    bit-identical to the 6-orbital one, the direct sigma (whose pair records
    keep orbitals in 7-bit fields) equals the 6-orbital stored sigma, and
    the eigenvalues are the 6-orbital ones."""
    cfg, ints, _ = fci6
    A = api.Hamiltonian(cfg, ints)
    nnz = A.build()
    cfg64, ints64 = inputs.embed(cfg, ints, [0, 1, 2, 61, 62, 63])
    B = api.Hamiltonian(cfg64, ints64)
    assert B.build() == nnz
    for a, b in zip(A.get_h(), B.get_h()):
        np.testing.assert_array_equal(a, b)
    X = np.random.default_rng(64).standard_normal((A.n, 4))
    ys = A.sigma(X, api.STORED, api.FP64)
    yd = B.sigma(X, api.DIRECT, api.FP64)
    assert np.abs(yd - ys).max() <= 1e-12 * np.abs(ys).max()
    ea = A.davidson(4, prec=api.FP64, mode=api.STORED, tol=1e-8)
    eb = B.davidson(4, prec=api.MIXED, mode=api.DIRECT, tol=1e-8)
    assert ea["converged"] and eb["converged"]
    np.testing.assert_allclose(eb["evals"], ea["evals"], atol=1e-9, rtol=0)
    A.close()
    B.close()


def open_list(nelec, ksym, min_open):
    """The 10-orbital configurations with at least min_open open shells."""
    full = inputs.fci_configs(10, nelec, jsym=1, ksym=ksym)
    keep = (full.occ == 1).sum(1) >= min_open
    return inputs.Configs(nbf=10, nelec=nelec, ksym=ksym, occ=full.occ[keep], jsym=full.jsym[keep],
                          title=f"{min_open} or more open shells")


def test_ten_open_shells_stored_and_direct():
    """8 and 10 open shells (1 x 256 + 90 x 64 SEFs): the stored H is the
    oracle's bit for bit; the direct sigma, which splits the 256-SEF
    configuration over four wavefronts, equals the stored sigma; the
    adjoint reproduces each root's energy (Euler: H is linear in the
    integrals)."""
    cfg = open_list(10, 1, 8)
    assert cfg.ncfg == 91 and (cfg.occ == 1).sum(1).max() == 10
    ints = inputs.make_integrals(10, 1010)
    H = api.Hamiltonian(cfg, ints)
    sp = oracle.OracleSpace(cfg, ints)
    sp.build_h(8)
    assert H.n == 256 + 90 * 64
    assert H.build() == len(sp.val)
    rowptr, col, val, diag = H.get_h()
    np.testing.assert_array_equal(rowptr, sp.rowptr)
    np.testing.assert_array_equal(col, sp.col)
    assert np.array_equal(val, sp.val) and np.array_equal(diag, sp.diag)
    X = np.random.default_rng(10).standard_normal((H.n, 16))
    ys = H.sigma(X, api.STORED, api.FP64)
    yd = H.sigma(X, api.DIRECT, api.FP64)
    assert np.abs(yd - ys).max() <= 1e-12 * np.abs(ys).max()
    yd32 = H.sigma(X.astype(np.float32), api.DIRECT, api.FP32)
    assert np.abs(yd32 - ys).max() <= 2e-5 * np.abs(ys).max()
    # H conserves |Ms|, so a solve started in one Ms sector (the lowest
    # diagonal elements) finds that sector's roots; here the lowest exact
    # level is a multiplet spread over sectors.  Each converged root is an
    # exact eigenpair of the oracle's matrix, and mixed direct equals FP64.
    Hd = sp.dense_h()
    w = np.linalg.eigvalsh(Hd)
    ref = H.davidson(4, prec=api.FP64, mode=api.STORED, tol=1e-8)
    out = H.davidson(4, prec=api.MIXED, mode=api.DIRECT, tol=1e-8)
    assert ref["converged"] and out["converged"]
    np.testing.assert_allclose(out["evals"], ref["evals"], atol=1e-9, rtol=0)
    for e in out["evals"]:
        assert np.abs(w - e).min() <= 1e-9
    V = out["evecs"]
    assert np.abs(Hd @ V - V * out["evals"]).max() <= 1e-7
    for r in range(4):
        wr = np.zeros(4)
        wr[r] = 1.0
        g1, g2, _ = H.rdm(V, wr)
        assert abs(g1 @ ints.valone + g2 @ ints.valtwo - out["evals"][r]) < 1e-9
    H.close()


@pytest.mark.parametrize("nelec,ksym", [(9, 1), (10, 3)])
def test_split_configurations_other_classes(nelec, ksym):
    """One determinant per SEF with 7 and 9 open shells (64 and 256 SEFs) and
    the ODD class with 8 and 10: direct sigma equal to the oracle's."""
    cfg = open_list(nelec, ksym, 7)
    ints = inputs.make_integrals(10, 2020 + nelec)
    H = api.Hamiltonian(cfg, ints)
    sp = oracle.OracleSpace(cfg, ints)
    sp.build_h(8)
    X = np.random.default_rng(nelec).standard_normal((H.n, 8))
    ref = sp.sigma(X)
    yd = H.sigma(X, api.DIRECT, api.FP64)
    assert np.abs(yd - ref).max() <= 1e-12 * np.abs(ref).max()
    H.close()


@pytest.mark.parametrize("occ", ["2200", "1111", "2110"])
def test_one_configuration(occ):
    """A one-configuration space: H is the configuration's diagonal block,
    so the eigenvalues are those of the oracle's block."""
    row = np.array([[int(c) for c in occ]], np.uint8)
    cfg = inputs.Configs(nbf=4, nelec=4, ksym=1, occ=row, jsym=np.ones(1, np.int32), title=occ)
    ints = inputs.make_integrals(4, 44)
    H = api.Hamiltonian(cfg, ints)
    sp = oracle.OracleSpace(cfg, ints)
    _, blk = sp.block(0, 0)
    w = np.linalg.eigvalsh(blk)
    assert H.n == len(w)
    X = np.random.default_rng(1).standard_normal((H.n, 2))
    H.build()
    np.testing.assert_allclose(H.sigma(X, api.STORED, api.FP64), blk @ X, rtol=0, atol=1e-13)
    np.testing.assert_allclose(H.sigma(X, api.DIRECT, api.FP64), blk @ X, rtol=0, atol=1e-13)
    out = H.davidson(H.n, prec=api.FP64, mode=api.DIRECT, tol=1e-9)
    assert out["converged"]
    np.testing.assert_allclose(out["evals"], w, atol=1e-10, rtol=0)
    with pytest.raises(api.PsociError) as e:
        H.davidson(H.n + 1, prec=api.FP64, mode=api.DIRECT)
    assert e.value.code == api.ERR_ARG
    H.close()


def test_empty_configuration_list():
    cfg = inputs.Configs(nbf=4, nelec=4, ksym=1, occ=np.zeros((0, 4), np.uint8), jsym=np.zeros(0, np.int32))
    with pytest.raises(api.PsociError) as e:
        api.Hamiltonian(cfg, inputs.make_integrals(4, 1))
    assert e.value.code == api.ERR_ARG


@pytest.mark.parametrize("nroots,extra", [(12, 0), (8, 6), (16, 0)])
def test_block_wider_than_half_the_space(nroots, extra):
    """A working block k with k < n < 2k (here n = 22): the subspace holds
    at most n columns, so after a collapse fewer columns are free than roots
    are selected; the corrections of the lowest roots are kept and the
    eigenpairs are exact: every returned value is an eigenvalue of the
    oracle's dense H (LAPACK; distinct eigenvalues matched one to one) and
    every returned vector satisfies the eigen-equation.  (H conserves |Ms|,
    so a solve started on unit vectors may stay out of a sector: the values
    are checked as a sub-multiset of the spectrum, not as its lowest k.)"""
    cfg = inputs.fci_configs(4, 4)
    ints = inputs.make_integrals(4, 7)
    H = api.Hamiltonian(cfg, ints)
    assert nroots + extra < H.n < 2 * (nroots + extra)
    sp = oracle.OracleSpace(cfg, ints)
    sp.build_h()
    w = np.linalg.eigvalsh(sp.dense_h())
    H.build()
    for mode, prec in ((api.STORED, api.FP64), (api.DIRECT, api.FP64), (api.DIRECT, api.MIXED)):
        out = H.davidson(nroots, prec=prec, mode=mode, tol=1e-8, extra_roots=extra, guess=1)
        assert out["converged"], (mode, prec)
        free = list(w)
        for e in out["evals"]:
            j = int(np.argmin(np.abs(np.array(free) - e)))
            assert abs(free[j] - e) < 1e-9, (mode, prec, e)
            free.pop(j)
        V = out["evecs"]
        np.testing.assert_allclose(V.T @ V, np.eye(nroots), atol=1e-9)
        R = sp.dense_h() @ V - V * out["evals"]
        assert np.linalg.norm(R, axis=0).max() < 1e-6, (mode, prec)
    H.close()
