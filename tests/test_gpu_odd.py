"""ODD-class spaces on the GPU through the C ABI (state symmetry class
different from the configurations': generateODDSpinDeterminants,
PsociDeterminants.C:1613-1658, 3634-4353): determinants and phases equal to
the oracle's, the stored Hamiltonian bit-identical to the oracle's packSefsef
rows, the direct sigma equal to the stored one (closed-shell / open-shell
pairs, which never connect in this class, are left out of its loops), and
Davidson eigenvalues equal to dense LAPACK on the oracle's matrix, for both
SEF signs (ksym 3: +, ksym 4: -)."""
import numpy as np
import pytest

import oracle
from psoci_amd import api, inputs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ksym", [3, 4])
def test_odd_fci6_dets_stored_h_and_spectrum(fci6, ksym):
    c6, ints, _ = fci6
    cfg = inputs.fci_configs(c6.nbf, c6.nelec, jsym=1, ksym=ksym)
    H = api.Hamiltonian(cfg, ints)
    sp = oracle.OracleSpace(cfg, ints)
    for I in range(cfg.ncfg):
        d, o = H.dets(I), sp.dets(I)
        for k in range(len(o["alpha"])):
            assert int(d["alpha"][k]) == sum(1 << p for p in o["alpha"][k])
            assert int(d["beta"][k]) == sum(1 << p for p in o["beta"][k])
        np.testing.assert_array_equal(d["ms"], o["ms"])
        np.testing.assert_array_equal(d["phase"], o["phase"])
    sp.build_h(1)
    assert H.build() == len(sp.val)
    rowptr, col, val, diag = H.get_h()
    np.testing.assert_array_equal(rowptr, sp.rowptr)
    np.testing.assert_array_equal(col, sp.col)
    assert np.array_equal(val, sp.val) and np.array_equal(diag, sp.diag)
    X = np.random.default_rng(ksym).standard_normal((H.n, 5))
    ys = H.sigma(X, api.STORED, api.FP64)
    assert np.abs(ys - sp.sigma(X)).max() <= 1e-12 * np.abs(ys).max()
    yd = H.sigma(X, api.DIRECT, api.FP64)
    assert np.abs(yd - ys).max() <= 1e-12 * np.abs(ys).max()
    w = np.linalg.eigvalsh(sp.dense_h())
    out = H.davidson(4, prec=api.FP64, mode=api.DIRECT, tol=1e-8)
    assert out["converged"]
    np.testing.assert_allclose(out["evals"], w[:4], atol=1e-9, rtol=0)
    H.close()


def test_odd_fci10_direct_and_stored(fci10):
    c10, ints, _ = fci10
    cfg = inputs.fci_configs(c10.nbf, c10.nelec, jsym=1, ksym=3)
    H = api.Hamiltonian(cfg, ints)
    sp = oracle.OracleSpace(cfg, ints)
    sp.build_h(8)
    H.build()
    rowptr, col, val, diag = H.get_h()
    np.testing.assert_array_equal(rowptr, sp.rowptr)
    np.testing.assert_array_equal(col, sp.col)
    assert np.array_equal(val, sp.val)
    for k in (1, 16):
        X = np.random.default_rng(k).standard_normal((H.n, k))
        ys = H.sigma(X, api.STORED, api.FP64)
        yd = H.sigma(X, api.DIRECT, api.FP64)
        assert np.abs(yd - ys).max() <= 1e-12 * np.abs(ys).max()
        yd32 = H.sigma(X.astype(np.float32), api.DIRECT, api.FP32)
        assert np.abs(yd32 - ys).max() <= 2e-5 * np.abs(ys).max()
    a = H.davidson(4, prec=api.FP64, mode=api.STORED, tol=1e-7)
    b = H.davidson(4, prec=api.MIXED, mode=api.DIRECT, tol=1e-7)
    assert a["converged"] and b["converged"]
    np.testing.assert_allclose(b["evals"], a["evals"], atol=1e-9, rtol=0)
    H.close()


def test_odd_mrci_direct_rows_match_oracle():
    """An MR-CISD list in the ODD class: sampled rows of the direct sigma
    against the oracle's blocks."""
    cfg = inputs.mrcisd_configs(4, 6, 4, ksym=3)
    ints = inputs.make_integrals(10, 99, n_internal=4)
    H = api.Hamiltonian(cfg, ints)
    sp = oracle.OracleSpace(cfg, ints)
    sp.build_h(8)
    X = np.random.default_rng(2).standard_normal((H.n, 4))
    ref = sp.sigma(X)
    Y = H.sigma(X, api.DIRECT, api.FP64)
    assert np.abs(Y - ref).max() <= 1e-12 * np.abs(ref).max()
    H.close()
