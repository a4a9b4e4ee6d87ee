"""Odd electron counts on the GPU through the C ABI (the fractional
determinants, generateSpinFractionalDeterminants, PsociDeterminants.C:
1508-1561, 1680-2443; transformCDCt keeps detdet, PsociHamiltonian.C:
1386-1391): determinants and phases equal to the oracle's, the stored
Hamiltonian bit-identical to the oracle's rows, the direct sigma equal to
the stored one, eigenvalues equal to dense LAPACK, and the adjoint's 1-RDM
equal to the oracle's CIDEN density -- for both spin orientations (jsym 1:
Ms = 1/2, -3/2, ...; jsym 3: alpha and beta swapped)."""
import numpy as np
import pytest

import oracle
from psoci_amd import api, inputs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("jsym", [1, 3])
def test_fractional_fci6_parity(jsym):
    cfg = inputs.fci_configs(6, 5, jsym=jsym, ksym=jsym)
    ints = inputs.make_integrals(6, 424242)
    H = api.Hamiltonian(cfg, ints)
    sp = oracle.OracleSpace(cfg, ints)
    for I in range(cfg.ncfg):
        d, o = H.dets(I), sp.dets(I)
        for k in range(len(o["alpha"])):
            assert int(d["alpha"][k]) == sum(1 << p for p in o["alpha"][k])
            assert int(d["beta"][k]) == sum(1 << p for p in o["beta"][k])
        np.testing.assert_array_equal(d["phase"], o["phase"])
    sp.build_h(1)
    assert H.build() == len(sp.val)
    rowptr, col, val, diag = H.get_h()
    np.testing.assert_array_equal(rowptr, sp.rowptr)
    np.testing.assert_array_equal(col, sp.col)
    assert np.array_equal(val, sp.val) and np.array_equal(diag, sp.diag)
    X = np.random.default_rng(jsym).standard_normal((H.n, 4))
    ys = H.sigma(X, api.STORED, api.FP64)
    assert np.abs(ys - sp.sigma(X)).max() <= 1e-12 * np.abs(ys).max()
    yd = H.sigma(X, api.DIRECT, api.FP64)
    assert np.abs(yd - ys).max() <= 1e-12 * np.abs(ys).max()
    w = np.linalg.eigvalsh(sp.dense_h())
    out = H.davidson(4, prec=api.FP64, mode=api.DIRECT, tol=1e-8)
    assert out["converged"]
    np.testing.assert_allclose(out["evals"], w[:4], atol=1e-9, rtol=0)
    # the adjoint: dE/dh_pp = D_pp, dE/dh_pq = 2 D_pq against the oracle's density
    c = out["evecs"][:, 0]
    g1, g2, _ = H.rdm(c)
    assert abs(g1 @ ints.valone + g2 @ ints.valtwo - out["evals"][0]) < 1e-10
    np.testing.assert_allclose(api.rdm1_from_grad(g1), sp.rdm1(c), atol=1e-10)
    H.close()


def test_fractional_mrci_direct_and_mixed():
    cfg = inputs.mrcisd_configs(6, 10, 7)
    ints = inputs.make_integrals(16, 5151, n_internal=6)
    H = api.Hamiltonian(cfg, ints)
    H.build()
    for k in (1, 8):
        X = np.random.default_rng(k).standard_normal((H.n, k))
        ys = H.sigma(X, api.STORED, api.FP64)
        yd = H.sigma(X, api.DIRECT, api.FP64)
        assert np.abs(yd - ys).max() <= 1e-12 * np.abs(ys).max()
    a = H.davidson(4, prec=api.FP64, mode=api.STORED, tol=1e-7)
    b = H.davidson(4, prec=api.MIXED, mode=api.DIRECT, tol=1e-7)
    assert a["converged"] and b["converged"]
    np.testing.assert_allclose(b["evals"], a["evals"], atol=1e-9, rtol=0)
    H.close()
