"""Adjoint (RDM) kernel through the C ABI.

Pinned three ways: (1) H is homogeneous of degree one in the integrals, so
sum_theta theta * dE/dtheta must reproduce E (Euler); (2) the one-electron
part must equal the oracle's CIDEN 1-RDM (grad_one[pp] = D_pp,
grad_one[pq] = 2 D_pq: a packed h_pq enters H_IJ and H_JI while the CI
density counts each configuration pair once); (3) central finite
differences of the eigenvalues (step 1e-5) along 20 seeded random
perturbations of the integrals, on the 31,650-SEF stored-H space."""
import numpy as np
import pytest

import oracle
from psoci_amd import api, inputs

pytestmark = pytest.mark.gpu


def test_rdm_homogeneity_and_oracle_density(fci6):
    cfg, ints, k = fci6
    H = api.Hamiltonian(cfg, ints)
    out = H.davidson(k, prec=api.FP64, mode=api.DIRECT, tol=1e-10)
    assert out["converged"]
    sp = oracle.OracleSpace(cfg, ints)
    for r in range(k):
        c = out["evecs"][:, r]
        w = np.zeros(k)
        w[r] = 1.0
        g1, g2, gso = H.rdm(out["evecs"], w)
        e = g1 @ ints.valone + g2 @ ints.valtwo
        assert abs(e - out["evals"][r]) < 1e-10, (r, e, out["evals"][r])
        assert not gso.any()
        D = sp.rdm1(c)
        nbf = cfg.nbf
        for p in range(nbf):
            for q in range(p + 1):
                want = D[inputs.pair_index(p, q)] * (1.0 if p == q else 2.0)
                assert abs(g1[inputs.pair_index(p, q)] - want) < 1e-10, (r, p, q)
    H.close()


def test_rdm_matches_oracle_density_fci10(fci10, oracle_fci10):
    """The adjoint's one-electron part for converged roots of the
    31,650-SEF space equals the oracle's CIDEN 1-RDM (the reference's
    constructDensityMatrix): grad_one[pp] = D_pp, grad_one[pq] = 2 D_pq."""
    cfg, ints, _ = fci10
    k = 4
    H = api.Hamiltonian(cfg, ints)
    H.build()
    out = H.davidson(k, prec=api.FP64, mode=api.STORED, tol=1e-9)
    assert out["converged"]
    sp = oracle_fci10
    nbf = cfg.nbf
    scale = np.array([1.0 if p == q else 2.0 for p in range(nbf) for q in range(p + 1)])
    for r in range(k):
        w = np.zeros(k)
        w[r] = 1.0
        g1, g2, _ = H.rdm(out["evecs"], w)
        D = sp.rdm1(out["evecs"][:, r])
        np.testing.assert_allclose(g1, D * scale, atol=1e-11, rtol=0)
        assert abs(g1 @ ints.valone + g2 @ ints.valtwo - out["evals"][r]) < 1e-9
    H.close()


def test_rdm_finite_differences_fci10(fci10):
    cfg, ints, _ = fci10
    # roots 3 and 4 of this space are an exactly degenerate pair (spin
    # components of one multiplet): the sum over the lowest five is smooth
    k = 5
    H = api.Hamiltonian(cfg, ints)
    H.build()
    base = H.davidson(k, prec=api.FP64, mode=api.STORED, tol=1e-9)
    assert base["converged"]
    g1, g2, _ = H.rdm(base["evecs"], np.ones(k))
    rng = np.random.default_rng(20261204)
    eps = 1e-5
    worst = 0.0
    for t in range(20):
        d1 = rng.standard_normal(ints.valone.shape)
        d2 = rng.standard_normal(ints.valtwo.shape) * 0.1
        es = []
        for sgn in (+1, -1):
            pert = inputs.Integrals(nbf=cfg.nbf, valone=ints.valone + sgn * eps * d1,
                                    valtwo=ints.valtwo + sgn * eps * d2)
            H.set_integrals(pert)
            H.build()
            o = H.davidson(k, prec=api.FP64, mode=api.STORED, tol=1e-9)
            assert o["converged"]
            es.append(o["evals"].sum())
        fd = (es[0] - es[1]) / (2 * eps)
        an = g1 @ d1 + g2 @ d2
        err = abs(fd - an) / max(1.0, abs(an))
        worst = max(worst, err)
        assert err < 1e-6, (t, fd, an)
    H.close()


def test_autograd_wrapper_matches_adjoint(fci6):
    import torch

    from psoci_amd.autograd import energies
    cfg, ints, k = fci6
    H = api.Hamiltonian(cfg, ints)
    one = torch.tensor(ints.valone, dtype=torch.float64, requires_grad=True)
    two = torch.tensor(ints.valtwo, dtype=torch.float64, requires_grad=True)
    so = torch.tensor(ints.so, dtype=torch.float64, requires_grad=True)
    E = energies(one, two, so, H, k, api.DIRECT, 1e-10)
    w = torch.tensor([1.0, -0.5, 0.25, 2.0], dtype=torch.float64)
    (E * w).sum().backward()
    out = H.davidson(k, prec=api.FP64, mode=api.DIRECT, tol=1e-10)
    g1, g2, _ = H.rdm(out["evecs"], w.numpy())
    np.testing.assert_allclose(one.grad.numpy(), g1, atol=1e-9)
    np.testing.assert_allclose(two.grad.numpy(), g2, atol=1e-9)
    assert not so.grad.any()
    H.close()
