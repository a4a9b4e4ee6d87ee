"""Spin-orbit blocks on the GPU (spaces whose configurations carry more
than one jsym): the stored Hamiltonian -- pattern, values and diagonal --
and single blocks equal the oracle's restatement of generateSOCIblocks /
socig bit for bit; the determinants of both classes equal the oracle's;
Davidson eigenpairs are eigenpairs of the oracle's dense H (LAPACK) and,
with an odd electron count and both fractional orientations, come in
Kramers pairs; the direct sigma and the adjoint refuse such spaces."""
import numpy as np
import pytest

import oracle
from psoci_amd import api, inputs

pytestmark = pytest.mark.gpu

SPACES = [(4, 4, (1, 3)), (4, 4, (1, 2, 3, 4)), (5, 3, (1, 3)), (5, 5, (2, 3)), (6, 4, (1, 3))]


def mixed_space(nbf, nelec, jsyms, seed=5):
    base = inputs.fci_configs(nbf, nelec)
    occ = np.concatenate([base.occ] * len(jsyms))
    jsym = np.concatenate([np.full(base.ncfg, j, np.int32) for j in jsyms])
    cfg = inputs.Configs(nbf=nbf, nelec=nelec, ksym=1, occ=occ, jsym=jsym, title="mixed jsym")
    return cfg, inputs.make_integrals(nbf, seed)


@pytest.mark.parametrize("nbf,nelec,jsyms", SPACES)
def test_stored_h_and_blocks_bit_identical(nbf, nelec, jsyms):
    cfg, ints = mixed_space(nbf, nelec, jsyms)
    sp = oracle.OracleSpace(cfg, ints)
    sp.build_h()
    H = api.Hamiltonian(cfg, ints)
    assert H.n == sp.n
    nnz = H.build()
    assert nnz == len(sp.val)
    rowptr, col, val, diag = H.get_h()
    np.testing.assert_array_equal(rowptr, sp.rowptr)
    np.testing.assert_array_equal(col, sp.col)
    assert np.array_equal(val, sp.val)
    assert np.array_equal(diag, sp.diag)
    assert np.array_equal(H.diag(), sp.diag)
    rng = np.random.default_rng(nbf * 10 + nelec)
    for I in rng.choice(cfg.ncfg, 12, replace=False):
        I = int(I)
        d = H.dets(I)
        o = sp.dets(I)
        for k in range(len(o["alpha"])):
            assert int(d["alpha"][k]) == sum(1 << p for p in o["alpha"][k])
            assert int(d["beta"][k]) == sum(1 << p for p in o["beta"][k])
        for J in range(cfg.ncfg):
            st, ref = sp.block(I, J)
            st2, out = H.block(I, J)
            assert st == st2 and np.array_equal(ref, out), (I, J)
    X = rng.standard_normal((H.n, 3))
    ref = sp.sigma(X)
    assert np.abs(H.sigma(X, api.STORED, api.FP64) - ref).max() <= 1e-12 * np.abs(ref).max()
    H.close()


@pytest.mark.parametrize("nbf,nelec,jsyms", [(5, 3, (1, 3)), (6, 4, (1, 3)), (5, 5, (1, 3))])
def test_davidson_eigenpairs(nbf, nelec, jsyms):
    cfg, ints = mixed_space(nbf, nelec, jsyms)
    sp = oracle.OracleSpace(cfg, ints)
    sp.build_h()
    Hd = sp.dense_h()
    w = np.linalg.eigvalsh(Hd)
    H = api.Hamiltonian(cfg, ints)
    H.build()
    for prec in (api.FP64, api.MIXED):
        out = H.davidson(6, prec=prec, mode=api.STORED, tol=1e-8)
        assert out["converged"], prec
        free = list(w)
        for e in out["evals"]:
            j = int(np.argmin(np.abs(np.array(free) - e)))
            assert abs(free[j] - e) < 1e-9, (prec, e)
            free.pop(j)
        V = out["evecs"]
        assert np.linalg.norm(Hd @ V - V * out["evals"], axis=0).max() < 1e-6
        if nelec % 2 == 1:
            # time reversal: Kramers pairs
            ev = out["evals"]
            assert np.abs(ev[0::2] - ev[1::2]).max() < 1e-9, ev
    with pytest.raises(api.PsociError) as e:
        H.sigma(np.ones(H.n), api.DIRECT)
    assert e.value.code == api.ERR_UNSUPPORTED
    with pytest.raises(api.PsociError) as e:
        H.rdm(np.ones((H.n, 1)) / np.sqrt(H.n))
    assert e.value.code == api.ERR_UNSUPPORTED
    H.close()
