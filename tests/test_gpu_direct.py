"""Direct sigma (elements regenerated on the fly) through the C ABI, both
engines (1: configuration pair records, 2: string-driven products): against
the stored Hamiltonian on the FCI spaces, and against the oracle's own
configuration-pair blocks, row by row, on the MR-CISD spaces (sampled rows:
a size-independent check)."""
import numpy as np
import pytest

import oracle
from psoci_amd import api, inputs

pytestmark = pytest.mark.gpu


def connected(cfg, I):
    """Configurations within two moved electrons and four open shells of I."""
    occ = cfg.occ.astype(np.int16)
    nopen = (cfg.occ == 1).sum(1)
    moved = np.maximum(occ[I] - occ, 0).sum(1)
    return np.where((moved <= 2) & (np.abs(nopen[I] - nopen) <= 4))[0]


def oracle_rows(sp, cfg, I, X):
    """y_I = sum_J H_IJ x_J from the oracle's blocks (reference orientation)."""
    off = sp.sef_off
    ni = int(sp.nsef[I])
    y = np.zeros((ni, X.shape[1]))
    for J in connected(cfg, I):
        J = int(J)
        if J <= I:
            st, B = sp.block(I, J)
        else:
            st, B = sp.block(J, I)
            B = B.T
        if st:
            y += B @ X[off[J]:off[J] + int(sp.nsef[J])]
    return y


@pytest.fixture(scope="module", params=[1, 2], ids=["records", "strings"])
def fci10_gpu(fci10, request):
    cfg, ints, _ = fci10
    H = api.Hamiltonian(cfg, ints)
    H.build()
    assert H.direct_engine(request.param) == request.param
    yield H
    H.close()


@pytest.mark.parametrize("k", [1, 2, 3, 4, 8, 16, 24, 32])
def test_direct_matches_stored_fci10(fci10_gpu, k):
    """Every kernel width the library instantiates (1, 2, 4, 8, 16, 24, 32;
    3 runs padded to 4), FP64 and FP32."""
    H = fci10_gpu
    X = np.random.default_rng(k).standard_normal((H.n, k))
    ys = H.sigma(X, api.STORED, api.FP64)
    yd = H.sigma(X, api.DIRECT, api.FP64)
    assert np.abs(yd - ys).max() <= 1e-12 * np.abs(ys).max()
    yd32 = H.sigma(X.astype(np.float32), api.DIRECT, api.FP32)
    assert np.abs(yd32 - ys).max() <= 2e-5 * np.abs(ys).max()


def test_direct_davidson_matches_stored_fci10(fci10_gpu):
    H = fci10_gpu
    a = H.davidson(4, prec=api.FP64, mode=api.STORED, tol=1e-6)
    b = H.davidson(4, prec=api.FP64, mode=api.DIRECT, tol=1e-6)
    assert a["converged"] and b["converged"]
    np.testing.assert_allclose(b["evals"], a["evals"], atol=1e-9, rtol=0)


@pytest.mark.parametrize("engine", [1, 2])
@pytest.mark.parametrize("k", [8, 24])
def test_direct_rows_match_oracle_mrci8x24(k, engine):
    """k = 24 is the width of an FP32 solve phase with 16 roots and 8 guard
    roots."""
    cfg, ints, _ = inputs.workload("mrci8x24")
    H = api.Hamiltonian(cfg, ints)
    assert H.direct_engine(engine) == engine
    sp = oracle.OracleSpace(cfg, ints)
    rng = np.random.default_rng(3)
    X = rng.standard_normal((H.n, k))
    Y = H.sigma(X, api.DIRECT, api.FP64)
    Y32 = H.sigma(X.astype(np.float32), api.DIRECT, api.FP32)
    nopen = (cfg.occ == 1).sum(1)
    rows = list(rng.choice(cfg.ncfg, 12, replace=False)) + list(np.where(nopen == 6)[0][:2]) + [0, cfg.ncfg - 1]
    for I in rows:
        I = int(I)
        ref = oracle_rows(sp, cfg, I, X)
        got = Y[sp.sef_off[I]:sp.sef_off[I] + int(sp.nsef[I])]
        scale = max(1.0, np.abs(ref).max())
        assert np.abs(got - ref).max() <= 1e-12 * scale, I
        got32 = Y32[sp.sef_off[I]:sp.sef_off[I] + int(sp.nsef[I])]
        assert np.abs(got32 - ref).max() <= 5e-5 * scale, I
    # symmetry of the operator: u^T H v == v^T H u
    u = rng.standard_normal(H.n)
    v = rng.standard_normal(H.n)
    Hu = H.sigma(u, api.DIRECT, api.FP64)
    Hv = H.sigma(v, api.DIRECT, api.FP64)
    assert abs(v @ Hu - u @ Hv) <= 1e-11 * abs(v @ Hu)
    H.close()


def test_direct_mixed_davidson_mrci8x24():
    """The 8-root direct-sigma workload: mixed precision agrees with FP64."""
    cfg, ints, k = inputs.workload("mrci8x24")
    H = api.Hamiltonian(cfg, ints)
    a = H.davidson(k, prec=api.FP64, mode=api.DIRECT, tol=1e-5)
    b = H.davidson(k, prec=api.MIXED, mode=api.DIRECT, tol=1e-5, switch_tol=1e-4)
    assert a["converged"] and b["converged"]
    np.testing.assert_allclose(b["evals"], a["evals"], atol=1e-8, rtol=0)
    # Ritz vectors satisfy the eigen-equation through an independent FP64 sigma
    V = a["evecs"]
    R = H.sigma(V, api.DIRECT, api.FP64) - V * a["evals"]
    assert np.linalg.norm(R, axis=0).max() < 1e-4
    H.close()


def test_direct_mixed_guard_roots_mrci8x24():
    """Guard roots in the FP32 phase (the benchmark's 16 + 8 block shape on
    the 8-root workload: 8 + 8): every wanted root within 1e-8 Eh of the
    all-FP64 solve."""
    cfg, ints, k = inputs.workload("mrci8x24")
    H = api.Hamiltonian(cfg, ints)
    a = H.davidson(k, prec=api.FP64, mode=api.DIRECT, tol=1e-5, want_vectors=False)
    b = H.davidson(k, prec=api.MIXED, mode=api.DIRECT, tol=1e-5, switch_tol=1e-4, extra_roots=8,
                   want_vectors=False)
    assert a["converged"] and b["converged"]
    assert b["stats"]["switch_iter"] > 0
    np.testing.assert_allclose(b["evals"], a["evals"], atol=1e-8, rtol=0)
    H.close()


@pytest.mark.parametrize("engine", [1, 2])
def test_direct_full_size_mrci10x24(engine):
    """The 47.6M-determinant workload at full size: sampled rows of the FP64
    and FP32 direct sigma against the oracle's blocks, and the symmetry of
    the operator (u^T H v = v^T H u) over all 11,894,370 SEFs."""
    cfg, ints, _ = inputs.workload("mrci10x24")
    H = api.Hamiltonian(cfg, ints)
    assert H.direct_engine(engine) == engine
    sp = oracle.OracleSpace(cfg, ints)
    rng = np.random.default_rng(17)
    X = rng.standard_normal((H.n, 4))
    Y = H.sigma(X, api.DIRECT, api.FP64)
    Y32 = H.sigma(X.astype(np.float32), api.DIRECT, api.FP32)
    nopen = (cfg.occ == 1).sum(1)
    rows = [int(r) for r in rng.choice(cfg.ncfg, 5, replace=False)] + [int(np.where(nopen == 8)[0][7])]
    rows += [0, cfg.ncfg - 1]
    for I in rows:
        ref = oracle_rows(sp, cfg, I, X)
        o = int(sp.sef_off[I])
        got = Y[o:o + int(sp.nsef[I])]
        scale = max(1.0, np.abs(ref).max())
        assert np.abs(got - ref).max() <= 1e-12 * scale, I
        assert np.abs(Y32[o:o + int(sp.nsef[I])] - ref).max() <= 5e-5 * scale, I
    u = X[:, 0].copy()
    v = X[:, 1].copy()
    assert abs(v @ Y[:, 0] - u @ Y[:, 1]) <= 1e-11 * abs(v @ Y[:, 0])
    H.close()


def test_det_pair_count_fci6(fci6):
    """The work counter of the direct sigma (connected determinant pairs per
    launch) equals a brute-force count: every ordered configuration pair
    within two moved electrons and four open shells (J = I included), every
    row determinant 2i (one per SEF) against every column determinant with
    the same alpha count and at most two spin-orbital replacements."""
    cfg, ints, _ = fci6
    H = api.Hamiltonian(cfg, ints)
    dets = [H.dets(I) for I in range(cfg.ncfg)]
    pc = np.vectorize(lambda x: bin(int(x)).count("1"))
    total = 0
    for I in range(cfg.ncfg):
        a = dets[I]
        rows = np.arange(0, len(a["alpha"]), 2)[:int(H.info.nsef[I])]
        aa, ab = a["alpha"][rows], a["beta"][rows]
        for J in connected(cfg, I):
            b = dets[int(J)]
            da = np.bitwise_xor.outer(aa, b["alpha"])
            db = np.bitwise_xor.outer(ab, b["beta"])
            same_ms = pc(aa)[:, None] == pc(b["alpha"])[None, :]
            total += int(((pc(da) + pc(db) <= 4) & same_ms).sum())
    assert H.direct_det_pairs() == total
    H.close()
