"""The string-driven direct sigma (engine 2: resolutions of the identity
over strings, dense products on MFMA) through the C ABI: against the
oracle's stored sigma on small complete RAS spaces (full CI, MR-CISD,
MR-CIS, both SEF signs), the library's stored sigma on fci10 and the
configuration-driven engine on mrci8x24, FP64 and FP32, at every column
chunk width; Davidson roots; engine selection and refusals."""
import numpy as np
import pytest

import oracle
from psoci_amd import api, inputs

pytestmark = pytest.mark.gpu

SMALL = [
    ("fci6", lambda: inputs.workload("fci6")[:2]),
    ("fci6 mult -1", lambda: (inputs.fci_configs(6, 6, jsym=1, ksym=2), inputs.make_integrals(6, 6))),
    ("fci 8/8", lambda: (inputs.fci_configs(8, 8), inputs.make_integrals(8, 5))),
    ("fci 7/2", lambda: (inputs.fci_configs(7, 2), inputs.make_integrals(7, 9))),
    ("mrcisd 4+5/4", lambda: (inputs.mrcisd_configs(4, 5, 4), inputs.make_integrals(9, 11, n_internal=4))),
    ("mrcisd 3+4/6", lambda: (inputs.mrcisd_configs(3, 4, 6), inputs.make_integrals(7, 12, n_internal=3))),
    ("mrcis 4+3/4", lambda: (inputs.mrcisd_configs(4, 3, 4, max_external=1),
                             inputs.make_integrals(7, 8, n_internal=4))),
    ("mrcisd 5+4/6 mult -1", lambda: (inputs.mrcisd_configs(5, 4, 6, jsym=1, ksym=2),
                                      inputs.make_integrals(9, 13, n_internal=5))),
]


@pytest.mark.parametrize("name,make", SMALL, ids=[c[0] for c in SMALL])
def test_ras_sigma_matches_oracle(name, make):
    """Column counts 1, 2, 3, 8, 16, 21: chunk widths 1, 2, 4, 8 and 16
    (FP32) / 8 (FP64), and a block over several chunks."""
    cfg, ints = make()
    H = api.Hamiltonian(cfg, ints)
    assert H.direct_engine(2) == 2
    sp = oracle.OracleSpace(cfg, ints)
    sp.build_h()
    rng = np.random.default_rng(11)
    for k in (1, 2, 3, 8, 16, 21):
        X = rng.standard_normal((H.n, k))
        ref = sp.sigma(X)
        scale = np.abs(ref).max()
        y = H.sigma(X, api.DIRECT, api.FP64).reshape(ref.shape)
        assert np.abs(y - ref).max() <= 1e-12 * scale, (name, k)
        y32 = H.sigma(X.astype(np.float32), api.DIRECT, api.FP32).reshape(ref.shape)
        assert np.abs(y32 - ref).max() <= 2e-5 * scale, (name, k)
    H.close()


@pytest.fixture(scope="module")
def fci10_strings(fci10):
    cfg, ints, _ = fci10
    H = api.Hamiltonian(cfg, ints)
    H.build()
    assert H.direct_engine(2) == 2
    yield H
    H.close()


@pytest.mark.parametrize("k", [2, 16, 24, 32])
def test_ras_sigma_matches_stored_fci10(fci10_strings, k):
    H = fci10_strings
    X = np.random.default_rng(k).standard_normal((H.n, k))
    ys = H.sigma(X, api.STORED, api.FP64)
    yd = H.sigma(X, api.DIRECT, api.FP64)
    assert np.abs(yd - ys).max() <= 1e-12 * np.abs(ys).max()
    yd32 = H.sigma(X.astype(np.float32), api.DIRECT, api.FP32)
    assert np.abs(yd32 - ys).max() <= 2e-5 * np.abs(ys).max()


def test_ras_plan_info_fci10(fci10_strings):
    info = fci10_strings.direct_ras_info()
    # full CI: no external orbitals; sectors n_alpha - n_beta = 0, 4, 8
    assert info["m"] == 0 and info["n_external"] == 0 and info["sectors"] == 3
    assert info["ndet"] == 44100 + 9450 + 45


def test_ras_davidson_fci10(fci10_strings):
    H = fci10_strings
    a = H.davidson(4, prec=api.FP64, mode=api.STORED, tol=1e-7, want_vectors=False)
    b = H.davidson(4, prec=api.FP64, mode=api.DIRECT, tol=1e-7, want_vectors=False)
    c = H.davidson(4, prec=api.MIXED, mode=api.DIRECT, tol=1e-7, want_vectors=False)
    assert a["converged"] and b["converged"] and c["converged"]
    np.testing.assert_allclose(b["evals"], a["evals"], atol=1e-9, rtol=0)
    np.testing.assert_allclose(c["evals"], a["evals"], atol=1e-8, rtol=0)


@pytest.mark.parametrize("mode", [api.DIRECT, api.STORED])
def test_davidson_fp32_basis_paths(fci10_strings, mode):
    """The FP32 phase keeps its basis in FP32: an FP32-only solve (vectors
    formed from the FP32 basis), mixed solves with caller vectors, Olsen's
    correction, guard roots and a small subspace cap (collapses of the FP32
    basis) all land on the FP64 roots."""
    H = fci10_strings
    ref = H.davidson(4, prec=api.FP64, mode=mode, tol=1e-8)
    assert ref["converged"]
    e64, v64 = ref["evals"], ref["evecs"]
    f = H.davidson(4, prec=api.FP32, mode=mode, tol=1e-4)
    assert f["converged"]
    np.testing.assert_allclose(f["evals"], e64, atol=1e-5, rtol=0)
    # the returned vectors (formed from the FP32 basis): their Rayleigh
    # quotients under the FP64 stored H are the roots, their residuals small
    V = f["evecs"]
    HV = H.sigma(V, api.STORED, api.FP64)
    rq = np.sum(V * HV, axis=0) / np.sum(V * V, axis=0)
    np.testing.assert_allclose(rq, e64, atol=1e-5, rtol=0)
    res = np.linalg.norm(HV - V * rq, axis=0) / np.linalg.norm(V, axis=0)
    assert np.all(res < 1e-3), res
    guess = v64 + 1e-3 * np.random.default_rng(5).standard_normal(v64.shape)
    for kw in (dict(guess=2, guess_vectors=guess), dict(precond=1), dict(extra_roots=4),
               dict(max_subspace=12)):
        c = H.davidson(4, prec=api.MIXED, mode=mode, tol=1e-7, want_vectors=False, **kw)
        assert c["converged"], kw
        np.testing.assert_allclose(c["evals"], e64, atol=1e-9, rtol=0, err_msg=str(kw))


@pytest.fixture(scope="module")
def mrci8x24_pair():
    cfg, ints, k = inputs.workload("mrci8x24")
    A = api.Hamiltonian(cfg, ints)
    B = api.Hamiltonian(cfg, ints)
    assert A.direct_engine(1) == 1
    assert B.direct_engine(2) == 2
    yield A, B, k
    A.close()
    B.close()


@pytest.mark.parametrize("k", [1, 8, 16, 24])
def test_ras_matches_records_mrci8x24(mrci8x24_pair, k):
    A, B, _ = mrci8x24_pair
    X = np.random.default_rng(100 + k).standard_normal((A.n, k))
    ya = A.sigma(X, api.DIRECT, api.FP64)
    yb = B.sigma(X, api.DIRECT, api.FP64)
    scale = np.abs(ya).max()
    assert np.abs(yb - ya).max() <= 1e-11 * scale
    yb32 = B.sigma(X.astype(np.float32), api.DIRECT, api.FP32)
    assert np.abs(yb32 - ya).max() <= 2e-5 * scale


def test_ras_davidson_mrci8x24(mrci8x24_pair):
    """The 8-root workload: FP64 roots of both engines agree, and the mixed
    solve on the string engine is within 1e-8 Eh of them."""
    A, B, k = mrci8x24_pair
    a = A.davidson(k, prec=api.FP64, mode=api.DIRECT, tol=1e-5, want_vectors=False)
    b = B.davidson(k, prec=api.FP64, mode=api.DIRECT, tol=1e-5, want_vectors=False)
    c = B.davidson(k, prec=api.MIXED, mode=api.DIRECT, tol=1e-5, switch_tol=1e-4, want_vectors=False)
    assert a["converged"] and b["converged"] and c["converged"]
    np.testing.assert_allclose(b["evals"], a["evals"], atol=1e-9, rtol=0)
    np.testing.assert_allclose(c["evals"], a["evals"], atol=1e-8, rtol=0)


def test_ras_engine_selection(fci6):
    cfg, ints, _ = fci6
    # auto picks the string engine for a full CI space
    H = api.Hamiltonian(cfg, ints)
    assert H.direct_engine() == 2
    X = np.random.default_rng(2).standard_normal((H.n, 3))
    y2 = H.sigma(X, api.DIRECT, api.FP64)
    assert H.direct_engine(1) == 1
    y1 = H.sigma(X, api.DIRECT, api.FP64)
    assert np.abs(y2 - y1).max() <= 1e-12 * np.abs(y1).max()
    assert H.direct_engine(0) == 2
    with pytest.raises(api.PsociError) as e:
        H.direct_engine(3)
    assert e.value.code == api.ERR_ARG
    H.close()
    # the ODD class and odd electron counts take the configuration engine
    for c in (inputs.fci_configs(6, 6, jsym=1, ksym=3), inputs.fci_configs(6, 5)):
        H = api.Hamiltonian(c, ints)
        assert H.direct_engine() == 1
        with pytest.raises(api.PsociError) as e:
            H.direct_engine(2)
        assert e.value.code == api.ERR_UNSUPPORTED
        H.close()
    # a list that is not complete RAS
    keep = np.ones(cfg.ncfg, bool)
    keep[cfg.ncfg // 2] = False
    part = inputs.Configs(nbf=cfg.nbf, nelec=cfg.nelec, ksym=cfg.ksym, occ=cfg.occ[keep].copy(),
                          jsym=cfg.jsym[keep].copy(), title="fci6 minus one")
    H = api.Hamiltonian(part, ints)
    assert H.direct_engine() == 1
    H.close()


def test_ras_zero_single_falls_back(fci6):
    """A configuration pair one electron apart whose single-excitation sum
    vnt[0] is exactly zero while its exchange terms are not: the reference
    skips that block (processSingleExcitation returns 0), so the string
    engine (which would not) is refused and auto takes the configuration
    engine, whose sigma matches the oracle's.  vnt[0] of a 0 -> 1 move is
    h_01 - (01|00) + sum_i occ_i (01|ii) - sum over orbitals doubly occupied
    in both configurations of (0i|i1): with h_01 = 0 and every (01|ii) = 0 it
    is exactly zero for pairs without a common doubly occupied orbital (e.g.
    (2,0,1,1,1,1) -> (1,1,1,1,1,1)), while (0c|c1) stays nonzero."""
    cfg, ints, _ = fci6
    one = ints.valone.copy()
    two = ints.valtwo.copy()
    one[inputs.pair_index(0, 1)] = 0.0
    for i in range(cfg.nbf):
        two[inputs.canonical_index(0, 1, i, i)] = 0.0
    assert two[inputs.canonical_index(0, 2, 2, 1)] != 0.0
    z = inputs.Integrals(nbf=ints.nbf, valone=one, valtwo=two, ecore=ints.ecore)
    H = api.Hamiltonian(cfg, z)
    assert H.direct_engine() == 1
    with pytest.raises(api.PsociError) as e:
        H.direct_engine(2)
    assert e.value.code == api.ERR_UNSUPPORTED
    sp = oracle.OracleSpace(cfg, z)
    sp.build_h()
    X = np.random.default_rng(4).standard_normal((H.n, 2))
    ref = sp.sigma(X)
    np.testing.assert_allclose(H.sigma(X, api.DIRECT, api.FP64), ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    H.close()
