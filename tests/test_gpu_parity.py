"""GPU parity through the C ABI: the HIP path against the oracle on the same
seeded inputs.  Integer/pattern results are compared exactly, FP64 elements
bit for bit (np.array_equal), sigma within 1e-12 relative, eigenvalues
within 1e-9 Eh."""
import numpy as np
import pytest

import oracle
from psoci_amd import api, inputs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu_fci6(fci6):
    cfg, ints, _ = fci6
    H = api.Hamiltonian(cfg, ints)
    H.build()
    yield H
    H.close()


@pytest.fixture(scope="module")
def gpu_fci10(fci10):
    cfg, ints, _ = fci10
    H = api.Hamiltonian(cfg, ints)
    H.build()
    yield H
    H.close()


def test_space_and_dets_match_oracle(gpu_fci6, oracle_fci6):
    H, sp = gpu_fci6, oracle_fci6
    assert H.n == sp.n == 246
    np.testing.assert_array_equal(H.info.nsef, sp.nsef)
    np.testing.assert_array_equal(H.info.sef_offset, sp.sef_off)
    for I in range(H.cfg.ncfg):
        d = H.dets(I)
        o = sp.dets(I)
        for k in range(len(o["alpha"])):
            assert int(d["alpha"][k]) == sum(1 << p for p in o["alpha"][k])
            assert int(d["beta"][k]) == sum(1 << p for p in o["beta"][k])
        np.testing.assert_array_equal(d["ms"], o["ms"])
        np.testing.assert_array_equal(d["phase"], o["phase"])


def test_blocks_bit_identical_fci6(gpu_fci6, oracle_fci6):
    H, sp = gpu_fci6, oracle_fci6
    rng = np.random.default_rng(5)
    rows = list(range(0, H.cfg.ncfg, 7)) + [H.cfg.ncfg - 1]
    n = 0
    for I in rows:
        js = sp.jlist(I)
        for J in rng.choice(js, min(12, len(js)), replace=False):
            st, ref = sp.block(I, int(J))
            st2, out = H.block(I, int(J))
            assert st == st2
            assert np.array_equal(ref, out), (I, J)
            n += 1
    assert n > 100


@pytest.mark.parametrize("which", ["fci6", "fci10"])
def test_stored_h_matches_reference_pattern(which, gpu_fci6, gpu_fci10, oracle_fci6, oracle_fci10):
    H, sp = (gpu_fci6, oracle_fci6) if which == "fci6" else (gpu_fci10, oracle_fci10)
    assert H.nnz_lower == len(sp.val)
    rowptr, col, val, diag = H.get_h()
    np.testing.assert_array_equal(rowptr, sp.rowptr)
    np.testing.assert_array_equal(col, sp.col)
    assert np.array_equal(val, sp.val)  # bit-identical elements
    assert np.array_equal(diag, sp.diag)
    assert np.array_equal(H.diag(), sp.diag)


@pytest.mark.parametrize("k", [1, 4, 16, 5, 24, 32])
def test_stored_sigma_fp64_and_fp32(k, gpu_fci10, oracle_fci10):
    H, sp = gpu_fci10, oracle_fci10
    rng = np.random.default_rng(k)
    X = rng.standard_normal((H.n, k))
    ref = sp.sigma(X)
    Y = H.sigma(X, api.STORED, api.FP64)
    # summation order differs from the reference's loop: tolerance 1e-12 relative
    assert np.abs(Y - ref).max() <= 1e-12 * np.abs(ref).max()
    Y32 = H.sigma(X.astype(np.float32), api.STORED, api.FP32)
    assert np.abs(Y32 - ref).max() <= 2e-5 * np.abs(ref).max()


def test_davidson_fp64_fci6_against_lapack(gpu_fci6, oracle_fci6):
    H = gpu_fci6
    w = np.linalg.eigvalsh(oracle_fci6.dense_h())
    out = H.davidson(4, prec=api.FP64, tol=1e-8)
    assert out["converged"]
    np.testing.assert_allclose(out["evals"], w[:4], atol=1e-9, rtol=0)
    V = out["evecs"]
    R = oracle_fci6.sigma(V) - V * out["evals"]
    assert np.linalg.norm(R, axis=0).max() < 1e-7
    np.testing.assert_allclose(V.T @ V, np.eye(4), atol=1e-10)


def test_davidson_reference_guess_fci10(gpu_fci10, oracle_fci10):
    """Unit guesses on the first k SEFs (the reference's guess) converge to
    the same roots as the oracle's reference-schedule Davidson."""
    H, sp = gpu_fci10, oracle_fci10
    ref = sp.davidson(4, rtol=1e-7)
    out = H.davidson(4, prec=api.FP64, tol=1e-7, guess=1)
    assert out["converged"]
    np.testing.assert_allclose(out["evals"], ref["evals"], atol=1e-9, rtol=0)


def test_mixed_precision_matches_fp64_fci10(gpu_fci10):
    H = gpu_fci10
    a = H.davidson(8, prec=api.FP64, tol=1e-6)
    b = H.davidson(8, prec=api.MIXED, tol=1e-6, switch_tol=1e-4)
    assert a["converged"] and b["converged"]
    assert b["stats"]["switch_iter"] > 0
    np.testing.assert_allclose(b["evals"], a["evals"], atol=1e-8, rtol=0)


def test_error_codes_not_faults(gpu_fci6):
    H = gpu_fci6
    with pytest.raises(api.PsociError) as e:
        H.block(10**6, 0)
    assert e.value.code == api.ERR_ARG
    # configurations of different spatial symmetry (spin-orbit blocks) run
    # stored only: the direct sigma refuses them
    mixed = inputs.fci_configs(4, 4)
    mixed.jsym[1] = 2
    Hm = api.Hamiltonian(mixed, inputs.make_integrals(4, 1))
    with pytest.raises(api.PsociError) as e:
        Hm.sigma(np.ones(Hm.n), api.DIRECT)
    assert e.value.code == api.ERR_UNSUPPORTED
    Hm.close()
    cfg, ints, _ = inputs.workload("fci6")
    H2 = api.Hamiltonian(cfg, ints)
    with pytest.raises(api.PsociError) as e:
        H2.sigma(np.ones(H2.n), api.STORED)
    assert e.value.code == api.ERR_STATE
    H2.close()


def test_olsen_preconditioner_same_roots_fci10(gpu_fci10):
    """precond 1 (the diagonal preconditioner with Olsen's correction)
    converges to the roots of the reference's plain diagonal one."""
    H = gpu_fci10
    a = H.davidson(8, prec=api.FP64, tol=1e-7, precond=0, want_vectors=False)
    b = H.davidson(8, prec=api.FP64, tol=1e-7, precond=1, want_vectors=False)
    c = H.davidson(8, prec=api.MIXED, tol=1e-7, precond=1, want_vectors=False)
    assert a["converged"] and b["converged"] and c["converged"]
    np.testing.assert_allclose(b["evals"], a["evals"], atol=1e-9, rtol=0)
    np.testing.assert_allclose(c["evals"], a["evals"], atol=1e-8, rtol=0)
    with pytest.raises(api.PsociError) as e:
        H.davidson(2, prec=api.FP64, precond=2)
    assert e.value.code == api.ERR_ARG


@pytest.mark.parametrize("mode", [api.STORED, api.DIRECT])
def test_select_rules_same_roots_fci10(gpu_fci10, mode):
    """select_rule 0 (the padded block width's free columns take the
    corrections of guard roots and of converged wanted roots) converges to
    the roots of the reference's rule 1 (one correction per unconverged
    root), with and without guard roots, FP64 and mixed."""
    H = gpu_fci10
    ref = H.davidson(6, prec=api.FP64, mode=mode, tol=1e-9, select_rule=1, want_vectors=False)
    assert ref["converged"]
    for prec, extra in ((api.FP64, 0), (api.FP64, 3), (api.MIXED, 0), (api.MIXED, 3)):
        a = H.davidson(6, prec=prec, mode=mode, tol=1e-7, select_rule=0, extra_roots=extra, want_vectors=False)
        b = H.davidson(6, prec=prec, mode=mode, tol=1e-7, select_rule=1, extra_roots=extra, want_vectors=False)
        assert a["converged"] and b["converged"]
        tol = 1e-9 if prec == api.FP64 else 1e-8
        np.testing.assert_allclose(a["evals"], ref["evals"], atol=tol, rtol=0)
        np.testing.assert_allclose(b["evals"], ref["evals"], atol=tol, rtol=0)
    with pytest.raises(api.PsociError) as e:
        H.davidson(2, prec=api.FP64, mode=mode, select_rule=2)
    assert e.value.code == api.ERR_ARG


@pytest.mark.parametrize("prec", [api.FP64, api.MIXED])
def test_graph_steps_equal_eager_steps_fci10(gpu_fci10, prec):
    """The stored-sigma steps captured as hipGraphs and replayed (many
    replays per solve, with eager launches and host work between them) give
    the eager launches' solve bit for bit: same eigenvalues, vectors and
    iteration count."""
    H = gpu_fci10
    a = H.davidson(8, prec=prec, tol=1e-7, graph=0)
    b = H.davidson(8, prec=prec, tol=1e-7, graph=1)
    assert a["converged"] and b["converged"]
    assert a["stats"]["graph_launches"] >= a["stats"]["iterations"] and b["stats"]["graph_launches"] == 0
    assert a["stats"]["iterations"] == b["stats"]["iterations"]
    assert np.array_equal(a["evals"], b["evals"]) and np.array_equal(a["evecs"], b["evecs"])


@pytest.mark.parametrize("force", ["0", "1"])
def test_stored_row_buckets(fci10, oracle_fci10, force, monkeypatch):
    """Row-length buckets of the stored sigma (fci10: 45 rows shorter than
    64, the rest 256-1023 long): the bucketed launches and the single one
    both give the oracle's sigma; a checkpoint carries the choice, so the
    loaded operator's sigma is the built one's bit for bit."""
    import os
    import tempfile
    monkeypatch.setenv("PSOCI_SPMM_BUCKETS", force)
    cfg, ints, _ = fci10
    H = api.Hamiltonian(cfg, ints)
    H.build()
    info = H.stored_buckets(api.FP32)
    assert info["rows"] == [45, 0, 31605, 0, 0]
    assert info["used"] == (force == "1") and H.stored_buckets(api.FP64)["used"] == (force == "1")
    X = np.random.default_rng(7).standard_normal((H.n, 12))
    ref = oracle_fci10.sigma(X)
    Y = H.sigma(X, api.STORED, api.FP64)
    assert np.abs(Y - ref).max() <= 1e-12 * np.abs(ref).max()
    Y32 = H.sigma(X.astype(np.float32), api.STORED, api.FP32)
    assert np.abs(Y32 - ref).max() <= 2e-5 * np.abs(ref).max()
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "h.bin")
        H.save_h(path)
        monkeypatch.delenv("PSOCI_SPMM_BUCKETS")
        G = api.Hamiltonian(cfg, ints)
        G.load_h(path)
        assert G.stored_buckets(api.FP32)["used"] == info["used"]
        assert np.array_equal(G.sigma(X, api.STORED, api.FP64), Y)
        G.close()
    H.close()
