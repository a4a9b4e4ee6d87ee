"""Restart paths on the GPU through the C ABI: the stored-Hamiltonian
checkpoint (the reference's DRA dump / --hamiltonian_restart,
PsociConstructHamiltonian.C:56-700) reproduces the built operator bit for
bit and refuses files written for other inputs or damaged on disk, and a
Davidson interrupted, written to a CI-vector file and resumed from it
(--vector_restart, driver.PSOCI.splittest.C:962-971, 1354-1358) reaches the
eigenvalues of an uninterrupted solve."""
import dataclasses

import numpy as np
import pytest

from psoci_amd import api

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def built(fci10, tmp_path_factory):
    cfg, ints, _ = fci10
    H = api.Hamiltonian(cfg, ints)
    H.build()
    path = str(tmp_path_factory.mktemp("ck") / "h.ck")
    H.save_h(path)
    yield H, path
    H.close()


def test_checkpoint_reproduces_operator(fci10, built):
    cfg, ints, _ = fci10
    H, path = built
    G = api.Hamiltonian(cfg, ints)
    assert G.load_h(path) == H.nnz_lower
    for a, b in zip(H.get_h(), G.get_h()):
        assert np.array_equal(a, b)
    assert np.array_equal(G.diag(), H.diag())
    for prec in (api.FP64, api.FP32):
        assert G.stored_shape(prec) == H.stored_shape(prec)
    X = np.random.default_rng(5).standard_normal((H.n, 16))
    assert np.array_equal(G.sigma(X, api.STORED, api.FP64), H.sigma(X, api.STORED, api.FP64))
    X32 = X.astype(np.float32)
    assert np.array_equal(G.sigma(X32, api.STORED, api.FP32), H.sigma(X32, api.STORED, api.FP32))
    a = H.davidson(4, prec=api.FP64, tol=1e-7)
    b = G.davidson(4, prec=api.FP64, tol=1e-7)
    assert a["converged"] and b["converged"]
    assert np.array_equal(a["evals"], b["evals"])
    G.close()


def test_checkpoint_refuses_other_inputs(fci6, fci10, built):
    cfg, ints, _ = fci10
    _, path = built
    ints2 = dataclasses.replace(ints, valtwo=ints.valtwo.copy())
    ints2.valtwo[7] += 1e-3
    G = api.Hamiltonian(cfg, ints2)
    with pytest.raises(api.PsociError, match="other integrals"):
        G.load_h(path)
    G.close()
    cfg6, ints6, _ = fci6
    G = api.Hamiltonian(cfg6, ints6)
    with pytest.raises(api.PsociError, match="another configuration list"):
        G.load_h(path)
    G.close()


def test_checkpoint_refuses_damaged_files(fci10, built, tmp_path):
    cfg, ints, _ = fci10
    H, path = built
    data = bytearray(open(path, "rb").read())
    G = api.Hamiltonian(cfg, ints)
    bad = tmp_path / "bad.ck"
    flipped = bytearray(data)
    flipped[len(data) // 2] ^= 0x10  # inside the values
    bad.write_bytes(bytes(flipped))
    with pytest.raises(api.PsociError, match="checksum"):
        G.load_h(str(bad))
    bad.write_bytes(bytes(data[:-100]))
    with pytest.raises(api.PsociError, match="file size"):
        G.load_h(str(bad))
    bad.write_bytes(b"PSOCIHCX" + bytes(data[8:]))
    with pytest.raises(api.PsociError, match="not a stored-Hamiltonian checkpoint"):
        G.load_h(str(bad))
    # the failed loads left no operator behind
    with pytest.raises(api.PsociError, match="build_h"):
        G.sigma(np.ones(H.n), api.STORED, api.FP64)
    G.close()


def test_resume_from_vector_file(fci10, built, tmp_path):
    H, _ = built
    full = H.davidson(4, prec=api.FP64, tol=1e-8, guess=1)
    assert full["converged"]
    part = H.davidson(4, prec=api.FP64, tol=1e-8, guess=1, max_iter=4)
    assert not part["converged"]
    path = str(tmp_path / api.restart_filename("civec.dat"))
    api.civec_write(path, part["evecs"])
    assert api.civec_info(path) == (H.n, 4, "default")
    res = H.davidson(4, prec=api.FP64, tol=1e-8, restart=path)
    assert res["converged"]
    np.testing.assert_allclose(res["evals"], full["evals"], atol=1e-9, rtol=0)
    assert res["stats"]["iterations"] < full["stats"]["iterations"]
    # converged vectors written and read back: the restart converges at once
    api.civec_write(path, res["evecs"], title="fci10 4 roots")
    again = H.davidson(4, prec=api.FP64, tol=1e-8, restart=path)
    assert again["converged"] and again["stats"]["iterations"] <= 2
    np.testing.assert_allclose(again["evals"], full["evals"], atol=1e-9, rtol=0)
    # fewer roots on file than sought: unit vectors complete the block
    api.civec_write(path, res["evecs"][:, :2])
    more = H.davidson(4, prec=api.FP64, tol=1e-8, restart=path)
    assert more["converged"]
    np.testing.assert_allclose(more["evals"], full["evals"], atol=1e-9, rtol=0)
