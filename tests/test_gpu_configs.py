"""psoci_set_configs_file on the GPU: a written configuration list read back
in file order gives the same space and the same direct sigma as
psoci_set_configs on the arrays; the sorted reader (PsociConfigs.C:143-300)
gives the same space permuted by configuration, so its sigma is the
file-order sigma with rows and columns permuted in SEF blocks; a malformed
file leaves its message, with the line, on the context."""
import numpy as np
import pytest

from psoci_amd import api, inputs

pytestmark = pytest.mark.gpu


def sef_permutation(info_a, perm_cfg, info_b):
    """SEF indices in space a of the SEFs of space b, whose configuration c is
    configuration perm_cfg[c] of space a."""
    idx = []
    for c, a in enumerate(perm_cfg):
        assert info_b.nsef[c] == info_a.nsef[a]
        idx.extend(range(info_a.sef_offset[a], info_a.sef_offset[a] + info_a.nsef[a]))
    return np.array(idx)


def test_set_configs_file_file_order_and_sorted(tmp_path):
    cfg = inputs.mrcisd_configs(4, 6, 4)
    rng = np.random.default_rng(5)
    perm = rng.permutation(cfg.ncfg)
    shuffled = inputs.Configs(nbf=cfg.nbf, nelec=cfg.nelec, ksym=cfg.ksym, occ=cfg.occ[perm], jsym=cfg.jsym[perm],
                              title="shuffled MR-CISD 4x6")
    path = str(tmp_path / "configs")
    inputs.write_configs(path, shuffled)
    ints = inputs.make_integrals(10, 77, n_internal=4)

    A = api.Hamiltonian(shuffled, ints)
    B = api.Hamiltonian(cfg, ints)
    B.set_configs_file(path)
    assert B.n == A.n
    np.testing.assert_array_equal(B.info.nsef, A.info.nsef)
    X = rng.standard_normal((A.n, 3))
    ya = A.sigma(X, api.DIRECT, api.FP64)
    yb = B.sigma(X, api.DIRECT, api.FP64)
    assert np.abs(yb - ya).max() <= 1e-13 * np.abs(ya).max()

    S = api.Hamiltonian(cfg, ints)
    S.set_configs_file(path, sorted=True)
    keys = [(int(j), "".join(map(str, r))) for j, r in zip(shuffled.jsym, shuffled.occ)]
    order = sorted(range(len(keys)), key=lambda c: keys[c])
    idx = sef_permutation(A.info, order, S.info)
    ys = S.sigma(X[idx], api.DIRECT, api.FP64)
    assert np.abs(ys - ya[idx]).max() <= 1e-12 * np.abs(ya).max()

    bad = tmp_path / "bad"
    bad.write_text("t\n10 4 1 2 4\n1 2200000000\n1 22000000001\n")
    with pytest.raises(api.PsociError) as e:
        S.set_configs_file(str(bad))
    assert e.value.code == api.ERR_ARG and "line 4" in str(e.value)
    for H in (A, B, S):
        H.close()


def test_from_files_equals_array_input(fci6, tmp_path):
    """Hamiltonian.from_files (psoci_set_configs_file + psoci_set_integrals_sifs,
    the reference's input path) builds the same stored H as array input."""
    cfg, ints, _ = fci6
    cpath, spath = str(tmp_path / "configs"), str(tmp_path / "moints")
    inputs.write_configs(cpath, cfg)
    inputs.write_sifs(spath, ints)
    A = api.Hamiltonian(cfg, ints)
    B = api.Hamiltonian.from_files(cpath, spath)
    assert B.n == A.n and A.build() == B.build()
    for a, b in zip(A.get_h(), B.get_h()):
        np.testing.assert_array_equal(a, b)
    A.close()
    B.close()
