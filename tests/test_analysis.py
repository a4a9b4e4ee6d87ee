"""CI analysis off the hot path: the oracle of the energy-contribution sums
(PsociCIAnalysis.C:126-142) pinned by closed forms, and the natural-orbital /
Mulliken chain (PsociNOpopulations.C:173-729, 1084-1790) through the C ABI
(context free, no GPU) against LAPACK and the identities it must satisfy."""
import numpy as np
import pytest

import oracle
from psoci_amd import api


def two_by_two(a, b, d):
    H = np.array([[a, b], [b, d]])
    w, V = np.linalg.eigh(H)
    return H, w, V


def test_energy_contrib_two_state_closed_form():
    # for an exact eigenvector of [[a, b], [b, d]], configuration 1 contributes
    # -c1^2 (d - E) / c0^2 = b c1 / c0 = E - a, and configuration 0 likewise E - d
    H, w, V = two_by_two(-1.3, 0.21, 0.4)
    de = oracle.energy_contrib([0, 1], np.diag(H), V, w)
    for r in range(2):
        assert abs(de[1, r] - (w[r] - H[0, 0])) < 1e-13
        assert abs(de[0, r] - (w[r] - H[1, 1])) < 1e-13


def test_energy_contrib_reset_and_zeros():
    diag = np.array([1.0, 2.0, 3.0, 4.0, 5.0])
    sef_off = [0, 2, 4]  # configurations of 2, 2 and 1 SEFs
    E = -0.5
    C = np.zeros((5, 2))
    C[0, 0] = 1.0          # c^2 == 1 resets the sum to E; the next SEF still adds
    C[1, 0] = 0.25
    C[4, 1] = -1.0
    de = oracle.energy_contrib(sef_off, diag, C, [E, E])
    x2 = 0.0625
    assert de[0, 0] == E - x2 * (diag[1] - E) / (1 - x2)
    assert de[1, 0] == 0.0 and de[2, 0] == 0.0
    assert de[2, 1] == E and de[0, 1] == 0.0 and de[1, 1] == 0.0


def fci6_density():
    from psoci_amd import inputs
    cfg, ints, _ = inputs.workload("fci6")
    sp = oracle.OracleSpace(cfg, ints)
    sp.build_h()
    w, V = np.linalg.eigh(sp.dense_h())
    return cfg, sp.rdm1(V[:, 0])


def unpack(p, nbf):
    D = np.zeros((nbf, nbf))
    for i in range(nbf):
        for j in range(i + 1):
            D[i, j] = D[j, i] = p[i * (i + 1) // 2 + j]
    return D


def test_natural_orbitals_against_lapack():
    cfg, rdm = fci6_density()
    nbf = cfg.nbf
    out = api.natural_orbitals(rdm)
    occ, U = out["occ"], out["no_mo"]
    D = unpack(rdm, nbf)
    ref = np.linalg.eigvalsh(D)[::-1]
    np.testing.assert_allclose(occ, ref, atol=1e-13)
    assert abs(occ.sum() - cfg.nelec) < 1e-12
    assert np.all(np.diff(occ) <= 1e-15) and occ.min() > -1e-12 and occ.max() < 2 + 1e-12
    np.testing.assert_allclose(U @ U.T, np.eye(nbf), atol=1e-13)
    np.testing.assert_allclose(U.T @ np.diag(occ) @ U, D, atol=1e-13)
    for i in range(nbf):  # sign convention: largest-magnitude coefficient positive
        assert U[i, np.argmax(np.abs(U[i]))] > 0


def test_natural_orbitals_closed_shell():
    nbf = 6
    D = np.diag([2.0, 0.0, 2.0, 0.0, 2.0, 0.0])
    p = np.array([D[i, j] for i in range(nbf) for j in range(i + 1)])
    out = api.natural_orbitals(p)
    assert np.array_equal(out["occ"], [2.0, 2.0, 2.0, 0.0, 0.0, 0.0])
    np.testing.assert_allclose(np.abs(out["no_mo"][:3]), np.eye(nbf)[[0, 2, 4]], atol=1e-15)


def orthonormal_mos(nbf, seed=0):
    """A positive definite AO overlap S and MO rows C with C S C^T = I."""
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((nbf, nbf)) * 0.2
    S = np.eye(nbf) + A @ A.T * 0.3
    e, W = np.linalg.eigh(S)
    S_mhalf = W @ np.diag(e ** -0.5) @ W.T
    Q, _ = np.linalg.qr(rng.standard_normal((nbf, nbf)))
    C = (S_mhalf @ Q).T
    return S, C


def pack(M):
    n = M.shape[0]
    return np.array([M[i, j] for i in range(n) for j in range(i + 1)])


def test_ao_density_and_mulliken():
    cfg, rdm = fci6_density()
    nbf = cfg.nbf
    S, C = orthonormal_mos(nbf)
    np.testing.assert_allclose(C @ S @ C.T, np.eye(nbf), atol=1e-12)
    out = api.natural_orbitals(rdm, mocoef=C)
    D = unpack(rdm, nbf)
    np.testing.assert_allclose(unpack(out["dens_ao"], nbf), C.T @ D @ C, atol=1e-13)
    np.testing.assert_allclose(out["no_ao"], out["no_mo"] @ C, atol=1e-13)
    # the AO natural orbitals are S-orthonormal
    np.testing.assert_allclose(out["no_ao"] @ S @ out["no_ao"].T, np.eye(nbf), atol=1e-12)
    pop, total = api.mulliken(out["dens_ao"], pack(S))
    assert abs(total - cfg.nelec) < 1e-12  # trace(P S) = trace(D C S C^T)
    np.testing.assert_allclose(pop, unpack(out["dens_ao"], nbf) * S, atol=1e-15)
    # condensed populations partition the total
    labels = ["C1_1s", "  2C1_2p", "C1_2p", "H1_1s", "1H2_1s", "H1_2p"]
    grp, ng = api.basis_groups(labels, api.ATOMIC)
    assert ng == 3 and list(grp) == [0, 0, 0, 1, 2, 1]  # C1 < H1 < H2
    gross, bond = api.group_populations(pop, grp, ng)
    assert abs(gross.sum() - total) < 1e-12
    assert abs(np.triu(bond).sum() - total) < 1e-12 and np.all(np.tril(bond, -1) == 0)
    for g in range(ng):
        m = grp == g
        assert abs(gross[g] - pop[:, m].sum()) < 1e-13
        assert abs(bond[g, g] - pop[np.ix_(m, m)].sum()) < 1e-13


def test_label_modes():
    labels = ["  1C1s", "C1_2px", "O12_d", "__x", "7"]
    _, ng_none = api.basis_groups(labels, api.NONE)
    assert ng_none == 5
    grp, ng = api.basis_groups(labels, api.SUBSHELL)
    # leading digits and non-alphanumerics dropped: C1s, C1_2px, O12_d, x, ""
    assert ng == 5 and list(grp) == [2, 1, 3, 4, 0]
    grp, ng = api.basis_groups(labels, api.ATOMIC)
    # cut at the first '_' after the stripped prefix: C1s, C1, O12, x (its '_' is in the prefix), ""
    assert ng == 5 and list(grp) == [2, 1, 3, 4, 0]


def test_ao_file_round_trip(tmp_path):
    S, _ = orthonormal_mos(5, seed=3)
    labels = ["C1_1s", "C1_2s", "C1_2p", "H1_1s", "H2_1s"]
    p = str(tmp_path / "aoints")
    api.sifs_write_ao(p, pack(S), labels, title="test AO file")
    s2, lab2 = api.sifs_read_ao(p)
    assert np.array_equal(s2, pack(S))
    assert [x.rstrip() for x in lab2] == labels
    with pytest.raises(api.PsociError):
        api.sifs_read_ao(str(tmp_path / "missing"))
