"""CI analysis on the GPU through the C ABI: the per-configuration energy
contributions (analyzeEnergyContributions, PsociCIAnalysis.C:126-142) are
bit-identical to the oracle's sequential sums, and the natural orbitals of a
converged root's density (psoci_rdm -> psoci_natural_orbitals) carry the
electron count."""
import numpy as np
import pytest

import oracle
from psoci_amd import api, inputs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["fci10", "mrci8x24"])
def test_energy_contrib_bit_exact(name):
    cfg, ints, k = inputs.workload(name)
    H = api.Hamiltonian(cfg, ints)
    mode = api.STORED if name == "fci10" else api.DIRECT
    if mode == api.STORED:
        H.build()
    res = H.davidson(k, prec=api.FP64, mode=mode, tol=1e-5)
    assert res["converged"]
    C, E = res["evecs"], res["evals"]
    got = H.energy_contrib(C, E)
    ref = oracle.energy_contrib(H.info.sef_offset, H.diag(), C, E)
    assert np.array_equal(got, ref)
    # a unit vector on one SEF: its configuration gets E exactly, the rest 0
    i = int(H.info.sef_offset[min(5, cfg.ncfg - 1)])
    e1 = np.zeros((H.n, 1))
    e1[i, 0] = 1.0
    de = H.energy_contrib(e1, [-3.25])
    I = int(np.searchsorted(H.info.sef_offset, i, side="right") - 1)
    assert de[I, 0] == -3.25 and np.count_nonzero(de) == 1
    report = api.analyze_energy_contributions(H, C[:, :2], E[:2], thresh=1e-3, ecore=ints.ecore)
    for r, rep in enumerate(report):
        vals = [c["deltae"] for c in rep["configs"]]
        assert all(abs(v) >= 1e-3 for v in vals)
        assert all(abs(a) >= abs(b) for a, b in zip(vals, vals[1:]))
        assert sum(b["count"] for b in rep["distribution"]) == cfg.ncfg
    H.close()


def test_natural_orbitals_of_a_root(fci10):
    cfg, ints, _ = fci10
    H = api.Hamiltonian(cfg, ints)
    H.build()
    res = H.davidson(2, prec=api.FP64, tol=1e-7)
    g1, _, _ = H.rdm(res["evecs"][:, 0])
    rdm = api.rdm1_from_grad(g1)
    sp = oracle.OracleSpace(cfg, ints)
    np.testing.assert_allclose(rdm, sp.rdm1(res["evecs"][:, 0]), atol=1e-11)
    out = api.natural_orbitals(rdm)
    assert abs(out["occ"].sum() - cfg.nelec) < 1e-10
    assert out["occ"].max() < 2 + 1e-10 and out["occ"].min() > -1e-10
    H.close()
