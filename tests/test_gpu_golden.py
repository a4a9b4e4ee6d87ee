"""The GPU path against the committed fixtures (no oracle at run time):
stored H of the 246-SEF space bit for bit, a sample of rows of the
31,650-SEF space bit for bit, and determinant lists and phases for 0..10
open shells."""
import json
import os

import numpy as np
import pytest

from psoci_amd import api, inputs

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_fci6_h_matches_fixture(fci6):
    cfg, ints, _ = fci6
    f = np.load(os.path.join(GOLD, "fci6_h.npz"))
    H = api.Hamiltonian(cfg, ints)
    H.build()
    rowptr, col, val, diag = H.get_h()
    assert np.array_equal(rowptr, f["rowptr"])
    assert np.array_equal(col, f["col"])
    assert np.array_equal(val, f["val"])
    assert np.array_equal(diag, f["diag"])
    out = H.davidson(4, prec=api.FP64, mode=api.STORED, tol=1e-9)
    assert np.abs(out["evals"] - f["evals"][:4]).max() < 1e-9
    H.close()


def test_fci10_rows_match_fixture(fci10):
    cfg, ints, _ = fci10
    f = np.load(os.path.join(GOLD, "fci10_rows.npz"))
    H = api.Hamiltonian(cfg, ints)
    H.build()
    rowptr, col, val, diag = H.get_h()
    for r, d in zip(f["rows"], f["diag"]):
        a, b = rowptr[r], rowptr[r + 1]
        sel = f["row"] == r
        assert np.array_equal(col[a:b], f["col"][sel]), r
        assert np.array_equal(val[a:b], f["val"][sel]), r
        assert diag[r] == d
    H.close()


def test_dets_match_fixture():
    tab = json.load(open(os.path.join(GOLD, "dets_by_open.json")))
    for s in range(0, 11, 2):
        rec = tab[str(s)]
        occ = np.array([rec["occ"]], np.uint8)
        cfg = inputs.Configs(nbf=occ.shape[1], nelec=int(occ.sum()), ksym=1, occ=occ, jsym=np.ones(1, np.int32))
        ints = inputs.make_integrals(cfg.nbf, 1)
        H = api.Hamiltonian(cfg, ints)
        d = H.dets(0)
        want_a = [sum(1 << o for o in lst) for lst in rec["alpha"]]
        want_b = [sum(1 << o for o in lst) for lst in rec["beta"]]
        assert [int(x) for x in d["alpha"]] == want_a, s
        assert [int(x) for x in d["beta"]] == want_b, s
        assert [int(x) for x in d["phase"]] == rec["phase"], s
        H.close()


def test_sifs_integrals_path(fci6, tmp_path):
    """Integrals through a SIFS-style file give the same Hamiltonian as the
    arrays (psoci_set_integrals_sifs vs psoci_set_integrals)."""
    cfg, ints, _ = fci6
    path = str(tmp_path / "fci6.sifs")
    inputs.write_sifs(path, ints)
    A = api.Hamiltonian(cfg, ints)
    B = api.Hamiltonian(cfg, ints)
    B.set_integrals_sifs(path)
    A.build()
    B.build()
    ra, ca, va, da = A.get_h()
    rb, cb, vb, db = B.get_h()
    assert np.array_equal(ra, rb) and np.array_equal(ca, cb) and np.array_equal(va, vb) and np.array_equal(da, db)
    A.close()
    B.close()
