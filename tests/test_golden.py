"""Regression fixtures (tests/golden/, written by make_fixtures.py from the
oracle after a brute-force check): the oracle still reproduces them bit for
bit, and the fci6 spectrum they hold lies in the brute-force FCI spectrum."""
import json
import os

import numpy as np

import oracle
from oracle import bruteforce as bf
from psoci_amd import inputs

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def test_fci6_stored_h_fixture(fci6, oracle_fci6):
    f = np.load(os.path.join(GOLD, "fci6_h.npz"))
    sp = oracle_fci6
    assert np.array_equal(sp.rowptr, f["rowptr"])
    assert np.array_equal(sp.col, f["col"])
    assert np.array_equal(sp.val, f["val"])
    assert np.array_equal(sp.diag, f["diag"])
    cfg, ints, _ = fci6
    h, g = ints.dense()
    full = np.linalg.eigvalsh(bf.fci_matrix(bf.Brute(h, g), cfg.nelec)[0])
    ev = f["evals"]
    assert len(ev) == sp.n
    assert np.abs(ev[:, None] - full[None, :]).min(axis=1).max() < 1e-9


def test_fci10_rows_fixture(oracle_fci10):
    f = np.load(os.path.join(GOLD, "fci10_rows.npz"))
    sp = oracle_fci10
    for r, d in zip(f["rows"], f["diag"]):
        a, b = sp.rowptr[r], sp.rowptr[r + 1]
        sel = f["row"] == r
        assert np.array_equal(sp.col[a:b], f["col"][sel])
        assert np.array_equal(sp.val[a:b], f["val"][sel])
        assert sp.diag[r] == d


def test_dets_by_open_fixture():
    tab = json.load(open(os.path.join(GOLD, "dets_by_open.json")))
    for s in range(0, 11, 2):
        rec = tab[str(s)]
        occ = np.array([rec["occ"]], np.uint8)
        nel = int(occ.sum())
        cfg = inputs.Configs(nbf=occ.shape[1], nelec=nel, ksym=1, occ=occ, jsym=np.ones(1, np.int32))
        d = oracle.OracleSpace(cfg).dets(0)
        assert [list(map(int, a)) for a in d["alpha"]] == rec["alpha"], s
        assert [list(map(int, b)) for b in d["beta"]] == rec["beta"], s
        assert [int(p) for p in d["phase"]] == rec["phase"], s
        assert len(rec["alpha"]) == (1 << max(s - 1, 0))
