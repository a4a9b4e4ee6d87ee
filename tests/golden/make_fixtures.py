"""Writes the regression fixtures of this directory from the oracle (the C++
restatement of the reference CPU path) and checks them against the
independent brute force before writing.

    python tests/golden/make_fixtures.py

Outputs (data only: inputs are regenerated from seeds by psoci_amd.inputs):
  fci6_h.npz       stored lower triangle of H for the 6-orbital / 6-electron
                   FCI space (packSefsef layout: rowptr, col, val, diag) and
                   its 246 eigenvalues (dense LAPACK)
  fci10_rows.npz   a sample of stored rows of the 10-orbital / 8-electron
                   FCI space (row index, its entries and diagonal)
  dets_by_open.json  determinant lists (alpha, beta orbitals) and phases of
                   one configuration per open-shell count 0..10, with doubly
                   occupied orbitals interleaved between the open shells
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from oracle import bruteforce as bf  # noqa: E402
from psoci_amd import inputs  # noqa: E402

FCI10_ROWS = [0, 1, 2, 777, 4096, 15000, 22222, 31649]


def single_config(nopen: int, ndoub: int = 2):
    nbf = nopen + 2 * ndoub + 1
    occ = np.zeros((1, nbf), np.uint8)
    dbl = list(range(1, 1 + 2 * ndoub, 2))[:ndoub]
    rest = [o for o in range(nbf) if o not in dbl][:nopen]
    occ[0, dbl] = 2
    occ[0, rest] = 1
    return inputs.Configs(nbf=nbf, nelec=2 * ndoub + nopen, ksym=1, occ=occ, jsym=np.ones(1, np.int32))


def main():
    oracle.build()
    # fci6: stored H and spectrum; the spectrum must lie in the brute-force
    # FCI spectrum of the same integrals (924 determinants)
    cfg, ints, _ = inputs.workload("fci6")
    sp = oracle.OracleSpace(cfg, ints)
    sp.build_h(1)
    H = sp.dense_h()
    ev = np.linalg.eigvalsh(H)
    h, g = ints.dense()
    full = np.linalg.eigvalsh(bf.fci_matrix(bf.Brute(h, g), cfg.nelec)[0])
    gap = np.abs(ev[:, None] - full[None, :]).min(axis=1).max()
    assert gap < 1e-9, gap
    np.savez(os.path.join(HERE, "fci6_h.npz"), rowptr=sp.rowptr, col=sp.col, val=sp.val, diag=sp.diag, evals=ev)

    cfg, ints, _ = inputs.workload("fci10")
    sp = oracle.OracleSpace(cfg, ints)
    sp.build_h(8)
    rows, cols, vals = [], [], []
    for r in FCI10_ROWS:
        a, b = sp.rowptr[r], sp.rowptr[r + 1]
        rows.append(np.full(b - a, r, np.int64))
        cols.append(sp.col[a:b])
        vals.append(sp.val[a:b])
    np.savez(os.path.join(HERE, "fci10_rows.npz"), rows=np.array(FCI10_ROWS, np.int64),
             row=np.concatenate(rows), col=np.concatenate(cols), val=np.concatenate(vals),
             diag=sp.diag[FCI10_ROWS])

    out = {"source": "oracle determinant generators (reference EVEN generators restated); one configuration "
                     "per open-shell count, doubly occupied orbitals at 1 and 3 (0-based)"}
    for s in range(0, 11, 2):
        c = single_config(s)
        sp = oracle.OracleSpace(c)
        d = sp.dets(0)
        out[str(s)] = {"occ": [int(x) for x in c.occ[0]],
                       "alpha": [[int(x) for x in a] for a in d["alpha"]],
                       "beta": [[int(x) for x in b] for b in d["beta"]],
                       "phase": [int(x) for x in d["phase"]]}
    with open(os.path.join(HERE, "dets_by_open.json"), "w") as f:
        json.dump(out, f, indent=0)
    print("fixtures written")


if __name__ == "__main__":
    main()
