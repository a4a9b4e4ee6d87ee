"""Oracle package -- TEST INFRASTRUCTURE ONLY.

``liboracle.so`` (built from ``psoci_oracle.cpp`` by ``oracle/Makefile``) is a
CPU restatement of the reference PSOCI arithmetic: determinant generation and
phases (PsociDeterminants.C:1467-4469), configuration-pair blocks
(PsociHamiltonian.C:98-1516), packSefsef (PsociGAhamiltonian.C:4077-4136),
stored sigma (:6921-7269), the production Davidson schedule
(driver.PSOCI.splittest.C:895-1340) and the CIDEN 1-RDM
(PsociGAhamiltonian.C:10027-10291, 11354-11784).  ``bruteforce.py`` is an
independent second-quantised Slater-Condon check of the same blocks.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package, and only as the checker or the timed CPU baseline.

Parity status of each oracle function (what pins it):
  * canonical/pair indices     -- brute-force permutation symmetry (tests)
  * determinant lists          -- the reference's comment tables for 2, 4 and
                                  6 open shells + invariants (count, occupations,
                                  Ms, Kramers pairing, distinct dets)
  * phases                     -- brute-force inversion counting
  * blocks (values and signs)  -- second-quantised brute force, elementwise
  * stored H / eigenvalues     -- dense LAPACK + FCI spectrum containment
  * stored sigma               -- dense matrix-vector product
  * Davidson                   -- converged eigenvalues against dense LAPACK
                                  (its iteration trajectory: parity unpinned)
  * CIDEN 1-RDM                -- brute-force <Psi|E_pq|Psi> from the
                                  second-quantised determinant expansion
  * spin-orbit blocks          -- Kramers pairs of odd-electron spectra,
                                  orientation symmetry, zero integrals, the
                                  single-jsym reduction, planted mistakes;
                                  the window count and phases inside socig:
                                  parity unpinned (tests/test_spin_orbit.py)
  * transition Hamiltonian     -- equals the stored H when both lists are the
                                  same; the index-equality and entry-point
                                  rules by hand-built cases (tests/test_oracle.py)
  * CI energy contributions    -- closed forms: a 2x2 eigenproblem (each
                                  configuration's contribution is E minus the
                                  other's diagonal), the c^2 == 1 reset, zero
                                  coefficients contribute nothing
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB) or os.path.getmtime(_LIB) < os.path.getmtime(
            os.path.join(_HERE, "psoci_oracle.cpp")):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(_LIB)
        P = ctypes.c_void_p
        L.orc_space_new.restype = P
        L.orc_space_new.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_long, P, P,
                                    ctypes.c_char_p, ctypes.c_int]
        L.orc_space_free.argtypes = [P]
        L.orc_space_nsef_total.restype = ctypes.c_long
        L.orc_space_nsef_total.argtypes = [P]
        L.orc_space_info.argtypes = [P, P, P, P, P]
        L.orc_dets.restype = ctypes.c_int
        L.orc_dets.argtypes = [P, ctypes.c_long, ctypes.c_int, P, P, P, P, P, P]
        L.orc_set_integrals.argtypes = [P, P, P]
        L.orc_block.restype = ctypes.c_int
        L.orc_block.argtypes = [P, ctypes.c_long, ctypes.c_long, P]
        L.orc_jlist.restype = ctypes.c_long
        L.orc_jlist.argtypes = [P, ctypes.c_long, P]
        L.orc_build_h.restype = ctypes.c_long
        L.orc_build_h.argtypes = [P, ctypes.c_int]
        L.orc_get_h.argtypes = [P, P, P, P, P]
        L.orc_sigma.argtypes = [P, ctypes.c_int, P, P]
        L.orc_sigma_csr.argtypes = [ctypes.c_long, P, P, P, ctypes.c_int, P, P]
        L.orc_davidson.restype = ctypes.c_int
        L.orc_davidson.argtypes = [P, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, P, P, P, P]
        L.orc_rdm1.argtypes = [P, P, P]
        L.orc_transition_h.restype = ctypes.c_long
        L.orc_transition_h.argtypes = [P, P, P, P, P, ctypes.c_long]
        L.orc_set_so_mutation.argtypes = [ctypes.c_int]
        L.orc_energy_contrib.argtypes = [ctypes.c_long, P, ctypes.c_long, ctypes.c_int, P, P, P, P]
        L.orc_direct_rows.restype = ctypes.c_int
        L.orc_direct_rows.argtypes = [P, ctypes.c_long, P, ctypes.c_int, P, P, ctypes.c_int]
        L.orc_two_e_index.restype = ctypes.c_int
        L.orc_two_e_index.argtypes = [ctypes.c_int] * 2
        L.orc_canonical_index.restype = ctypes.c_int
        L.orc_canonical_index.argtypes = [ctypes.c_int] * 4
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class OracleSpace:
    """Determinant/SEF space + Hamiltonian of the oracle."""

    def __init__(self, cfg, ints=None):
        L = lib()
        self.cfg = cfg
        self.nbf = cfg.nbf
        err = ctypes.create_string_buffer(256)
        occ = np.ascontiguousarray(cfg.occ, dtype=np.uint8)
        js = np.ascontiguousarray(cfg.jsym, dtype=np.int32)
        self._h = L.orc_space_new(cfg.nbf, cfg.nelec, cfg.ksym, cfg.ncfg, _p(occ), _p(js), err, 256)
        if not self._h:
            raise ValueError(err.value.decode())
        n = cfg.ncfg
        self.nopen = np.zeros(n, np.int32)
        self.ndet = np.zeros(n, np.int32)
        self.nsef = np.zeros(n, np.int32)
        self.sef_off = np.zeros(n, np.int64)
        L.orc_space_info(self._h, _p(self.nopen), _p(self.ndet), _p(self.nsef), _p(self.sef_off))
        self.n = int(L.orc_space_nsef_total(self._h))
        self._ints = None
        if ints is not None:
            self.set_integrals(ints)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_space_free(self._h)
            self._h = None

    def set_integrals(self, ints):
        self._ints = (np.ascontiguousarray(ints.valone, np.float64), np.ascontiguousarray(ints.valtwo, np.float64))
        lib().orc_set_integrals(self._h, _p(self._ints[0]), _p(self._ints[1]))

    def dets(self, I: int, raw: bool = False):
        nd = int(self.ndet[I])
        nbf = self.nbf
        a = np.zeros((nd, nbf), np.int32)
        b = np.zeros((nd, nbf), np.int32)
        ms = np.zeros(nd, np.int32)
        ph = np.zeros(nd, np.int32)
        sa = np.zeros(nd, np.int32)
        sb = np.zeros(nd, np.int32)
        lib().orc_dets(self._h, I, int(raw), _p(a), _p(b), _p(ms), _p(ph), _p(sa), _p(sb))
        alpha = [list(r[r >= 0]) for r in a]
        beta = [list(r[r >= 0]) for r in b]
        return dict(alpha=alpha, beta=beta, ms=ms, phase=ph, suma=sa, sumb=sb)

    def block(self, I: int, J: int):
        out = np.zeros((int(self.nsef[I]), int(self.nsef[J])), np.float64)
        st = lib().orc_block(self._h, I, J, _p(out))
        return st, out

    def jlist(self, I: int):
        out = np.zeros(I + 1, np.int64)
        n = lib().orc_jlist(self._h, I, _p(out))
        return out[:n]

    def build_h(self, nthreads: int = 0):
        nnz = lib().orc_build_h(self._h, nthreads)
        if nnz < 0:
            raise RuntimeError(f"oracle build failed ({nnz})")
        self.rowptr = np.zeros(self.n + 1, np.int64)
        self.col = np.zeros(nnz, np.int32)
        self.val = np.zeros(nnz, np.float64)
        self.diag = np.zeros(self.n, np.float64)
        lib().orc_get_h(self._h, _p(self.rowptr), _p(self.col), _p(self.val), _p(self.diag))
        return nnz

    def dense_h(self):
        """Full symmetric H from the stored lower triangle."""
        n = self.n
        H = np.zeros((n, n))
        rows = np.repeat(np.arange(n), np.diff(self.rowptr))
        H[rows, self.col] = self.val
        H[self.col, rows] = self.val
        return H

    def sigma(self, X: np.ndarray):
        X = np.ascontiguousarray(X, np.float64)
        k = 1 if X.ndim == 1 else X.shape[1]
        Y = np.zeros_like(X)
        lib().orc_sigma(self._h, k, _p(X), _p(Y))
        return Y

    def davidson(self, k: int, rtol: float = 1e-6, maxit: int = 400, guess: int = 1):
        """Guess codes as the library's psoci_dav_params.guess: 1 (default)
        the reference's unit guesses on the first k SEFs; 0 unit guesses on
        the k lowest diagonal elements."""
        ev = np.zeros(k)
        vecs = np.zeros((self.n, k))
        rn = np.zeros(k)
        ns = ctypes.c_int(0)
        it = lib().orc_davidson(self._h, k, rtol, maxit, guess, _p(ev), _p(vecs), _p(rn), ctypes.byref(ns))
        return dict(evals=ev, evecs=vecs, rnorms=rn, iterations=it, nsigma=ns.value)

    def direct_rows(self, rows, X: np.ndarray, nthreads: int = 0):
        """Rows of the direct sigma for the listed configurations, packed in
        list order ([sum nsef][k])."""
        rows = np.ascontiguousarray(rows, np.int64)
        X = np.ascontiguousarray(X, np.float64)
        k = X.shape[1]
        Y = np.zeros((int(self.nsef[rows].sum()), k))
        rc = lib().orc_direct_rows(self._h, len(rows), _p(rows), k, _p(X), _p(Y), nthreads)
        if rc != 0:
            raise RuntimeError(f"oracle direct rows failed ({rc})")
        return Y

    def transition_h(self, other: "OracleSpace"):
        """Transition Hamiltonian with self's list as rows and other's as
        columns (constructGlobalTransitionHamiltonian2WAY + packSefsef):
        (rowptr, col, val) over self's SEFs.  self holds the integrals."""
        L = lib()
        nnz = L.orc_transition_h(self._h, other._h, None, None, None, 0)
        if nnz < 0:
            raise RuntimeError(f"oracle transition build failed ({nnz})")
        rowptr = np.zeros(self.n + 1, np.int64)
        col = np.zeros(nnz, np.int32)
        val = np.zeros(nnz, np.float64)
        L.orc_transition_h(self._h, other._h, _p(rowptr), _p(col), _p(val), nnz)
        return rowptr, col, val

    def rdm1(self, c: np.ndarray):
        c = np.ascontiguousarray(c, np.float64)
        D = np.zeros(self.nbf * (self.nbf + 1) // 2)
        lib().orc_rdm1(self._h, _p(c), _p(D))
        return D


def energy_contrib(sef_off, diag, C, evals):
    """deltae [ncfg][k] of analyzeEnergyContributions for C [n][k]."""
    sef_off = np.ascontiguousarray(sef_off, np.int64)
    diag = np.ascontiguousarray(diag, np.float64)
    C = np.ascontiguousarray(C, np.float64)
    if C.ndim == 1:
        C = C[:, None]
    evals = np.ascontiguousarray(evals, np.float64)
    n, k = C.shape
    out = np.zeros((len(sef_off), k))
    lib().orc_energy_contrib(len(sef_off), _p(sef_off), n, k, _p(C), _p(diag), _p(evals), _p(out))
    return out
