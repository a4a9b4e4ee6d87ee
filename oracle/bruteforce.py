"""Independent brute force -- TEST INFRASTRUCTURE ONLY.

Second-quantised spin-free Hamiltonian over spin-orbital determinants, built
from textbook Slater-Condon rules by applying creation/annihilation operators
to bit strings.  It shares no code with the oracle or the product: it pins
the block values and the excitation phases.

Conventions.  Spin-orbital index: alpha p -> p, beta p -> nbf + p.  A
determinant bit string b stands for a+_{i1} a+_{i2} ... |0> with
i1 < i2 < ... (alpha string ascending, then beta string ascending), the
ordering in which the reference's sorted determinant lists are read.  A
reference determinant with generator phase ph is (-1)^ph times that string.

H = sum_pq h_pq sum_s a+_ps a_qs + 1/2 sum_pqrs (pq|rs) sum_st a+_ps a+_rt a_st a_qs.
"""
from __future__ import annotations

import itertools

import numpy as np


def _apply(ops, det):
    """Apply operators right to left: ops = [(k, +1 create | -1 annihilate), ...]
    in left-to-right order.  Returns (sign, det) or (0, None)."""
    sign = 1
    for k, kind in reversed(ops):
        bit = 1 << k
        below = bin(det & (bit - 1)).count("1")
        if kind < 0:
            if not det & bit:
                return 0, None
            det ^= bit
        else:
            if det & bit:
                return 0, None
            det |= bit
        if below & 1:
            sign = -sign
    return sign, det


class Brute:
    def __init__(self, h: np.ndarray, g: np.ndarray):
        self.h = h  # [n, n]
        self.g = g  # [n, n, n, n] chemists' (pq|rs)
        self.nbf = h.shape[0]

    def _sp(self, k):
        return k % self.nbf, k // self.nbf

    def element(self, a: int, b: int) -> float:
        """<a|H|b> for spin-orbital bit strings a, b."""
        diff = a ^ b
        nd = bin(diff).count("1")
        if nd > 4:
            return 0.0
        nbf = self.nbf
        h, g = self.h, self.g
        occ_b = [k for k in range(2 * nbf) if b >> k & 1]
        if nd == 0:
            e = 0.0
            for k in occ_b:
                p, _ = self._sp(k)
                e += h[p, p]
            for k in occ_b:
                for l in occ_b:
                    p, s1 = self._sp(k)
                    q, s2 = self._sp(l)
                    e += 0.5 * g[p, p, q, q]
                    if s1 == s2:
                        e -= 0.5 * g[p, q, q, p]
            return e
        add = [k for k in range(2 * nbf) if (a >> k & 1) and not (b >> k & 1)]  # in a, not in b
        rem = [k for k in range(2 * nbf) if (b >> k & 1) and not (a >> k & 1)]  # in b, not in a
        if nd == 2:
            P, Q = add[0], rem[0]
            p, sp = self._sp(P)
            q, sq = self._sp(Q)
            if sp != sq:
                return 0.0
            sign, d = _apply([(P, 1), (Q, -1)], b)
            assert d == a
            e = h[p, q]
            for k in occ_b:
                if k == Q:
                    continue
                r, sr = self._sp(k)
                e += g[p, q, r, r]
                if sr == sp:
                    e -= g[p, r, r, q]
            return sign * e
        # nd == 4: two spin-orbitals differ
        P, R = add
        Q, S = rem
        e = 0.0
        # 1/2 sum (pq|rs) a+_P a+_R a_S a_Q over the 4 orderings reduces to two terms
        for (X, Y) in ((Q, S), (S, Q)):
            p, sp = self._sp(P)
            r, sr = self._sp(R)
            x, sx = self._sp(X)
            y, sy = self._sp(Y)
            if sp != sx or sr != sy:
                continue
            sign, d = _apply([(P, 1), (R, 1), (Y, -1), (X, -1)], b)
            assert d == a
            e += sign * g[p, x, r, y]
        return e


def det_bits(alpha, beta, nbf):
    b = 0
    for p in alpha:
        b |= 1 << p
    for p in beta:
        b |= 1 << (nbf + p)
    return b


def sef_vectors(dets: dict, nbf: int, nopen: int, ncp: int = 1, fractional: bool = False):
    """SEF k = (D_2k +/- D_2k+1)/sqrt(2) (D_0 for closed shells; every
    determinant by itself for odd electron counts), each reference
    determinant being (-1)^phase times its sorted bit string.  Returns a list
    of {bitstring: coefficient} dicts."""
    terms = []
    for a, b_, ph in zip(dets["alpha"], dets["beta"], dets["phase"]):
        terms.append((det_bits(a, b_, nbf), -1.0 if ph % 2 else 1.0))
    if fractional:
        return [{d: c} for d, c in terms]
    if nopen == 0:
        return [{terms[0][0]: terms[0][1]}]
    m = -1.0 if ncp == 2 else 1.0
    r = np.sqrt(0.5)
    out = []
    for k in range(len(terms) // 2):
        d0, s0 = terms[2 * k]
        d1, s1 = terms[2 * k + 1]
        v = {}
        v[d0] = v.get(d0, 0.0) + r * s0
        v[d1] = v.get(d1, 0.0) + r * m * s1
        out.append(v)
    return out


def block(brute: Brute, sefs_i, sefs_j):
    out = np.zeros((len(sefs_i), len(sefs_j)))
    cache = {}
    for i, vi in enumerate(sefs_i):
        for j, vj in enumerate(sefs_j):
            s = 0.0
            for da, ca in vi.items():
                for db, cb in vj.items():
                    key = (da, db)
                    if key not in cache:
                        cache[key] = brute.element(da, db)
                    s += ca * cb * cache[key]
            out[i, j] = s
    return out


def fci_matrix(brute: Brute, nelec: int):
    """Dense H over all spin-orbital determinants with nelec electrons."""
    nbf = brute.nbf
    dets = [sum(1 << k for k in c) for c in itertools.combinations(range(2 * nbf), nelec)]
    n = len(dets)
    H = np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1):
            H[i, j] = H[j, i] = brute.element(dets[i], dets[j])
    return H, dets


def one_rdm(c_sef, sefs_all, nbf):
    """gamma_pq = <Psi| sum_s a+_ps a_qs |Psi> from an SEF-basis vector."""
    psi = {}
    for coef, v in zip(c_sef, sefs_all):
        for d, cd in v.items():
            psi[d] = psi.get(d, 0.0) + coef * cd
    gam = np.zeros((nbf, nbf))
    for d, cd in psi.items():
        for s in range(2):
            for p in range(nbf):
                for q in range(nbf):
                    sign, nd = _apply([(p + s * nbf, 1), (q + s * nbf, -1)], d)
                    if sign and nd in psi:
                        gam[p, q] += psi[nd] * sign * cd
    return gam
