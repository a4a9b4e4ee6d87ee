"""Seeded synthetic inputs in the reference's data layouts.

The reference reads integrals from a Columbus SIFS file into two packed
arrays (PsociIntegrals.C:426-540, 623-762):

* ``valone[NNDAF(nbf)]``: one-electron h_pq at the 1-based lower-triangle index
  ``max(p,q)*(max(p,q)-1)/2 + min(p,q)`` (p, q 1-based), stored 0-based;
* ``valtwo[NNDAF(NNDAF(nbf))]``: two-electron (pq|rs) at the canonical 8-fold
  index (``PsociIntegrals::canonicalIndex``, PsociIntegrals.C:892-901).

and configurations from a text file (PsociConfigs.C:143-245): a title line, a
line ``nbf nelec ksym ncfg nopen`` and one ``jsym occstring`` line per
configuration, whose file order defines the SEF numbering
(``readConfigsToGA``, PsociConfigs.C:336-495).

This module makes such inputs from a seed, with the distributions of the
benchmark workloads: h_pp sorted uniform[-2, 0.5] Eh (external h_pp
uniform[0.5, 3]); h_pq ~ N(0, 0.05^2); (pp|qq) uniform[0.3, 0.8] (including
(pp|pp)); exchange (pq|qp) uniform[0.01, 0.1]; other (pq|rs) ~ N(0, 0.01^2) with
8-fold symmetry.  The spin-orbit coupling is drawn separately (N(0, 0.005^2))
and returned but not applied: in a C1 double group every configuration has
the same symmetry, so the reference never reaches its spin-orbit branch
(PsociHamiltonian.C:223-238).
"""
from __future__ import annotations

import builtins
import itertools
from dataclasses import dataclass, field
from math import comb

import numpy as np


def nndaf(i: int) -> int:
    return i * (i + 1) // 2


def pair_index(p: int, q: int) -> int:
    """0-based packed index of (p, q), p, q 0-based (TwoEIndex - 1)."""
    a, b = (p, q) if p >= q else (q, p)
    return a * (a + 1) // 2 + b


def canonical_index(p: int, q: int, r: int, s: int) -> int:
    """0-based canonical index of (pq|rs) (canonicalIndex - 1)."""
    ij = pair_index(p, q) + 1
    kl = pair_index(r, s) + 1
    a, b = (ij, kl) if ij >= kl else (kl, ij)
    return (a - 1) * a // 2 + b - 1


@dataclass
class Integrals:
    nbf: int
    valone: np.ndarray  # float64 [NNDAF(nbf)]
    valtwo: np.ndarray  # float64 [NNDAF(NNDAF(nbf))]
    ecore: float = 0.0
    so: np.ndarray | None = None  # float64 [NNDAF(nbf)], not applied (C1)

    def h(self, p, q):
        return self.valone[pair_index(p, q)]

    def g(self, p, q, r, s):
        return self.valtwo[canonical_index(p, q, r, s)]

    def dense(self):
        """h[p,q] and (pq|rs) as dense arrays (for brute-force checks)."""
        n = self.nbf
        h = np.zeros((n, n))
        for p in range(n):
            for q in range(n):
                h[p, q] = self.h(p, q)
        pairs = np.array([[pair_index(p, q) for q in range(n)] for p in range(n)])
        a = pairs[:, :, None, None] + 1
        b = pairs[None, None, :, :] + 1
        hi = np.maximum(a, b)
        lo = np.minimum(a, b)
        g = self.valtwo[(hi - 1) * hi // 2 + lo - 1]
        return h, g


def make_integrals(nbf: int, seed: int, n_internal: int | None = None) -> Integrals:
    """Seeded integrals.  Orbitals >= n_internal are external (h_pp in [0.5, 3])."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nint = nbf if n_internal is None else n_internal
    diag_int = np.sort(rng.uniform(-2.0, 0.5, size=nint))
    diag_ext = np.sort(rng.uniform(0.5, 3.0, size=nbf - nint))
    hdiag = np.concatenate([diag_int, diag_ext])
    n1 = nndaf(nbf)
    valone = np.empty(n1)
    off = rng.normal(0.0, 0.05, size=n1)
    for p in range(nbf):
        for q in range(p + 1):
            k = pair_index(p, q)
            valone[k] = hdiag[p] if p == q else off[k]
    # two-electron: classify every canonical (ij|kl), ij >= kl
    n2 = nndaf(n1)
    pi, pj = np.tril_indices(nbf)  # pair k -> (i >= j), k = i(i+1)/2 + j
    order = np.argsort(pi * (pi + 1) // 2 + pj)
    pi, pj = pi[order], pj[order]
    IJ, KL = np.tril_indices(n1)
    idx = IJ * (IJ + 1) // 2 + KL
    same_pair = IJ == KL
    i_diag = pi[IJ] == pj[IJ]
    k_diag = pi[KL] == pj[KL]
    coul = rng.uniform(0.3, 0.8, size=n2)
    exch = rng.uniform(0.01, 0.1, size=n2)
    other = rng.normal(0.0, 0.01, size=n2)
    vals = np.where(i_diag & k_diag, coul[idx], np.where(same_pair, exch[idx], other[idx]))
    valtwo = np.empty(n2)
    valtwo[idx] = vals
    so = rng.normal(0.0, 0.005, size=n1)
    return Integrals(nbf=nbf, valone=valone, valtwo=valtwo, ecore=0.0, so=so)


@dataclass
class Configs:
    nbf: int
    nelec: int
    ksym: int
    occ: np.ndarray  # uint8 [ncfg][nbf], values 0/1/2, file order
    jsym: np.ndarray  # int32 [ncfg]
    title: str = "synthetic"
    nopen_max: int = field(default=0)

    def __post_init__(self):
        self.occ = np.ascontiguousarray(self.occ, dtype=np.uint8)
        self.jsym = np.ascontiguousarray(self.jsym, dtype=np.int32)
        if self.nopen_max == 0 and len(self.occ):
            self.nopen_max = int((self.occ == 1).sum(axis=1).max())

    @property
    def ncfg(self) -> int:
        return int(self.occ.shape[0])

    def nopen(self) -> np.ndarray:
        return (self.occ == 1).sum(axis=1).astype(np.int32)

    def nsef(self) -> np.ndarray:
        s = self.nopen()
        h = np.maximum(s - 1, 0) // 2
        return (1 << h) * (1 << h)

    def ndet(self) -> np.ndarray:
        return 1 << np.maximum(self.nopen() - 1, 0)


def _occupations(norb: int, nel: int):
    """All 0/1/2 occupation tuples over norb orbitals with nel electrons, in
    lexicographic order of the tuple read as digits (highest first)."""
    out = []
    for t in itertools.product((2, 1, 0), repeat=norb):
        if sum(t) == nel:
            out.append(t)
    return out


def fci_configs(nbf: int, nelec: int, jsym: int = 1, ksym: int = 1) -> Configs:
    occ = np.array(_occupations(nbf, nelec), dtype=np.uint8).reshape(-1, nbf)
    return Configs(nbf=nbf, nelec=nelec, ksym=ksym, occ=occ, jsym=np.full(len(occ), jsym, np.int32),
                   title=f"FCI {nbf} orbitals {nelec} electrons")


def _ext_patterns(next_: int, nel: int):
    """External occupation patterns with nel (0, 1 or 2) electrons."""
    if nel == 0:
        return [()]
    pats = []
    if nel == 1:
        for a in range(next_):
            pats.append(((a, 1),))
    elif nel == 2:
        for a in range(next_):
            pats.append(((a, 2),))
        for b in range(next_):
            for a in range(b):
                pats.append(((a, 1), (b, 1)))
    return pats


def mrcisd_configs(n_internal: int, n_external: int, nelec: int, max_external: int = 2,
                   jsym: int = 1, ksym: int = 1) -> Configs:
    """Internal full CI plus every configuration with at most max_external
    electrons in the external orbitals.  Order: by number of external
    electrons, then external pattern, then internal occupation."""
    nbf = n_internal + n_external
    blocks = []
    for ne in range(0, max_external + 1):
        internal = np.array(_occupations(n_internal, nelec - ne), dtype=np.uint8).reshape(-1, n_internal)
        if len(internal) == 0:
            continue
        for pat in _ext_patterns(n_external, ne):
            ext = np.zeros(n_external, dtype=np.uint8)
            for a, v in pat:
                ext[a] = v
            blocks.append(np.hstack([internal, np.broadcast_to(ext, (len(internal), n_external))]))
    occ = np.ascontiguousarray(np.vstack(blocks), dtype=np.uint8)
    return Configs(nbf=nbf, nelec=nelec, ksym=ksym, occ=occ, jsym=np.full(len(occ), jsym, np.int32),
                   title=f"MR-CISD {n_internal}+{n_external} orbitals {nelec} electrons")


def write_configs(path: str, cfg: Configs) -> None:
    """Write the reference's configuration file format (PsociConfigs.C:143-245)."""
    digits = (np.asarray(cfg.occ, np.uint8) + ord("0")).astype(np.uint8)
    with open(path, "w") as f:
        f.write(cfg.title + "\n")
        f.write(f"{cfg.nbf} {cfg.nelec} {cfg.ksym} {cfg.ncfg} {cfg.nopen_max}\n")
        f.writelines(f"{int(js)} {row.tobytes().decode()}\n" for js, row in zip(cfg.jsym, digits))


class ConfigFileError(ValueError):
    def __init__(self, msg: str, line: int):
        super().__init__(f"line {line}: {msg}")
        self.line = line


def read_configs(path: str, sorted: bool = False) -> Configs:
    """Read a configuration file by the reference's rules: file order
    (PsociConfigs::readConfigsToGA, PsociConfigs.C:336-495), or with
    sorted=True in std::set<pair<jsym, string>> order with duplicates dropped
    (PsociConfigs::readConfigs, PsociConfigs.C:143-300).  Every line after
    the header is a configuration line; the reported line number counts the
    title and header lines; the count must match the header unless
    duplicates were dropped; nopen_max is the largest open-shell count found
    (the reference adjusts the header's value down to it)."""
    with open(path, newline="") as f:
        lines = f.read().split("\n")
    if lines and lines[-1] == "":
        lines.pop()
    lines = [ln[:-1] if ln.endswith("\r") else ln for ln in lines]
    if not lines:
        raise ConfigFileError("missing title", 1)
    title = lines[0]
    try:
        nbf, nelec, ksym, ncfg, nopen = (int(x) for x in lines[1].split()[:5])
    except (IndexError, ValueError):
        raise ConfigFileError("expected 5 integers: nbf nelec ksym ncfg nopen", 2) from None
    occ, js = [], []
    for n, line in enumerate(lines[2:], start=3):
        parts = line.split()
        if len(parts) < 2:
            raise ConfigFileError("expected 'jsym occupation'", n)
        try:
            sym = int(parts[0])
        except ValueError:
            raise ConfigFileError(f"bad symmetry '{parts[0]}'", n) from None
        s = parts[1]
        if len(s) != nbf:
            raise ConfigFileError(f"configuration length {len(s)} != nbf {nbf}", n)
        if any(ch not in "012" for ch in s):
            raise ConfigFileError("occupations must be 0, 1 or 2", n)
        row = [int(ch) for ch in s]
        if row.count(1) > nopen:
            raise ConfigFileError(f"{row.count(1)} open shells, the header allows {nopen}", n)
        if sum(row) != nelec:
            raise ConfigFileError(f"electron count {sum(row)} != {nelec}", n)
        occ.append(row)
        js.append(sym)
    dropped = False
    if sorted:
        keyed = {(sym, bytes(row)): None for sym, row in zip(js, occ)}
        dropped = len(keyed) < len(occ)
        order = builtins.sorted(keyed)
        js = [k[0] for k in order]
        occ = [list(k[1]) for k in order]
    if len(occ) != ncfg and not dropped:
        raise ConfigFileError(f"found {len(occ)} configurations, header says {ncfg}", len(lines))
    nopen_found = max((row.count(1) for row in occ), default=0)
    return Configs(nbf=nbf, nelec=nelec, ksym=ksym, occ=np.array(occ, dtype=np.uint8).reshape(-1, nbf),
                   jsym=np.array(js, dtype=np.int32), title=title, nopen_max=nopen_found)


# The five benchmark workloads.  Seeds are fixed per workload.
WORKLOADS = {
    "fci6": dict(kind="fci", nbf=6, nelec=6, roots=4, seed=20261203),
    "fci10": dict(kind="fci", nbf=10, nelec=8, roots=4, seed=20261204),
    "mrci8x24": dict(kind="mrci", nint=8, next=24, nelec=6, roots=8, seed=20261205),
    "mrci10x24": dict(kind="mrci", nint=10, next=24, nelec=8, roots=16, seed=20261207),
}


def workload(name: str):
    w = WORKLOADS[name]
    if w["kind"] == "fci":
        cfg = fci_configs(w["nbf"], w["nelec"])
        ints = make_integrals(w["nbf"], w["seed"])
    else:
        cfg = mrcisd_configs(w["nint"], w["next"], w["nelec"])
        ints = make_integrals(w["nint"] + w["next"], w["seed"], n_internal=w["nint"])
    return cfg, ints, w["roots"]


def embed(cfg: Configs, ints: Integrals, orbs, nbf: int = 64, seed: int = 7):
    """The same CI problem on more orbitals: orbital p of (cfg, ints) becomes
    orbital orbs[p] (orbs ascending, so determinant order and phases are
    unchanged) of an nbf-orbital basis whose other integrals are random and
    never reached (those orbitals stay empty)."""
    orbs = [int(o) for o in orbs]
    assert len(orbs) == cfg.nbf and orbs == sorted(orbs) and orbs[-1] < nbf
    rng = np.random.Generator(np.random.PCG64(seed))
    occ = np.zeros((cfg.ncfg, nbf), np.uint8)
    occ[:, orbs] = cfg.occ
    n1 = nndaf(nbf)
    one = rng.normal(0.0, 0.3, n1)
    two = rng.normal(0.0, 0.1, nndaf(n1))
    m = cfg.nbf
    for p in range(m):
        for q in range(p + 1):
            one[pair_index(orbs[p], orbs[q])] = ints.valone[pair_index(p, q)]
    for p, q, r, s in itertools.product(range(m), repeat=4):
        two[canonical_index(orbs[p], orbs[q], orbs[r], orbs[s])] = ints.valtwo[canonical_index(p, q, r, s)]
    return (Configs(nbf=nbf, nelec=cfg.nelec, ksym=cfg.ksym, occ=occ, jsym=cfg.jsym.copy(),
                    title=cfg.title + f" on orbitals {orbs}"),
            Integrals(nbf=nbf, valone=one, valtwo=two, ecore=ints.ecore))


def count_space(cfg: Configs):
    """(ncfg, sum ndet, sum nsef, sum over configs of 2^nopen)."""
    s = cfg.nopen()
    return cfg.ncfg, int(cfg.ndet().sum()), int(cfg.nsef().sum()), int((1 << s).sum())


__all__ = ["Integrals", "Configs", "make_integrals", "fci_configs", "mrcisd_configs", "write_configs",
           "read_configs", "ConfigFileError", "workload", "WORKLOADS", "count_space", "nndaf", "pair_index",
           "canonical_index", "comb", "embed"]

def write_sifs(path: str, ints: Integrals, title: str = "synthetic integrals") -> None:
    """Write integrals as a SIFS-style file through the library: an overlap
    record (identity; the reader skips it), the one-electron integrals as the
    kinetic block with an all-zero potential block (the reader sums them back
    to valone exactly), no spin-orbit records (with one configuration symmetry
    the reference model never uses them), and the two-electron integrals."""
    from .api import sifs_write
    n1 = nndaf(ints.nbf)
    s1 = np.zeros(n1)
    for p in range(ints.nbf):
        s1[pair_index(p, p)] = 1.0
    sifs_write(path, ints.nbf, ints.valtwo, t1=ints.valone, v1=np.zeros(n1), s1=s1, ecore=ints.ecore, title=title)


def read_sifs(path: str) -> Integrals:
    """Integrals from a SIFS-style file, processed as the reference does."""
    from .api import sifs_read
    nbf, one, two, ecore = sifs_read(path)
    return Integrals(nbf=nbf, valone=one, valtwo=two, ecore=ecore)
