"""Python mirror of the C ABI (include/psoci.h) over ctypes.

The class names follow the reference's objects for this path:
``Hamiltonian`` stands for PsociGAhamiltonian (build, stored/direct sigma,
diagonal), ``davidson`` for the driver loop over PsociGAbasis/PsociGAsubspace
(driver.PSOCI.splittest.C:895-1380), ``rdm`` for the CI density
(PsociNOpopulations.C:858-989) and its two-electron extension.  Every call
goes to libpsoci_amd.so on the GPU; there is no CPU fallback: without the
library or a HIP device the calls raise.
"""
from __future__ import annotations

import ctypes
import os
import re
from dataclasses import dataclass

import numpy as np

from .inputs import ConfigFileError, Configs, Integrals

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpsoci_amd.so")
HEADER = os.path.join(HERE, "..", "include", "psoci.h")

OK, ERR_ARG, ERR_UNSUPPORTED, ERR_STATE, ERR_HIP, ERR_NOMEM, ERR_CONVERGE, ERR_NUMERIC = 0, -1, -2, -3, -4, -5, -6, -7
FP64, FP32, MIXED = 0, 1, 2
STORED, DIRECT = 0, 1
CONFIGS_FILE_ORDER, CONFIGS_SORTED = 0, 1

_lib = None


class PsociError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"psoci error {code}: {msg}")
        self.code = code


class DavParams(ctypes.Structure):
    _fields_ = [("nroots", ctypes.c_int), ("prec", ctypes.c_int), ("sigma_mode", ctypes.c_int),
                ("max_iter", ctypes.c_int), ("max_subspace", ctypes.c_int), ("guess", ctypes.c_int),
                ("tol", ctypes.c_double), ("switch_tol", ctypes.c_double), ("extra_roots", ctypes.c_int),
                ("precond", ctypes.c_int), ("graph", ctypes.c_int), ("select_rule", ctypes.c_int)]


class DavStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("sigma_calls", ctypes.c_int), ("sigma_vectors", ctypes.c_int),
                ("switch_iter", ctypes.c_int), ("converged", ctypes.c_int), ("seconds", ctypes.c_double),
                ("sigma_seconds", ctypes.c_double), ("max_rnorm", ctypes.c_double), ("eig_seconds", ctypes.c_double),
                ("graph_launches", ctypes.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def header_functions() -> list[str]:
    """Names of the entry points declared in include/psoci.h."""
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(psoci_\w+)\s*\(", txt, re.M)))


def load(path: str | None = None):
    """Load libpsoci_amd.so (raises if it is missing: no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(f"{p} is not built: run python -m psoci_amd.build")
    # One HIP runtime per process: torch ships its own libamdhip64.so.7.  When
    # torch is importable it is loaded first, so the library binds to the
    # same runtime (same soname) instead of bringing in a second one.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(p)
    V, I, I64, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
    sig = {
        "psoci_ctx_create": (I, [I, ctypes.POINTER(V)]),
        "psoci_ctx_destroy": (I, [V]),
        "psoci_last_error": (ctypes.c_char_p, [V]),
        "psoci_set_integrals": (I, [V, I, V, V, D]),
        "psoci_set_integrals_sifs": (I, [V, ctypes.c_char_p]),
        "psoci_sifs_read": (I, [ctypes.c_char_p, V, V, V, V]),
        "psoci_sifs_write": (I, [ctypes.c_char_p, I, V, V, V, V, V, D, ctypes.c_char_p]),
        "psoci_set_configs": (I, [V, I, I, I, I64, V, V]),
        "psoci_configs_read": (I, [ctypes.c_char_p, I, V, V, V, V, V, V, V, ctypes.c_char_p, I, V]),
        "psoci_set_configs_file": (I, [V, ctypes.c_char_p, I]),
        "psoci_space_info": (I, [V, V, V, V, V]),
        "psoci_dets": (I, [V, I64, V, V, V, V]),
        "psoci_block": (I, [V, I64, I64, V, V]),
        "psoci_build_h": (I, [V, V]),
        "psoci_get_h": (I, [V, V, V, V, V]),
        "psoci_sigma": (I, [V, I, I, I, V, V, I]),
        "psoci_diag": (I, [V, V]),
        "psoci_davidson": (I, [V, ctypes.POINTER(DavParams), V, V, ctypes.POINTER(DavStats)]),
        "psoci_rdm": (I, [V, I, V, V, V, V, V]),
        "psoci_device_ptr": (I, [V, I, ctypes.POINTER(V)]),
        "psoci_project": (I, [V, I64, I, I, V, I, V, I, V]),
        "psoci_combine": (I, [V, I64, I, I, V, I, V, D, V, I]),
        "psoci_build_transition": (I, [V, I64, V, V, V]),
        "psoci_get_transition": (I, [V, V, V, V]),
        "psoci_direct_info": (I, [V, V, V]),
        "psoci_direct_det_pairs": (I, [V, V]),
        "psoci_direct_engine": (I, [V, I, V]),
        "psoci_direct_ras_info": (I, [V, V, V, V, V, V, V, V]),
        "psoci_direct_ras_timing": (I, [V, I, I, I, V, V]),
        "psoci_build_h_timing": (I, [V, V]),
        "psoci_direct_schedule": (I, [V, V, V, V]),
        "psoci_direct_schedule_prec": (I, [V, I, V, V, V]),
        "psoci_stored_shape": (I, [V, I, V, V, V]),
        "psoci_stored_buckets": (I, [V, I, V, V, V, V, V]),
        "psoci_sync": (I, [V]),
        "psoci_h_save": (I, [V, ctypes.c_char_p]),
        "psoci_h_load": (I, [V, ctypes.c_char_p, V]),
        "psoci_civec_write": (I, [ctypes.c_char_p, I64, I, V, ctypes.c_char_p]),
        "psoci_civec_info": (I, [ctypes.c_char_p, V, V, ctypes.c_char_p, I64]),
        "psoci_civec_read": (I, [ctypes.c_char_p, I64, I, V, V]),
        "psoci_sifs_read_ao": (I, [ctypes.c_char_p, V, V, V]),
        "psoci_sifs_write_ao": (I, [ctypes.c_char_p, I, V, ctypes.c_char_p, ctypes.c_char_p]),
        "psoci_energy_contrib": (I, [V, I, V, V, V]),
        "psoci_natural_orbitals": (I, [I, V, V, V, V, V, V]),
        "psoci_mulliken": (I, [I, V, V, V, V]),
        "psoci_basis_groups": (I, [I, ctypes.c_char_p, I, V, V]),
        "psoci_group_populations": (I, [I, V, V, I, V, V]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


@dataclass
class SpaceInfo:
    nsef_total: int
    ndet_total: int
    nsef: np.ndarray
    sef_offset: np.ndarray


class Hamiltonian:
    """One GPU context holding a determinant/SEF space, its integrals and
    (optionally) the stored Hamiltonian."""

    def __init__(self, cfg, ints, device: int = 0):
        self._open(device)
        self.cfg = cfg
        self.ints = ints
        self._check(self.L.psoci_set_configs(self._h, cfg.nbf, cfg.nelec, cfg.ksym, cfg.ncfg, _p(cfg.occ),
                                             _p(cfg.jsym)))
        self.set_integrals(ints)
        self._space_info()

    @classmethod
    def from_files(cls, configs_path: str, sifs_path: str, sorted: bool = False, device: int = 0):
        """The reference's input path: a configuration file
        (psoci_set_configs_file) and a SIFS-style integral file
        (psoci_set_integrals_sifs), both read by the library."""
        self = cls.__new__(cls)
        self._open(device)
        self._check(self.L.psoci_set_configs_file(self._h, os.fsencode(configs_path),
                                                  CONFIGS_SORTED if sorted else CONFIGS_FILE_ORDER))
        self.cfg = read_configs_file(configs_path, sorted=sorted)
        self.set_integrals_sifs(sifs_path)
        self._space_info()
        return self

    def _open(self, device: int):
        self.L = load()
        h = ctypes.c_void_p()
        rc = self.L.psoci_ctx_create(device, ctypes.byref(h))
        if rc != OK:
            raise PsociError(rc, "psoci_ctx_create failed (no HIP device?)")
        self._h = h

    def _space_info(self):
        n, nd = ctypes.c_int64(), ctypes.c_int64()
        nsef = np.zeros(self.cfg.ncfg, np.int32)
        off = np.zeros(self.cfg.ncfg, np.int64)
        self._check(self.L.psoci_space_info(self._h, ctypes.byref(n), ctypes.byref(nd), _p(nsef), _p(off)))
        self.info = SpaceInfo(n.value, nd.value, nsef, off)
        self.n = n.value
        self.nnz_lower = None

    def _check(self, rc):
        if rc != OK:
            raise PsociError(rc, self.L.psoci_last_error(self._h).decode())
        return rc

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self.L.psoci_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_integrals(self, ints):
        one = np.ascontiguousarray(ints.valone, np.float64)
        two = np.ascontiguousarray(ints.valtwo, np.float64)
        self._check(self.L.psoci_set_integrals(self._h, ints.nbf, _p(one), _p(two), float(ints.ecore)))
        self.ints = ints
        self.nnz_lower = None

    def set_integrals_sifs(self, path: str):
        """Integrals from a SIFS-style file (psoci_set_integrals_sifs)."""
        self._check(self.L.psoci_set_integrals_sifs(self._h, os.fsencode(path)))
        nbf, one, two, ecore = sifs_read(path)
        self.ints = Integrals(nbf=nbf, valone=one, valtwo=two, ecore=ecore)
        self.nnz_lower = None

    def set_configs_file(self, path: str, sorted: bool = False):
        """Replace the configuration list by a file's (psoci_set_configs_file;
        sorted=True: the sorted, de-duplicated reader)."""
        self._check(self.L.psoci_set_configs_file(self._h, os.fsencode(path),
                                                  CONFIGS_SORTED if sorted else CONFIGS_FILE_ORDER))
        self.cfg = read_configs_file(path, sorted=sorted)
        self._space_info()

    def dets(self, I: int):
        nd = 1 << max(int((self.cfg.occ[I] == 1).sum()) - 1, 0)
        a = np.zeros(nd, np.uint64)
        b = np.zeros(nd, np.uint64)
        ms = np.zeros(nd, np.int32)
        ph = np.zeros(nd, np.int32)
        self._check(self.L.psoci_dets(self._h, I, _p(a), _p(b), _p(ms), _p(ph)))
        return dict(alpha=a, beta=b, ms=ms, phase=ph)

    def block(self, I: int, J: int):
        ncfg = self.cfg.ncfg
        if not (0 <= I < ncfg and 0 <= J < ncfg):
            # let the library reject the indices (PSOCI_ERR_ARG)
            dummy = np.zeros(1)
            self._check(self.L.psoci_block(self._h, I, J, _p(dummy), None))
        out = np.zeros((int(self.info.nsef[I]), int(self.info.nsef[J])), np.float64)
        st = ctypes.c_int()
        self._check(self.L.psoci_block(self._h, I, J, _p(out), ctypes.byref(st)))
        return st.value, out

    def build(self) -> int:
        nnz = ctypes.c_int64()
        self._check(self.L.psoci_build_h(self._h, ctypes.byref(nnz)))
        self.nnz_lower = nnz.value
        return nnz.value

    def build_timing(self) -> dict:
        """Seconds of the last build(): the build proper and the stored
        sigma's shape autotuner after it."""
        sec = np.zeros(2)
        self._check(self.L.psoci_build_h_timing(self._h, _p(sec)))
        return {"build_s": float(sec[0]), "tune_s": float(sec[1])}

    def transition(self, cfg2):
        """Transition Hamiltonian with this list as rows and cfg2's as
        columns (psoci_build_transition + psoci_get_transition): rowptr,
        col, val over this list's SEFs."""
        occ = np.ascontiguousarray(cfg2.occ, np.uint8)
        js = np.ascontiguousarray(cfg2.jsym, np.int32)
        nnz = ctypes.c_int64()
        self._check(self.L.psoci_build_transition(self._h, cfg2.ncfg, _p(occ), _p(js), ctypes.byref(nnz)))
        rowptr = np.zeros(self.n + 1, np.int64)
        col = np.zeros(nnz.value, np.int32)
        val = np.zeros(nnz.value, np.float64)
        self._check(self.L.psoci_get_transition(self._h, _p(rowptr), _p(col), _p(val)))
        return rowptr, col, val

    def save_h(self, path: str):
        """Checkpoint the stored Hamiltonian (the reference's DRA dump)."""
        self._check(self.L.psoci_h_save(self._h, os.fsencode(path)))

    def load_h(self, path: str) -> int:
        """Load a checkpoint written for the same configurations and integrals
        in place of build() (the reference's --hamiltonian_restart)."""
        nnz = ctypes.c_int64()
        self._check(self.L.psoci_h_load(self._h, os.fsencode(path), ctypes.byref(nnz)))
        self.nnz_lower = nnz.value
        return nnz.value

    def get_h(self):
        n, nnz = self.n, self.nnz_lower
        rowptr = np.zeros(n + 1, np.int64)
        col = np.zeros(nnz, np.int32)
        val = np.zeros(nnz, np.float64)
        diag = np.zeros(n, np.float64)
        self._check(self.L.psoci_get_h(self._h, _p(rowptr), _p(col), _p(val), _p(diag)))
        return rowptr, col, val, diag

    def diag(self):
        d = np.zeros(self.n, np.float64)
        self._check(self.L.psoci_diag(self._h, _p(d)))
        return d

    def sigma(self, X: np.ndarray, mode: int = STORED, prec: int = FP64):
        dt = np.float32 if prec == FP32 else np.float64
        X = np.ascontiguousarray(X, dt)
        squeeze = X.ndim == 1
        if squeeze:
            X = X[:, None]
        Y = np.zeros_like(X)
        self._check(self.L.psoci_sigma(self._h, mode, prec, X.shape[1], _p(X), _p(Y), 0))
        return Y[:, 0] if squeeze else Y

    def sigma_device(self, X, Y, mode: int = STORED, prec: int = FP64):
        """X, Y: torch tensors on the GPU, [n][k] contiguous.  The library's
        stream is not ordered after torch's, so torch's pending work (what
        wrote X) is finished first; call sync() before reading Y."""
        import torch
        torch.cuda.current_stream().synchronize()
        k = 1 if X.dim() == 1 else X.shape[1]
        self._check(self.L.psoci_sigma(self._h, mode, prec, k, ctypes.c_void_p(X.data_ptr()),
                                       ctypes.c_void_p(Y.data_ptr()), 1))
        return Y

    def davidson(self, nroots: int, prec: int = FP64, mode: int = STORED, tol: float = 1e-6,
                 switch_tol: float = 1e-4, max_iter: int = 200, max_subspace: int = 0, guess: int = 0,
                 want_vectors: bool = True, guess_vectors=None, extra_roots: int = 0,
                 restart: str | None = None, precond: int = 0, graph: int = 0, select_rule: int = 0):
        """Block Davidson for the nroots lowest roots.  guess: 0 lowest diagonal,
        1 first SEFs (the reference's), 2 guess_vectors ([n][nroots]);
        restart: a CI-vector file (civec_write) read as guess_vectors, missing
        roots completed with unit vectors (the reference's --vector_restart);
        extra_roots guard roots ride along in the block (not tested; with
        MIXED only in the FP32 phase); precond 0: the reference's diagonal
        preconditioner, 1: the same with Olsen's correction; graph 0:
        stored-sigma steps as captured hipGraphs, 1: eager launches;
        select_rule 0: the padded kernel width's free columns take more
        corrections, 1: the reference's one per unconverged root."""
        if restart is not None:
            guess_vectors, _ = civec_read(restart, self.n, nroots)
            guess = 2
        p = DavParams(nroots, prec, mode, max_iter, max_subspace, guess, tol, switch_tol, extra_roots, precond, graph,
                      select_rule)
        ev = np.zeros(nroots)
        vec = np.zeros((self.n, nroots)) if want_vectors else None
        if guess == 2:
            if guess_vectors is None:
                raise ValueError("guess=2 needs guess_vectors [n][nroots]")
            vec = np.ascontiguousarray(guess_vectors, dtype=np.float64).reshape(self.n, nroots).copy()
        st = DavStats()
        rc = self.L.psoci_davidson(self._h, ctypes.byref(p), _p(ev), _p(vec), ctypes.byref(st))
        if rc not in (OK, ERR_CONVERGE):
            self._check(rc)
        return dict(evals=ev, evecs=vec, stats=st.as_dict(), converged=rc == OK)

    def energy_contrib(self, C: np.ndarray, evals) -> np.ndarray:
        """deltae [ncfg][k]: each configuration's energy contribution to each
        root (analyzeEnergyContributions, PsociCIAnalysis.C:126-142)."""
        C = np.ascontiguousarray(C, np.float64)
        if C.ndim == 1:
            C = C[:, None]
        k = C.shape[1]
        ev = np.ascontiguousarray(evals, np.float64).reshape(k)
        out = np.zeros((self.cfg.ncfg, k))
        self._check(self.L.psoci_energy_contrib(self._h, k, _p(C), _p(ev), _p(out)))
        return out

    def project(self, A, B):
        """A^T B over the rows of two GPU tensors (FP64 MFMA); torch's
        pending work on them is finished first (the library's stream is not
        ordered after torch's)."""
        import torch
        torch.cuda.current_stream().synchronize()
        n, m = A.shape
        b = B.shape[1]
        out = np.zeros((m, b))
        self._check(self.L.psoci_project(self._h, n, m, b, ctypes.c_void_p(A.data_ptr()), A.stride(0),
                                         ctypes.c_void_p(B.data_ptr()), B.stride(0), _p(out)))
        return out

    def combine(self, A, Y, Out, alpha: float = 0.0):
        """Out = alpha Out + A Y over the rows of GPU tensors (FP64 MFMA);
        Y host [m][k]; torch's pending work on A and Out is finished first."""
        import torch
        torch.cuda.current_stream().synchronize()
        n, m = A.shape
        k = Out.shape[1]
        Yh = np.ascontiguousarray(Y, np.float64).reshape(m, k)
        self._check(self.L.psoci_combine(self._h, n, m, k, ctypes.c_void_p(A.data_ptr()), A.stride(0), _p(Yh),
                                         float(alpha), ctypes.c_void_p(Out.data_ptr()), Out.stride(0)))
        return Out

    def direct_info(self):
        """(connected pairs, structural SEF elements) per direct sigma."""
        a, b = ctypes.c_int64(), ctypes.c_int64()
        self._check(self.L.psoci_direct_info(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def direct_engine(self, request: int = -1, build: bool = True):
        """Direct sigma engine: request -1 keeps the choice, 0 auto (the
        string-driven engine where the space qualifies), 1 the
        configuration-driven records, 2 string-driven only.  Returns the
        engine in use (1 or 2; the operator is built) or None with
        build=False."""
        e = ctypes.c_int()
        self._check(self.L.psoci_direct_engine(self._h, request, ctypes.byref(e) if build else None))
        return e.value if build else None

    def direct_ras_info(self):
        """The string-driven engine's plan: dict(m, n_external, sectors,
        blocks, ndet, w_per_column, flops_per_column)."""
        m, ne, ns, nb = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        nd, wu, fl = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_double()
        self._check(self.L.psoci_direct_ras_info(self._h, ctypes.byref(m), ctypes.byref(ne), ctypes.byref(ns),
                                                 ctypes.byref(nb), ctypes.byref(nd), ctypes.byref(wu),
                                                 ctypes.byref(fl)))
        return {"m": m.value, "n_external": ne.value, "sectors": ns.value, "blocks": nb.value, "ndet": nd.value,
                "w_per_column": wu.value, "flops_per_column": fl.value}

    def direct_ras_timing(self, prec: int, k: int, reps: int = 3):
        """Average ms per string-driven sigma call by launch group (scatter,
        alpha-beta, alpha-alpha, beta-beta products, gather) and the
        products' flops per call by kind."""
        ms = np.zeros(5)
        fl = np.zeros(3)
        self._check(self.L.psoci_direct_ras_timing(self._h, prec, k, reps, _p(ms), _p(fl)))
        return {"scatter_ms": ms[0], "gemm_ms": list(ms[1:4]), "gather_ms": ms[4], "gemm_flops": list(fl)}

    def direct_det_pairs(self) -> int:
        """Connected determinant pairs per direct sigma (counted once)."""
        v = ctypes.c_int64()
        self._check(self.L.psoci_direct_det_pairs(self._h, ctypes.byref(v)))
        return v.value

    def direct_schedule(self, prec: int = FP32):
        """The direct sigma's autotuned schedule of a precision: dict(order,
        chunk, tune_ms)."""
        o, ch, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        self._check(self.L.psoci_direct_schedule_prec(self._h, prec, ctypes.byref(o), ctypes.byref(ch),
                                                      ctypes.byref(ms)))
        return {"order": o.value, "chunk": ch.value, "tune_ms": ms.value}

    def stored_shape(self, prec: int = FP32):
        """The stored sigma's autotuned shape: dict(lanes_per_nz, rows_per_wave, tune_ms)."""
        g, r, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        self._check(self.L.psoci_stored_shape(self._h, prec, ctypes.byref(g), ctypes.byref(r), ctypes.byref(ms)))
        return {"lanes_per_nz": g.value, "rows_per_wave": r.value, "tune_ms": ms.value}

    def stored_buckets(self, prec: int = FP32):
        """The stored sigma's row-length buckets: dict(used, rows, lanes_per_nz,
        rows_per_wave, bucketed_ms)."""
        used, ms = ctypes.c_int(), ctypes.c_double()
        rows = np.zeros(5, np.int64)
        g = np.zeros(5, np.int32)
        r = np.zeros(5, np.int32)
        self._check(self.L.psoci_stored_buckets(self._h, prec, ctypes.byref(used), _p(rows), _p(g), _p(r),
                                                ctypes.byref(ms)))
        return {"used": bool(used.value), "rows": rows.tolist(), "lanes_per_nz": g.tolist(),
                "rows_per_wave": r.tolist(), "bucketed_ms": ms.value}

    def sync(self):
        """Wait for the work queued on the library's stream (psoci_sync)."""
        self._check(self.L.psoci_sync(self._h))

    def stream_handle(self) -> int:
        """The hipStream_t the library launches on (as an integer)."""
        ptr = ctypes.c_void_p()
        self._check(self.L.psoci_device_ptr(self._h, 1, ctypes.byref(ptr)))
        return ptr.value or 0

    def rdm(self, C: np.ndarray, weights=None):
        """Gradients of sum_k w_k E_k with respect to the packed integrals."""
        C = np.ascontiguousarray(C, np.float64)
        if C.ndim == 1:
            C = C[:, None]
        k = C.shape[1]
        w = None if weights is None else np.ascontiguousarray(weights, np.float64)
        nbf = self.cfg.nbf
        n1 = nbf * (nbf + 1) // 2
        g1 = np.zeros(n1)
        g2 = np.zeros(n1 * (n1 + 1) // 2)
        gso = np.zeros(n1)
        self._check(self.L.psoci_rdm(self._h, k, _p(C), _p(w), _p(g1), _p(g2), _p(gso)))
        return g1, g2, gso


def _ctxfree(rc: int):
    if rc != OK:
        msg = load().psoci_last_error(None)
        raise PsociError(rc, msg.decode() if msg else "")


def read_configs_file(path: str, sorted: bool = False) -> Configs:
    """A configuration file through the library's reader (psoci_configs_read):
    file order, or sorted by (jsym, occupation) with duplicates dropped.  A
    malformed line raises ConfigFileError carrying its line number."""
    L = load()
    order = CONFIGS_SORTED if sorted else CONFIGS_FILE_ORDER
    nbf, nelec, ksym, nopen = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    ncfg, line = ctypes.c_int64(), ctypes.c_int64()

    def call(occ, jsym, title, cap):
        rc = L.psoci_configs_read(os.fsencode(path), order, ctypes.byref(nbf), ctypes.byref(nelec),
                                  ctypes.byref(ksym), ctypes.byref(ncfg), ctypes.byref(nopen), occ, jsym, title,
                                  cap, ctypes.byref(line))
        if rc != OK:
            msg = (L.psoci_last_error(None) or b"").decode()
            if line.value > 0:
                raise ConfigFileError(msg.split(": ", 1)[-1], line.value)
            raise PsociError(rc, msg)

    call(None, None, None, 0)
    occ = np.zeros((ncfg.value, nbf.value), np.uint8)
    jsym = np.zeros(ncfg.value, np.int32)
    title = ctypes.create_string_buffer(4096)
    call(_p(occ), _p(jsym), title, len(title))
    return Configs(nbf=nbf.value, nelec=nelec.value, ksym=ksym.value, occ=occ, jsym=jsym,
                   title=title.value.decode(errors="replace"), nopen_max=nopen.value)


def sifs_read(path: str):
    """(nbf, valone, valtwo, ecore) from a SIFS-style integral file, processed
    as the reference does (T + V summed, overlap skipped, spin-orbit records
    overwrite antisymmetrically, core energy = energy[0])."""
    L = load()
    nbf = ctypes.c_int()
    _ctxfree(L.psoci_sifs_read(os.fsencode(path), ctypes.byref(nbf), None, None, None))
    n = nbf.value
    one = np.zeros(n * (n + 1) // 2)
    two = np.zeros(len(one) * (len(one) + 1) // 2)
    ecore = ctypes.c_double()
    _ctxfree(L.psoci_sifs_read(os.fsencode(path), ctypes.byref(nbf), _p(one), _p(two), ctypes.byref(ecore)))
    return n, one, two, ecore.value


def sifs_write(path: str, nbf: int, valtwo, t1=None, v1=None, s1=None, so=None, ecore: float = 0.0,
               title: str | None = None):
    """Write a SIFS-style integral file: overlap s1, kinetic t1 and potential
    v1 one-electron blocks (packed like valone; the reader sums t1 + v1),
    optional spin-orbit so [3][NNDAF(nbf)], two-electron valtwo."""
    L = load()
    arr = lambda a: None if a is None else np.ascontiguousarray(a, np.float64)  # noqa: E731
    t1, v1, s1, so, two = arr(t1), arr(v1), arr(s1), arr(so), arr(valtwo)
    ptr = lambda a: None if a is None else _p(a)  # noqa: E731
    _ctxfree(L.psoci_sifs_write(os.fsencode(path), nbf, ptr(s1), ptr(t1), ptr(v1), ptr(so), _p(two), float(ecore),
                                None if title is None else title.encode()))


def civec_write(path: str, C, title: str | None = None):
    """Write CI vectors C [n][k] in the reference's binary restart format."""
    C = np.ascontiguousarray(C, np.float64)
    if C.ndim == 1:
        C = C[:, None]
    _ctxfree(load().psoci_civec_write(os.fsencode(path), C.shape[0], C.shape[1], _p(C),
                                      None if title is None else title.encode()))


def civec_info(path: str):
    """(nbasis, nroots, title) of a CI-vector restart file."""
    nb, nr = ctypes.c_int64(), ctypes.c_int64()
    buf = ctypes.create_string_buffer(4096)
    _ctxfree(load().psoci_civec_info(os.fsencode(path), ctypes.byref(nb), ctypes.byref(nr), buf, len(buf)))
    return nb.value, nr.value, buf.value.decode(errors="replace")


def civec_read(path: str, n: int, k: int):
    """(C [n][k], roots read): the first min(nroots, k) roots of the file and
    unit vectors e_j in the remaining columns j, as the reference's
    readVectorScalable + specifyDiagonalGuessVectors do."""
    C = np.zeros((n, k))
    got = ctypes.c_int()
    _ctxfree(load().psoci_civec_read(os.fsencode(path), n, k, _p(C), ctypes.byref(got)))
    return C, got.value


def restart_filename(name: str, rank: int = 0) -> str:
    """The name PsociVector gives a file for rank `rank` (PsociVector.C:185-208):
    "-<rank>" inserted before the last "." of the string, else appended."""
    i = name.rfind(".")
    tag = f"-{rank}"
    return name[:i] + tag + name[i:] if i >= 0 else name + tag


AU_TO_EV = 27.21138505  # PsociCIAnalysis.hpp:42


def analyze_energy_contributions(H: "Hamiltonian", C, evals, thresh: float = 1e-4, ecore: float = 0.0):
    """The reference's CI energy analysis (analyzeEnergyContributions,
    PsociCIAnalysis.C:42-267) for the roots C [n][k]: per root the energies
    (valence, core, total, E(i) - E(1) in eV) and the configurations with
    |deltae| >= thresh, largest first (stable for ties, as the reference's
    list sort), each with its SEFs' coefficients and per-SEF contributions."""
    C = np.ascontiguousarray(C, np.float64)
    if C.ndim == 1:
        C = C[:, None]
    evals = np.asarray(evals, np.float64)
    de = H.energy_contrib(C, evals)
    diag = H.diag()
    off = H.info.sef_offset
    nsef = H.info.nsef
    roots = []
    for r in range(C.shape[1]):
        order = np.argsort(-np.abs(de[:, r]), kind="stable")
        keep = order[np.abs(de[order, r]) >= thresh]
        confs = []
        for I in keep:
            I = int(I)
            i0, i1 = int(off[I]), int(off[I]) + int(nsef[I])
            v = C[i0:i1, r]
            v2 = v * v
            with np.errstate(divide="ignore", invalid="ignore"):
                dl = np.where(v2 == 1.0, evals[r], v2 * (evals[r] - diag[i0:i1]) / (1.0 - v2))
            confs.append(dict(config=I, deltae=float(de[I, r]), occ="".join(str(int(x)) for x in H.cfg.occ[I]),
                              sefs=list(range(i0, i1)), coef=v.copy(), del_sef=dl))
        roots.append(dict(valence=float(evals[r]), core=float(ecore), total=float(evals[r] + ecore),
                          excitation_ev=float((evals[r] - evals[0]) * AU_TO_EV), configs=confs,
                          distribution=energy_distribution(de[:, r])))
    return roots


def energy_distribution(deltae, bins=(1e-8, 1e-7, 1e-6, 1e-5, 1e-4, 1e-3, 1e-2, 1e-1, 1.0, 10.0)):
    """Count and sum of the configuration contributions per |deltae| decade
    (the bins of energyDistributionSpatial, PsociCIAnalysis.C:267-335):
    entry b covers bins[b-1] <= |deltae| < bins[b] (b = 0: below bins[0],
    the last entry: at or above bins[-1])."""
    a = np.abs(np.asarray(deltae))
    edges = np.concatenate([[0.0], bins, [np.inf]])
    out = []
    for lo, hi in zip(edges[:-1], edges[1:]):
        m = (a >= lo) & (a < hi)
        out.append(dict(lo=float(lo), hi=float(hi), count=int(m.sum()), sum=float(np.asarray(deltae)[m].sum())))
    return out


def rdm1_from_grad(grad_one):
    """The reference's packed CI density (off-diagonal once per orbital pair)
    from psoci_rdm's dE/d(valone): the off-diagonal halved."""
    g = np.asarray(grad_one, np.float64).copy()
    n1 = len(g)
    nbf = int((np.sqrt(8 * n1 + 1) - 1) / 2)
    for p in range(nbf):
        for q in range(p):
            g[p * (p + 1) // 2 + q] *= 0.5
    return g


def natural_orbitals(rdm1, mocoef=None):
    """Natural orbitals of a packed CI density: dict(occ, no_mo[, no_ao,
    dens_ao]) (constructNaturalOrbitals, PsociNOpopulations.C:569-729);
    mocoef [MO][AO] adds the AO-basis NOs and the packed AO density."""
    rdm1 = np.ascontiguousarray(rdm1, np.float64)
    nbf = int((np.sqrt(8 * len(rdm1) + 1) - 1) / 2)
    occ = np.zeros(nbf)
    no_mo = np.zeros((nbf, nbf))
    out = dict(occ=occ, no_mo=no_mo)
    mo = no_ao = dens = None
    if mocoef is not None:
        mo = np.ascontiguousarray(mocoef, np.float64).reshape(nbf, nbf)
        no_ao = np.zeros((nbf, nbf))
        dens = np.zeros(len(rdm1))
        out.update(no_ao=no_ao, dens_ao=dens)
    _ctxfree(load().psoci_natural_orbitals(nbf, _p(rdm1), _p(mo), _p(occ), _p(no_mo), _p(no_ao), _p(dens)))
    return out


def mulliken(dens_ao, overlap):
    """(population [nbf][nbf], total) of constructMullikenPopulationDensAO."""
    d = np.ascontiguousarray(dens_ao, np.float64)
    s = np.ascontiguousarray(overlap, np.float64)
    nbf = int((np.sqrt(8 * len(d) + 1) - 1) / 2)
    pop = np.zeros((nbf, nbf))
    tot = ctypes.c_double()
    _ctxfree(load().psoci_mulliken(nbf, _p(d), _p(s), _p(pop), ctypes.byref(tot)))
    return pop, tot.value


ATOMIC, SUBSHELL, NONE = 0, 1, 2


def _labels_bytes(labels):
    return b"".join(lab.encode()[:8].ljust(8) for lab in labels)


def basis_groups(labels, mode: int = ATOMIC):
    """(group per basis function, number of groups) of clusterBasisLabels."""
    nbf = len(labels)
    grp = np.zeros(nbf, np.int32)
    ng = ctypes.c_int32()
    _ctxfree(load().psoci_basis_groups(nbf, _labels_bytes(labels), mode, _p(grp), ctypes.byref(ng)))
    return grp, ng.value


def group_populations(population, group, ngroups: int):
    """(gross [ngroups], bond [ngroups][ngroups]) condensed Mulliken populations."""
    pop = np.ascontiguousarray(population, np.float64)
    grp = np.ascontiguousarray(group, np.int32)
    gross = np.zeros(ngroups)
    bond = np.zeros((ngroups, ngroups))
    _ctxfree(load().psoci_group_populations(len(grp), _p(pop), _p(grp), ngroups, _p(gross), _p(bond)))
    return gross, bond


def sifs_write_ao(path: str, overlap, labels=None, title: str | None = None):
    """Write an AO-basis file: packed overlap and 8-character basis labels."""
    s = np.ascontiguousarray(overlap, np.float64)
    nbf = int((np.sqrt(8 * len(s) + 1) - 1) / 2)
    _ctxfree(load().psoci_sifs_write_ao(os.fsencode(path), nbf, _p(s),
                                        None if labels is None else _labels_bytes(labels),
                                        None if title is None else title.encode()))


def sifs_read_ao(path: str):
    """(packed overlap, basis labels) of an AO-basis file."""
    L = load()
    nbf = ctypes.c_int()
    _ctxfree(L.psoci_sifs_read_ao(os.fsencode(path), ctypes.byref(nbf), None, None))
    n = nbf.value
    s = np.zeros(n * (n + 1) // 2)
    lab = ctypes.create_string_buffer(8 * n)
    _ctxfree(L.psoci_sifs_read_ao(os.fsencode(path), ctypes.byref(nbf), _p(s), lab))
    raw = lab.raw
    return s, [raw[8 * i:8 * i + 8].decode() for i in range(n)]
