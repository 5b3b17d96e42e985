"""Oracle for the exact-diagonalisation hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the cpu_baseline leg of bench.py may
import this package, and only as the checker (or, for bench.py, as the timed CPU
baseline).  The product package ``ed_solver_amd`` never imports it.

Two independent parts:

* ``OracleSector`` -- ctypes binding of libed_oracle.so, the plain C + OpenMP
  restatement of the reference algorithms (each function in ed_oracle.c cites the
  reference file:line it follows).
* ``oracle.brute`` -- NumPy/SciPy brute force from second quantisation with a
  global Jordan-Wigner convention, used to pin the C restatement.

Parity status of each oracle function is listed in the header of ed_oracle.c and
pinned by tests/test_oracle.py (closed forms, library eigen-solvers, brute force).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from ed_solver_amd.aim import AIMParams, CModel

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libed_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


def build(force: bool = False) -> str:
    """Compile oracle/libed_oracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH) or \
            os.path.getmtime(_LIB_PATH) < os.path.getmtime(os.path.join(_HERE, "ed_oracle.c")):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_sector_create.restype = ctypes.c_void_p
        L.orc_sector_create.argtypes = [ctypes.POINTER(CModel), ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_sector_destroy.argtypes = [ctypes.c_void_p]
        L.orc_sector_dim.restype = ctypes.c_int64
        L.orc_sector_dim.argtypes = [ctypes.c_void_p]
        L.orc_sector_dup.argtypes = [ctypes.c_void_p]
        L.orc_sector_ddn.argtypes = [ctypes.c_void_p]
        L.orc_sector_nnz.restype = ctypes.c_int64
        L.orc_sector_nnz.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_set_threads.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_sector_tables.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _dp, _ip, _ip, _dp]
        L.orc_hmult.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_lanczos_diagonalize.restype = ctypes.c_int
        L.orc_lanczos_diagonalize.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                              _dp, _dp, _dp, _dp, _dp, _ip]
        L.orc_lanczos_gs.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp]
        L.orc_lanczos_run.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp]
        L.orc_lanczos_step.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, ctypes.c_int, _dp]
        L.orc_tridiag_eigvals.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int, _dp]
        L.orc_tridiag_eigvec.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_double, _dp]
        L.orc_cfrac.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.orc_apply_c.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_dynamic.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, _dp, ctypes.c_int, ctypes.c_double,
                                  ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_num_procs.restype = ctypes.c_int
        L.orc_sector_states.argtypes = [ctypes.c_int, ctypes.c_int, _ip]
        L.orc_rank_table.argtypes = [ctypes.c_int, _ip]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def sector_states(ns: int, n: int) -> np.ndarray:
    out = np.zeros(1 << ns, dtype=np.int32)
    c = lib().orc_sector_states(ns, n, _i(out))
    return out[:c].copy()


def rank_table(ns: int) -> np.ndarray:
    """1-based rank of every state inside its popcount bucket (reference convention)."""
    out = np.zeros(1 << ns, dtype=np.int32)
    lib().orc_rank_table(ns, _i(out))
    return out


def start_vector(dim: int) -> np.ndarray:
    """Deterministic Lanczos start vector shared with the device path.

    The reference fills v1(i) = ran_tab(i mod 10000) from gfortran's RANDOM_NUMBER
    rounded to 3 decimals (rcvector_class.f90:43-60, random.f90:661-706); that
    stream cannot be reproduced without gfortran's runtime, so both paths use the
    same table built from a fixed integer hash, rounded to 3 decimals, period 10000.
    """
    i = np.arange(dim, dtype=np.uint64) % np.uint64(10000)
    x = (i + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
    x ^= x >> np.uint64(29)
    x *= np.uint64(0xBF58476D1CE4E5B9)
    x ^= x >> np.uint64(32)
    m = x >> np.uint64(11)                                   # 53 random bits
    k = (m * np.uint64(1000)) >> np.uint64(53)               # floor(1000 * m / 2^53) in [0, 999]
    return (k + np.uint64(1)).astype(np.float64) / 1000.0    # in (0, 1], never all-zero


class OracleSector:
    """CPU restatement of one (N_up, N_dn) sector: tables, H.v, Lanczos, G."""

    def __init__(self, model: AIMParams, nup: int, ndn: int, nthreads: int = 1):
        self.model = model
        self._cm = CModel.from_params(model)
        self.nup, self.ndn = nup, ndn
        self.h = lib().orc_sector_create(ctypes.byref(self._cm), nup, ndn, nthreads)
        self.dim = lib().orc_sector_dim(self.h)
        self.dup = lib().orc_sector_dup(self.h)
        self.ddn = lib().orc_sector_ddn(self.h)

    def close(self):
        if self.h:
            lib().orc_sector_destroy(self.h)
            self.h = None

    __del__ = close

    def set_threads(self, n: int):
        lib().orc_set_threads(self.h, n)

    def tables(self, spin: int) -> dict:
        d = self.ddn if spin else self.dup
        nnz = lib().orc_sector_nnz(self.h, spin)
        st = np.zeros(d, np.int32); dg = np.zeros(d); no = np.zeros(d, np.int32)
        rk = np.zeros(max(nnz, 1), np.int32); of = np.zeros(max(nnz, 1))
        lib().orc_sector_tables(self.h, spin, _i(st), _d(dg), _i(no), _i(rk), _d(of))
        return {"states": st, "diag": dg, "noff": no, "rankoff": rk[:nnz], "offdiag": of[:nnz]}

    def hmult(self, v: np.ndarray) -> np.ndarray:
        v = np.ascontiguousarray(v, dtype=np.float64)
        assert v.shape == (self.dim,)
        out = np.empty_like(v)
        lib().orc_hmult(self.h, _d(v), _d(out))
        return out

    def lanczos(self, nitermax=128, neigen=1, tol=1e-12, start=None):
        """Lanczos_fast_diagonalize: returns dict(iters, evals, alpha, beta, vecp)."""
        start = start_vector(self.dim) if start is None else np.ascontiguousarray(start, np.float64)
        n = int(min(self.dim - 1, nitermax))
        alpha = np.zeros(max(n, 1)); beta = np.zeros(max(n, 1))
        neig = int(min(neigen, self.dim, n))
        ev = np.zeros(max(neig, 1)); vecp = np.zeros(max(neig, 1) * max(n, 1)); conv = np.zeros(max(neig, 1), np.int32)
        it = lib().orc_lanczos_diagonalize(self.h, nitermax, neigen, tol, _d(start), _d(alpha), _d(beta),
                                           _d(ev), _d(vecp), _i(conv))
        if it <= 0:
            raise RuntimeError(f"oracle Lanczos stopped with status {it}")
        return {"iters": it, "evals": ev[:neig].copy(), "alpha": alpha[:it].copy(), "beta": beta[:it].copy(),
                "vecp": vecp[:neig * it].reshape(neig, it).copy(), "converged": conv[:neig].astype(bool)}

    def ground_state(self, iters: int, vecp_row: np.ndarray, start=None) -> np.ndarray:
        start = start_vector(self.dim) if start is None else np.ascontiguousarray(start, np.float64)
        psi = np.zeros(self.dim)
        vecp_row = np.ascontiguousarray(vecp_row, np.float64)
        lib().orc_lanczos_gs(self.h, iters, _d(vecp_row), _d(start), _d(psi))
        return psi

    def lanczos_run(self, niter: int, start: np.ndarray):
        start = np.ascontiguousarray(start, np.float64)
        a = np.zeros(niter); b = np.zeros(niter)
        st = lib().orc_lanczos_run(self.h, niter, _d(start), _d(a), _d(b))
        if st < 0:
            raise RuntimeError(f"oracle Lanczos stopped with status {st}")
        return a, b

    def dynamic(self, iph: int, energy: float, phi: np.ndarray, nitergreen: int, wn: np.ndarray,
                cutoff_dynamic: float = 1e-8):
        phi = np.ascontiguousarray(phi, np.float64)
        wn = np.ascontiguousarray(wn, np.float64)
        re = np.zeros(len(wn)); im = np.zeros(len(wn))
        n = max(1, min(nitergreen, self.dim))
        a = np.zeros(n); b = np.zeros(n)
        st = lib().orc_dynamic(self.h, iph, energy, _d(phi), nitergreen, cutoff_dynamic, len(wn), _d(wn),
                               _d(re), _d(im), _d(a), _d(b))
        if st < 0:
            raise RuntimeError(f"oracle compute_dynamic stopped with status {st}")
        return re + 1j * im, a[:st], b[:st]


def lowest_eigs(sector: OracleSector, k: int = 8, tol: float = 1e-14):
    """The k lowest eigenpairs of the restated H: LAPACK (numpy eigh) on the dense matrix
    built column by column from the restated H.v when dim <= 3000, else ARPACK (scipy
    eigsh, which='SA') on the H.v as a linear operator.  The library routine is the
    ground truth for Block_Lanczos_diagonalize and ARPACK_diago (ED_ARPACK.f90:475-570)."""
    import scipy.sparse.linalg as sla
    n = sector.dim
    if n <= 3000:
        H = np.stack([sector.hmult(e) for e in np.eye(n)], axis=1)
        w, v = np.linalg.eigh(0.5 * (H + H.T))
        return w[:k], v[:, :k].T.copy()
    op = sla.LinearOperator((n, n), matvec=sector.hmult, dtype=np.float64)
    w, v = sla.eigsh(op, k=k, which="SA", tol=tol, v0=start_vector(n), ncv=max(3 * k, 40))
    order = np.argsort(w)
    return w[order], v[:, order].T.copy()


def apply_c(src: OracleSector, dst: OracleSector, site: int, spin: int, pm: int, psi: np.ndarray) -> np.ndarray:
    psi = np.ascontiguousarray(psi, np.float64)
    phi = np.zeros(dst.dim)
    lib().orc_apply_c(src.h, dst.h, site, spin, pm, _d(psi), _d(phi))
    return phi


def tridiag_eigvals(a, b, k=None):
    a = np.ascontiguousarray(a, np.float64); b = np.ascontiguousarray(b, np.float64)
    n = len(a); k = n if k is None else k
    w = np.zeros(k)
    lib().orc_tridiag_eigvals(n, _d(a), _d(b), k, _d(w))
    return w


def tridiag_eigvec(a, b, lam):
    a = np.ascontiguousarray(a, np.float64); b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(len(a))
    lib().orc_tridiag_eigvec(len(a), _d(a), _d(b), float(lam), _d(x))
    return x


def cfrac(a, b, z: complex) -> complex:
    a = np.ascontiguousarray(a, np.float64); b = np.ascontiguousarray(b, np.float64)
    re = np.zeros(1); im = np.zeros(1)
    st = lib().orc_cfrac(len(a), _d(a), _d(b), z.real, z.imag, _d(re), _d(im))
    if st < 0:
        raise ZeroDivisionError("continued fraction denominator below epsilon")
    return complex(re[0], im[0])


def matsubara(beta: float, nw: int) -> np.ndarray:
    """Imaginary parts of i w_n = i (2n-1) pi / beta, n = 1..nw (frequency_axis.f90:165-170)."""
    n = np.arange(1, nw + 1, dtype=np.float64)
    return (2.0 * np.pi / beta) * n - np.pi / beta
