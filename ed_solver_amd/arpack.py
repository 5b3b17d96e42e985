"""ARPACK_diago (ED_ARPACK.f90:475-570).

The reference calls ARPACK96's dsaupd/dseupd through arpack_eigenvector_sym_matrix
(ED_ARPACK.f90:35-250) with bmat='I', which='BE' (real; 'SR' complex), nev =
Neigen_arpack - 1, ncv = Neigen_arpack, maxitr = 300, tol = tolerance, mode 1, and
av = Hmultr; after dseupd it forms the relative residuals ||A x - lambda x|| / |lambda|
and ARPACK_diago keeps the converged pairs up to the first one outside the window
[E_1, E_1 + dEmax].

backend="device" (the default) runs the implicitly restarted Lanczos method of dsaupd on
the device (ed_arpack in libed_hip.so): basis, H.v, reorthogonalisation and restart
rotation stay in HBM, only the ncv x ncv tridiagonal work runs on the host.
backend="scipy" drives the ARPACK routines SciPy ships (scipy.sparse.linalg.eigsh) with
the device H.v through the host-vector C-ABI (ed_hmult_f64, two PCIe copies per H.v),
which is how the unchanged reference would call the library through Hmultr.
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np

from ._lib import ED_ARPACK_BE, ED_ARPACK_LA, ED_ARPACK_SA, ArpackOpts, check, dptr, iptr, lib
from .hclass import Sector, start_vector

__all__ = ["ArpackResult", "arpack_diago"]

_WHICH = {"BE": ED_ARPACK_BE, "SA": ED_ARPACK_SA, "LA": ED_ARPACK_LA}


@dataclasses.dataclass
class ArpackResult:
    evals: np.ndarray      # VALP, ascending, kept pairs only
    vectors: np.ndarray    # VECP, shape (nconv, dim)
    rel_resid: np.ndarray  # ||H x - E x|| / |E|
    nconv: int
    matvecs: int
    iters: int = 0         # implicit restarts + 1 (device backend)


def _window(w, v, res, dEmax):
    nconv = len(w)
    for i in range(1, nconv):   # first pair outside [E_1, E_1 + dEmax] ends the list
        if not (w[0] <= w[i] <= w[0] + dEmax):
            nconv = i
            break
    return nconv


def arpack_diago(sec: Sector, neigen_arpack: int, tolerance: float = 1e-12, dEmax: float = np.inf,
                 which: str = "BE", maxitr: int = 300, ncv: int | None = None, backend: str = "device",
                 start: np.ndarray | None = None) -> ArpackResult:
    """nev = min(dim - 1, Neigen_arpack) - 1 Ritz values by `which`; ncv defaults to the
    reference's Neigen_arpack (one more than nev, so a single shift per restart: few pairs
    converge within maxitr); a larger ncv converges faster.  start defaults to
    create_fix_initial_vector's deterministic vector."""
    n = sec.dim
    neig = min(n - 1, neigen_arpack)
    if neig < 2:
        raise ValueError("ARPACK_diago needs Neigen_arpack >= 2 and a sector of dimension >= 3")
    nev = neig - 1
    ncv = neig if ncv is None else min(n, max(neig, ncv))
    if which not in _WHICH:
        raise ValueError("which must be 'BE', 'SA' or 'LA'")
    if backend == "scipy":
        return _arpack_scipy(sec, nev, ncv, which, maxitr, tolerance, dEmax, start)
    if backend != "device":
        raise ValueError("backend must be 'device' or 'scipy'")
    o = ArpackOpts(int(nev), int(ncv), _WHICH[which], float(tolerance), int(maxitr))
    evals = np.zeros(nev)
    resid = np.zeros(nev)
    vecs = np.zeros((nev, n))
    nconv, iters, mv = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    st = None if start is None else np.ascontiguousarray(start, np.float64)
    if st is not None and st.shape != (n,):
        raise ValueError("start must have the sector's dimension")
    check(lib().ed_arpack(sec.handle, ctypes.byref(o), dptr(st) if st is not None else None, dptr(evals),
                          dptr(resid), dptr(vecs), n, ctypes.byref(nconv), ctypes.byref(iters), ctypes.byref(mv)))
    k = nconv.value
    w, v, res = evals[:k], vecs[:k], resid[:k]
    keep = _window(w, v, res, dEmax)
    return ArpackResult(w[:keep].copy(), v[:keep].copy(), res[:keep].copy(), keep, mv.value, iters.value)


def _arpack_scipy(sec: Sector, nev: int, ncv: int, which: str, maxitr: int, tolerance: float, dEmax: float,
                  start) -> ArpackResult:
    import scipy.sparse.linalg as sla
    n = sec.dim
    count = [0]

    def av(x):
        count[0] += 1
        return sec.hmult(np.asarray(x, np.float64).ravel())

    op = sla.LinearOperator((n, n), matvec=av, dtype=np.float64)
    which_sp = {"BE": "BE", "SA": "SA", "LA": "LA"}[which]
    v0 = start_vector(n) if start is None else np.asarray(start, np.float64)
    try:
        w, v = sla.eigsh(op, k=nev, which=which_sp, ncv=ncv, maxiter=maxitr, tol=tolerance, v0=v0)
    except sla.ArpackNoConvergence as e:   # info = 1: keep the iparam(5) converged pairs
        w, v = e.eigenvalues, e.eigenvectors
        if len(w) == 0:
            return ArpackResult(np.zeros(0), np.zeros((0, n)), np.zeros(0), 0, count[0])
    order = np.argsort(w)
    w, v = w[order], v[:, order].T.copy()
    res = np.array([np.linalg.norm(sec.hmult(v[j]) - w[j] * v[j]) / abs(w[j]) for j in range(len(w))])
    keep = _window(w, v, res, dEmax)
    return ArpackResult(w[:keep], v[:keep], res[:keep], keep, count[0])
