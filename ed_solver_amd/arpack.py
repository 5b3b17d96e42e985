"""ARPACK_diago (ED_ARPACK.f90:475-570): implicitly restarted Lanczos from ARPACK with the
device H.v as the operator.

The reference calls ARPACK96's dsaupd/dseupd through arpack_eigenvector_sym_matrix
(ED_ARPACK.f90:35-250) with bmat='I', which='BE' (real; 'SR' complex), nev =
Neigen_arpack - 1, ncv = Neigen_arpack, maxitr = 300, tol = tolerance, mode 1, and
av = Hmultr; after dseupd it forms the relative residuals ||A x - lambda x|| / |lambda|
and ARPACK_diago keeps the converged pairs up to the first one outside the window
[E_1, E_1 + dEmax].  Here the same ARPACK routines (the copy SciPy ships, as
scipy.sparse.linalg.eigsh) drive the H.v of libed_hip.so through its C-ABI (ed_hmult_f64:
host vectors in and out, as av = Hmultr does); the residuals use the device H.v too.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .hclass import Sector, start_vector

__all__ = ["ArpackResult", "arpack_diago"]


@dataclasses.dataclass
class ArpackResult:
    evals: np.ndarray      # VALP, ascending, kept pairs only
    vectors: np.ndarray    # VECP, shape (nconv, dim)
    rel_resid: np.ndarray  # ||H x - E x|| / |E|
    nconv: int
    matvecs: int


def arpack_diago(sec: Sector, neigen_arpack: int, tolerance: float = 1e-12, dEmax: float = np.inf,
                 which: str = "BE", maxitr: int = 300, ncv: int | None = None) -> ArpackResult:
    """ncv defaults to the reference's Neigen_arpack (one more than nev, so a single shift
    per restart: few pairs converge within maxitr); a larger ncv converges faster."""
    import scipy.sparse.linalg as sla
    n = sec.dim
    neig = min(n - 1, neigen_arpack)
    if neig < 2:
        raise ValueError("ARPACK_diago needs Neigen_arpack >= 2 and a sector of dimension >= 3")
    nev = neig - 1
    ncv = neig if ncv is None else min(n, max(neig, ncv))
    count = [0]

    def av(x):
        count[0] += 1
        return sec.hmult(np.asarray(x, np.float64).ravel())

    op = sla.LinearOperator((n, n), matvec=av, dtype=np.float64)
    try:
        w, v = sla.eigsh(op, k=nev, which=which, ncv=ncv, maxiter=maxitr, tol=tolerance, v0=start_vector(n))
    except sla.ArpackNoConvergence as e:   # info = 1: keep the iparam(5) converged pairs
        w, v = e.eigenvalues, e.eigenvectors
        if len(w) == 0:
            return ArpackResult(np.zeros(0), np.zeros((0, n)), np.zeros(0), 0, count[0])
    order = np.argsort(w)
    w, v = w[order], v[:, order].T.copy()
    res = np.array([np.linalg.norm(sec.hmult(v[j]) - w[j] * v[j]) / abs(w[j]) for j in range(len(w))])
    nconv = len(w)
    for i in range(1, nconv):   # first pair outside [E_1, E_1 + dEmax] ends the list
        if not (w[0] <= w[i] <= w[0] + dEmax):
            nconv = i
            break
    return ArpackResult(w[:nconv], v[:nconv], res[:nconv], nconv, count[0])
