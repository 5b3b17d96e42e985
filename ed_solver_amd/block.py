"""Block_Lanczos_diagonalize (Block_Lanczos.f90:37-355) on the device.

The reference drives BLZPACK's BLZDRD by reverse communication and keeps the converged
eigenpairs within dEmax of the lowest one (:300-320).  Here the whole block Lanczos
runs in libed_hip.so (ed_block_lanczos): block H.v, FP64 MFMA Gram / update
contractions for the full reorthogonalisation and Cholesky-QR2, Ritz values of the band
matrix T on the host.
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np

from ._lib import BlockOpts, check, lib
from .hclass import Sector

__all__ = ["BlockResult", "block_lanczos", "block_opts"]


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@dataclasses.dataclass
class BlockResult:
    evals: np.ndarray      # VALP(:,1)
    resid: np.ndarray      # VALP(:,2): ||H x_k - E_k x_k||
    vectors: np.ndarray    # VECP, shape (nev, dim)
    nconv: int             # ISTOR(67)
    steps: int             # ISTOR(19)
    full_passes: int = 0   # steps that projected against every earlier block
    ms: float = 0.0        # host wall time of the call
    ritz_ms: float = 0.0   # of which in the Ritz tests (band solver)


def block_opts(nblock: int = 8, neig: int = 8, nitermax: int = 128, tolerance: float = 1e-12,
               check_every: int = 2, reorth: str = "partial") -> BlockOpts:
    o = BlockOpts()
    check(lib().ed_block_default_opts(ctypes.byref(o)))
    o.block_size, o.nev, o.max_steps, o.tolerance, o.check_every = nblock, neig, nitermax, tolerance, check_every
    o.reorth = {"partial": 1, "full": 0}[reorth]
    return o


def block_lanczos(sec: Sector, nblock: int = 8, neig: int = 8, nitermax: int = 128, tolerance: float = 1e-12,
                  check_every: int = 2, dEmax: float | None = None, vectors: bool = True, reorth: str = "partial"):
    """(evals, vectors) of the lowest `neig` states; with dEmax, only those within dEmax of
    the lowest (Block_Lanczos.f90:300-320).  reorth: "partial" (the block omega-recurrence
    decides when to project against every earlier block, as BLZPACK) or "full" (every
    step).  The full result is in `.last`."""
    nev = min(neig, sec.dim)
    o = block_opts(min(nblock, sec.dim), nev, nitermax, tolerance, check_every, reorth)
    ev, res = np.zeros(nev), np.zeros(nev)
    vec = np.zeros((nev, sec.dim)) if vectors else None
    nconv, steps = ctypes.c_int(), ctypes.c_int()
    if dEmax is not None and vectors:
        # only the vectors within dEmax of the lowest state cross to the host
        # (ed_block_lanczos_window); the other rows of .last.vectors stay zero
        nkept = ctypes.c_int()
        check(lib().ed_block_lanczos_window(sec.handle, ctypes.byref(o), float(dEmax), _dptr(ev), _dptr(res),
                                            _dptr(vec), sec.dim, ctypes.byref(nconv), ctypes.byref(steps),
                                            ctypes.byref(nkept)))
    else:
        check(lib().ed_block_lanczos(sec.handle, ctypes.byref(o), _dptr(ev), _dptr(res),
                                     _dptr(vec) if vectors else None, sec.dim, ctypes.byref(nconv),
                                     ctypes.byref(steps)))
    full, ms, rms = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
    check(lib().ed_block_last_stats(None, ctypes.byref(full), ctypes.byref(ms), ctypes.byref(rms)))
    block_lanczos.last = BlockResult(ev, res, vec, nconv.value, steps.value, full.value, ms.value, rms.value)
    keep = np.ones(nev, bool) if dEmax is None else (ev - ev[0] <= dEmax)   # ascending: a leading run
    return ev[keep], (vec[keep] if vectors else None)
