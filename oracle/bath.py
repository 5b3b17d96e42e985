"""Bath-fit distance and the reference's finite-difference gradient in NumPy --
TEST INFRASTRUCTURE ONLY (the checker of ed_bath_distance).

* ``delta``     bath2hybrid, normal state, diagonal bath (bath_class_hybrid.f90:587-903):
                Delta_ij(i w_n) = sum_k Vbc(k,i) Vbc(k,j) / (i w_n - eps_k).
* ``distance``  distance_func -> diff_correl with default options (no asymmetry term,
                no frequency window, fit_nw = N_w; correl_class.f90:1061-1160):
                |sum_n sum_ij |Delta_ij - T_ij|^p / (|w_n - w_1| + s)^q| / (N_w N^2), with
                s = |w_1| when |fit_shift| < 1e-10 and N = 2 nc (both spin blocks; the
                spin off-diagonal blocks of the target are zero and contribute nothing).
* ``fd_gradient_civelli``  the gradient minimize_civelli takes (conj_grad_civelli.f90:
                282-303): step z_i = search_step * xm_i with xm_i = |x_i| + 1e-12
                (tools_minimization_wrapping.f90:374); forward differences
                (f(x + z e_i) - f(x)) / z while idiff = 1, central differences
                (f(x + z e_i) - f(x - z e_i)) / (2 z) once idiff = 2.

Parity: ``distance`` is pinned by its closed form on a one-site bath (tests/test_oracle_bath.py:
a single pole gives |V^2 / (i w - e) - T|^p summed by hand) and the FD gradient by a
quadratic function whose gradient is known exactly.
"""
from __future__ import annotations

import numpy as np


def delta(eps: np.ndarray, vbc: np.ndarray, wn: np.ndarray) -> np.ndarray:
    """[nw, nc, nc] complex."""
    eps = np.asarray(eps, float)
    vbc = np.asarray(vbc, float)
    g = 1.0 / (1j * np.asarray(wn, float)[:, None] - eps[None, :])            # [nw, nb]
    return np.einsum("ki,kj,nk->nij", vbc, vbc, g)


def distance(eps, vbc, wn, target, weight_expo=2.0, fit_weight_power=0.0, fit_shift=0.0) -> float:
    """target: [nw, 2, nc, nc] complex (spin up, spin down)."""
    wn = np.asarray(wn, float)
    nc = np.asarray(vbc).shape[1]
    d = delta(eps, vbc, wn)
    shift = abs(wn[0]) if abs(fit_shift) < 1e-10 else fit_shift
    wgt = 1.0 / (np.abs(wn - wn[0]) + shift) ** fit_weight_power
    s = 0.0
    for sg in range(2):
        s += np.sum(np.abs(d - target[:, sg]) ** weight_expo * wgt[:, None, None])
    return abs(s) / (len(wn) * (2 * nc) ** 2)


def fd_gradient_civelli(f, x: np.ndarray, search_step: float, idiff: int = 2) -> np.ndarray:
    x = np.array(x, float)
    f0 = f(x)
    g = np.zeros_like(x)
    for i in range(len(x)):
        z = search_step * (abs(x[i]) + 1e-12)
        w = x.copy()
        w[i] += z
        f1 = f(w)
        if idiff == 1:
            g[i] = (f1 - f0) / z
        else:
            w[i] -= 2 * z
            g[i] = (f1 - f(w)) / (2 * z)
    return g
