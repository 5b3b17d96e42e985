"""Derivatives: Hellmann-Feynman gradients of the ground-state energy and the bath-fit
distance with its analytic gradient (both computed by HIP kernels in libed_hip.so), plus
``torch.autograd.Function`` wrappers over them.

Parameterisation of the autograd wrappers (the synthetic Kanamori family of the
workloads): theta = (eps_0..eps_{nb-1}, V_0..V_{nb-1}, U), with bath site k coupled to
orbital k mod nc by V_k in both spins, U_ii = U, U_ij = U - 2J, J and the impurity level
fixed.  The reference fits the bath with finite differences of distance_func
(conj_grad_civelli.f90:282-303); these wrappers give the same derivatives analytically.

torch is used on CPU tensors only.  libed_hip.so is loaded before torch so that its HIP
calls bind to the ROCm runtime it was built against, not to the one bundled with torch.
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np

from ._lib import check, lib
from .aim import AIMParams, kanamori
from .hclass import Sector

lib()   # before torch (see the module docstring)

import torch  # noqa: E402

__all__ = ["HFGradient", "hf_gradient", "bath_distance", "GroundStateEnergy", "BathDistance", "kanamori_theta"]


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if a is not None else None


@dataclasses.dataclass
class HFGradient:
    level: np.ndarray   # [ns]: d E / d (level of site bit a): bath a < nb, impurity orbital a - nb
    v: np.ndarray       # [nb, nc]: d E / d Vbc(k, c)
    u: float            # d E / d U (U_ii = U, U_ij = U - 2J)


def hf_gradient(sec: Sector, psi: np.ndarray) -> HFGradient:
    """<psi| dH/dtheta |psi> / <psi|psi> on the GPU (ed_hf_gradient)."""
    m = sec.model
    psi = np.ascontiguousarray(psi, np.float64)
    lev = np.zeros(m.ns)
    v = np.zeros(m.nb * m.nc)
    u = ctypes.c_double()
    check(lib().ed_hf_gradient(sec.handle, sec.dim, _dp(psi), _dp(lev), _dp(v) if m.nb else None, ctypes.byref(u)))
    return HFGradient(lev, v.reshape(m.nb, m.nc), u.value)


def bath_distance(eps, vbc, wn, target, weight_expo: float = 2.0, fit_weight_power: float = 0.0,
                  fit_shift: float = 0.0, grad: bool = True):
    """(dist, d dist/d eps [nb], d dist/d Vbc [nb, nc]) from ed_bath_distance; target is
    [nw, 2, nc, nc] complex."""
    eps = np.ascontiguousarray(eps, np.float64)
    vbc = np.ascontiguousarray(vbc, np.float64)
    nb, nc = vbc.shape
    wn = np.ascontiguousarray(wn, np.float64)
    tgt = np.ascontiguousarray(np.asarray(target, np.complex128).reshape(len(wn), 2, nc, nc)).view(np.float64)
    d = ctypes.c_double()
    ge = np.zeros(nb) if grad else None
    gv = np.zeros(nb * nc) if grad else None
    check(lib().ed_bath_distance(nc, nb, _dp(eps), _dp(vbc), len(wn), _dp(wn), _dp(tgt), weight_expo,
                                 fit_weight_power, fit_shift, ctypes.byref(d), _dp(ge), _dp(gv)))
    return d.value, ge, (gv.reshape(nb, nc) if grad else None)


def kanamori_theta(theta: np.ndarray, nc: int, nb: int, J: float, eimp: float, cutoff: float = 1e-8) -> AIMParams:
    """The Kanamori model of theta = (eps[nb], V[nb], U)."""
    theta = np.asarray(theta, np.float64)
    return kanamori(nc, nb, float(theta[2 * nb]), J, theta[:nb], theta[nb:2 * nb], eimp=eimp, cutoff=cutoff)


def _theta_grad(g: HFGradient, nc: int, nb: int) -> np.ndarray:
    dv = np.array([g.v[k, k % nc] for k in range(nb)])
    return np.concatenate([g.level[:nb], dv, [g.u]])


class GroundStateEnergy(torch.autograd.Function):
    """E0 of sector (nup, ndn) of the Kanamori model of theta; backward = Hellmann-Feynman.

    spec: dict(nc, nb, J, eimp, sector=(nup, ndn), nitermax=..., tolerance=...)."""

    @staticmethod
    def forward(ctx, theta, spec):
        nc, nb = spec["nc"], spec["nb"]
        m = kanamori_theta(theta.detach().cpu().numpy(), nc, nb, spec["J"], spec["eimp"])
        with Sector(m, *spec["sector"]) as sec:
            res = sec.lanczos(nitermax=spec.get("nitermax", 300), neigen=1, tolerance=spec.get("tolerance", 1e-13))
            psi = sec.ground_state(res, 0)
            g = hf_gradient(sec, psi)
        ctx.save_for_backward(torch.as_tensor(_theta_grad(g, nc, nb), dtype=theta.dtype))
        return theta.new_tensor(float(res.evals[0]))

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out * g, None


class BathDistance(torch.autograd.Function):
    """distance_func of the bath theta = (eps[nb], V[nb]) (V_k on orbital k mod nc) to a
    target hybridisation; backward = the analytic gradient from the device kernel.

    spec: dict(nc, nb, wn, target, weight_expo, fit_weight_power, fit_shift)."""

    @staticmethod
    def forward(ctx, theta, spec):
        nc, nb = spec["nc"], spec["nb"]
        th = theta.detach().cpu().numpy()
        vbc = np.zeros((nb, nc))
        for k in range(nb):
            vbc[k, k % nc] = th[nb + k]
        d, ge, gv = bath_distance(th[:nb], vbc, spec["wn"], spec["target"], spec.get("weight_expo", 2.0),
                                  spec.get("fit_weight_power", 0.0), spec.get("fit_shift", 0.0))
        g = np.concatenate([ge, [gv[k, k % nc] for k in range(nb)]])
        ctx.save_for_backward(torch.as_tensor(g, dtype=theta.dtype))
        return theta.new_tensor(d)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out * g, None
