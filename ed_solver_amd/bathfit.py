"""The bath fit: the reference's quasi-Newton loop over the device distance and gradient.

The reference fits the bath with `minimize_civelli` (conj_grad_civelli.f90:8-306), a
variable-metric method that keeps the Hessian approximation as L D L^T, takes its
gradient by finite differences (forward, then central after the first stall: 2 or 4 *
nparam distance calls per gradient, :282-303) and is driven by
tools_minimization_wrapping.f90:229-233,362-374 (dfn = -1/2, xm = |x| + 1e-12,
hh = search_step, eps = dist_max, mode 1 = start from the identity).

Here the same method runs on the host with the distance AND its analytic gradient
from one device call (`ed_bath_distance`), so a gradient costs one call instead of
2 nparam; `gradient="fd"` keeps the reference's finite differences (over the device
distance) for comparison.  The method, as restated:

  direction   s = -(L D L^T)^{-1} g
  line search alpha0 = min(1, -2 df / g.s); while the value drops, accept and double
              (stop doubling when the new value, the current one and the one before the
              search look convex: f1 + f2 >= 2 f and 7 f1 + 5 f2 > 12 f); when it does
              not drop, halve, then interpolate
              quadratically (factor >= 0.1); stop when the step falls below
              eps / max_i(|s_i| / xm_i)
  update      with y = g+ - g, dgs = y.s and gs0 = g.s: if dgs <= 0 no update; else the
              BFGS update as two rank-one modifications of L D L^T (Gill-Murray), with
              the safeguarded pair of rank-one terms when dgs + alpha gs0 > 0.
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np

from ._lib import check, lib


def _dp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@dataclasses.dataclass
class FitResult:
    x: np.ndarray
    f: float
    iexit: int          # 1 step too small (converged), 2 no descent direction, 3 maxfn reached
    nfev: int           # distance evaluations (a gradient counts as one in analytic mode)
    niter: int


def _ldl_solve(L, D, g):
    """s = -(L D L^T)^{-1} g with L unit lower triangular."""
    n = len(g)
    y = np.empty(n)
    for i in range(n):
        y[i] = -g[i] - L[i, :i] @ y[:i]
    z = y / D
    s = np.empty(n)
    for i in range(n - 1, -1, -1):
        s[i] = z[i] - L[i + 1:, i] @ s[i + 1:]
    return s


def _rank1(L, D, sig, u, dmin):
    """L D L^T + sig u u^T in place (Gill-Murray); returns the new dmin."""
    n = len(u)
    u = u.copy()
    v = np.empty(n)
    for i in range(n):
        v[i] = u[i] - L[i, :i] @ v[:i]
    b = np.empty(n)
    for i in range(n):
        z = D[i] + sig * v[i] * v[i]
        if z <= 0.0:
            z = dmin
        dmin = min(dmin, z)
        b[i] = v[i] * sig / z
        sig = sig - b[i] * b[i] * z
        D[i] = z
    for i in range(n - 1):
        for j in range(i + 1, n):
            u[j] -= L[j, i] * v[i]
            L[j, i] += b[i] * u[j]
    return dmin


def minimize_qn(fun, x0, grad=None, dfn=-0.5, xm=None, hh=1e-5, eps=1e-8, maxfn=1000) -> FitResult:
    """Minimise fun(x) from x0.  grad(x) -> (f, g) gives the analytic gradient; without it
    the gradient is taken by forward differences of step hh * xm_i, then central ones
    after the first stall, as the reference does."""
    x = np.array(x0, dtype=np.float64)
    n = x.size
    xm = np.abs(x) + 1e-12 if xm is None else np.asarray(xm, np.float64)
    L = np.zeros((n, n))
    D = np.ones(n)
    dmin = 1.0
    nfev = 0
    idiff = 1

    def gradient(xp, fp):
        nonlocal nfev
        if grad is not None:
            nfev += 1
            return grad(xp)[1]
        g = np.empty(n)
        w = xp.copy()
        for i in range(n):
            z = hh * xm[i]
            if idiff == 1:
                w[i] += z
                g[i] = (fun(w) - fp) / z
                w[i] -= z
            else:
                w[i] += z
                f1 = fun(w)
                w[i] -= 2 * z
                f2 = fun(w)
                w[i] += z
                g[i] = (f1 - f2) / (2 * z)
        nfev += n * idiff
        return g

    f0 = 0.0
    f = fun(x)
    nfev += 1
    df = dfn if dfn > 0 else (abs(dfn * f) if dfn < 0 else f - f0)
    if df <= 0:
        df = 1.0
    g = gradient(x, f)
    itn = 0
    while True:
        if nfev >= maxfn:
            return FitResult(x, f, 3, nfev, itn)
        itn += 1
        s = _ldl_solve(L, D, g)
        zmax = np.max(np.abs(s) / xm)
        aeps = eps / zmax if zmax > 0 else np.inf
        gs0 = g @ s
        stalled = gs0 >= 0.0
        exit_code = 2
        if not stalled:
            exit_code = 1
            alpha = min(1.0, -2.0 * df / gs0)
            ff = f
            tot = 0.0
            mode = 0           # 0: may double, 1: after interpolation, 2: stop doubling
            f1 = f2 = f
            while True:        # line search
                if nfev >= maxfn:
                    return FitResult(x, f, 3, nfev, itn)
                w = x + alpha * s
                f1 = fun(w)
                nfev += 1
                if f1 < f:
                    f2 = f
                    tot += alpha
                    x, f = w, f1
                    if mode == 0:
                        while True:   # doubling
                            if nfev >= maxfn:
                                return FitResult(x, f, 3, nfev, itn)
                            w = x + alpha * s
                            f1 = fun(w)
                            nfev += 1
                            if f1 >= f:
                                break
                            if f1 + f2 >= f + f and 7.0 * f1 + 5.0 * f2 > 12.0 * f:
                                mode = 2
                            tot += alpha
                            alpha *= 2.0
                            x, f = w, f1
                            if mode == 2:
                                break
                        break
                    if mode == 1 and tot < aeps:
                        stalled = True
                    break
                # no decrease: halve, then interpolate
                if alpha < aeps:
                    stalled = True
                    break
                if nfev >= maxfn:
                    return FitResult(x, f, 3, nfev, itn)
                alpha *= 0.5
                w = x + alpha * s
                f2 = fun(w)
                nfev += 1
                if f2 < f:
                    tot += alpha
                    x, f = w, f2
                    if tot < aeps:
                        stalled = True
                    break
                z = 0.1
                if f1 + f > f2 + f2:
                    z = 1.0 + 0.5 * (f - f1) / (f + f1 - f2 - f2)
                alpha *= max(z, 0.1)
                mode = 1
        if stalled:
            if grad is None and idiff == 1:
                idiff = 2          # the reference switches to central differences and restarts
                g = gradient(x, f)
                continue
            return FitResult(x, f, exit_code, nfev, itn)
        alpha = tot
        g_old = g
        g = gradient(x, f)
        df = ff - f
        dgs = g @ s - gs0
        if dgs <= 0.0:
            continue
        if dgs + alpha * gs0 <= 0.0:
            dmin = _rank1(L, D, 1.0 / (alpha * dgs), g - g_old, dmin)
            dmin = _rank1(L, D, 1.0 / gs0, g_old, dmin)
        else:
            zz = alpha / (dgs - alpha * gs0)
            z = dgs * zz - 1.0
            dmin = _rank1(L, D, 1.0 / (zz * dgs * dgs), z * g_old + g, dmin)
            dmin = _rank1(L, D, -zz, g_old, dmin)


def fit_bath(eps0, vbc0, wn, target, weight_expo: float = 2.0, fit_weight_power: float = 0.0,
             fit_shift: float = 0.0, vmask=None, gradient: str = "analytic", search_step: float = 1e-5,
             dist_max: float = 1e-10, maxfn: int = 2000):
    """Fit a diagonal bath (levels eps[nb], hybridisations vbc[nb, nc]) to the target
    hybridisation target[nw, 2, nc, nc] with the distance of diff_correl
    (bath_class_hybrid.f90:1151-1204).  vmask[nb, nc] (default: the nonzero entries of
    vbc0) selects the free hybridisations.  Returns (eps, vbc, FitResult)."""
    eps0 = np.asarray(eps0, np.float64)
    vbc0 = np.asarray(vbc0, np.float64)
    nb, nc = vbc0.shape
    vmask = (vbc0 != 0.0) if vmask is None else np.asarray(vmask, bool)
    wn = np.ascontiguousarray(wn, np.float64)
    tgt = np.ascontiguousarray(np.asarray(target, np.complex128).reshape(len(wn), 2, nc, nc)).view(np.float64)

    def unpack(x):
        e = x[:nb].copy()
        v = np.zeros((nb, nc))
        v[vmask] = x[nb:]
        return e, v

    def dist_grad(x, want_grad=True):
        e, v = unpack(x)
        d = ctypes.c_double()
        ge = np.zeros(nb) if want_grad else None
        gv = np.zeros(nb * nc) if want_grad else None
        check(lib().ed_bath_distance(nc, nb, _dp(e), _dp(np.ascontiguousarray(v)), len(wn), _dp(wn), _dp(tgt),
                                     weight_expo, fit_weight_power, fit_shift, ctypes.byref(d), _dp(ge), _dp(gv)))
        if not want_grad:
            return d.value, None
        return d.value, np.concatenate([ge, gv.reshape(nb, nc)[vmask]])

    x0 = np.concatenate([eps0, vbc0[vmask]])
    fun = lambda x: dist_grad(x, False)[0]
    res = minimize_qn(fun, x0, grad=dist_grad if gradient == "analytic" else None, dfn=-0.5,
                      xm=np.abs(x0) + 1e-12, hh=search_step, eps=dist_max, maxfn=maxfn)
    e, v = unpack(res.x)
    return e, v, res
