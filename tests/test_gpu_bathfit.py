"""The bath-fit loop (minimize_civelli's quasi-Newton method) over the device distance:
analytic gradient against the reference's finite differences, and the minimiser itself
on functions with a known minimum."""
import numpy as np
import pytest

import oracle.bath as ob
from ed_solver_amd.bathfit import fit_bath, minimize_qn

pytestmark = pytest.mark.gpu


def test_minimiser_on_known_functions(gpu):
    """Rosenbrock and a shifted quadratic: the restated method reaches the minimum with
    the analytic gradient and with the reference's finite differences."""
    def rosen(x):
        return 100.0 * (x[1] - x[0] ** 2) ** 2 + (1 - x[0]) ** 2

    def rosen_g(x):
        return rosen(x), np.array([-400 * x[0] * (x[1] - x[0] ** 2) - 2 * (1 - x[0]), 200 * (x[1] - x[0] ** 2)])

    r = minimize_qn(rosen, [-1.2, 1.0], grad=rosen_g, eps=1e-12, maxfn=5000)
    assert np.max(np.abs(r.x - 1.0)) < 1e-5
    r_fd = minimize_qn(rosen, [-1.2, 1.0], hh=1e-7, eps=1e-12, maxfn=20000)
    assert np.max(np.abs(r_fd.x - 1.0)) < 1e-3
    A = np.array([[3.0, 0.5, 0.0], [0.5, 2.0, 0.3], [0.0, 0.3, 1.0]])
    b = np.array([1.0, -2.0, 0.5])
    q = lambda x: 0.5 * x @ A @ x - b @ x
    rq = minimize_qn(q, np.ones(3), grad=lambda x: (q(x), A @ x - b), eps=1e-14, maxfn=500)
    assert np.max(np.abs(rq.x - np.linalg.solve(A, b))) < 1e-7


def test_fit_recovers_a_bath(gpu):
    """Target = the hybridisation of a known 2-orbital bath (4 sites, 4096 w_n): from a
    perturbed start the analytic-gradient fit drives the distance down by 8 orders with
    fewer device calls than the finite-difference loop of the reference reaching less."""
    nc, nb = 2, 4
    r = np.random.default_rng(3)
    eps_t = np.array([-1.0, -0.3, 0.4, 1.1])
    vbc_t = np.zeros((nb, nc))
    for k in range(nb):
        vbc_t[k, k % nc] = [0.45, 0.3, 0.35, 0.5][k]
    beta = 50.0
    wn = np.pi / beta * (2 * np.arange(1, 513) - 1)
    dt = ob.delta(eps_t, vbc_t, wn)
    target = np.stack([dt, dt], axis=1)        # [nw, 2 spins, nc, nc]
    eps0 = eps_t + r.uniform(-0.2, 0.2, nb)
    vbc0 = vbc_t * (1 + r.uniform(-0.3, 0.3, (nb, nc)))
    d0 = ob.distance(eps0, vbc0, wn, target)
    e, v, res = fit_bath(eps0, vbc0, wn, target, dist_max=1e-14, maxfn=3000)
    print(f"start {d0:.3e}  analytic fit {res.f:.3e} in {res.nfev} calls, {res.niter} iterations", flush=True)
    assert res.f < 1e-8 * d0
    assert abs(ob.distance(e, v, wn, target) - res.f) <= 1e-12 * max(res.f, 1e-300) + 1e-15
    e_fd, v_fd, res_fd = fit_bath(eps0, vbc0, wn, target, gradient="fd", dist_max=1e-14, maxfn=res.nfev)
    print(f"finite differences with the same number of calls: {res_fd.f:.3e}", flush=True)
    assert res.f <= res_fd.f
