"""Derivatives on the GPU, through the C-ABI: Hellmann-Feynman gradients against brute
force <psi|dH|psi> and central finite differences of the GPU ground-state energy
(h = 1e-5), the bath-fit distance and its analytic gradient against the NumPy oracle and
the reference's finite-difference rule, and the torch.autograd wrappers (gradcheck)."""
import json
import os

import numpy as np
import pytest

import oracle.brute as bf
from oracle import bath as obath
from ed_solver_amd.aim import AIMParams, kanamori

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "models.json")


def golden(name):
    with open(GOLD) as f:
        d = json.load(f)[name]
    return AIMParams.from_json(d["model"]), tuple(d["sector"])


def unit_models(m: AIMParams):
    """dH/dtheta for every parameter as the Hamiltonian of a model holding only that term
    at unit strength (H is linear in each level, hybridisation and the U scale)."""
    nc, nb = m.nc, m.nb
    z = dict(ec=np.zeros((2, nc, nc)), eb=np.zeros((2, nb, nb)), vbc=np.zeros((2, nb, nc)), u=np.zeros((nc, nc)),
             jj=np.zeros((nc, nc)), jhund=0.0, cutoff=1e-12)
    out = {}
    for a in range(m.ns):
        d = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in z.items()}
        if a < nb:
            d["eb"][:, a, a] = 1.0
        else:
            d["ec"][:, a - nb, a - nb] = 1.0
        out[("level", a)] = AIMParams(nc, nb, **d)
    for k in range(nb):
        for c in range(nc):
            d = {kk: (v.copy() if isinstance(v, np.ndarray) else v) for kk, v in z.items()}
            d["vbc"][:, k, c] = 1.0
            out[("v", k, c)] = AIMParams(nc, nb, **d)
    d = {kk: (v.copy() if isinstance(v, np.ndarray) else v) for kk, v in z.items()}
    d["u"] = (m.u != 0).astype(float)
    out[("u",)] = AIMParams(nc, nb, **d)
    return out


@pytest.mark.parametrize("case", [(2, 3, (2, 3)), (1, 5, (3, 3)), (3, 3, (3, 3))])
def test_hf_gradient_vs_brute_force(gpu, case):
    from ed_solver_amd.grad import hf_gradient
    nc, nb, sec = case
    rng = np.random.default_rng(sum(sec) + nc)
    m = kanamori(nc, nb, 2.5, 0.4 if nc > 1 else 0.0, rng.uniform(-1, 1, nb), rng.uniform(0.2, 0.6, nb))
    s = gpu.Sector(m, *sec)
    H = bf.sector_hamiltonian(m, *sec).toarray()
    w, V = np.linalg.eigh(H)
    dH = {key: bf.sector_hamiltonian(mm, *sec) for key, mm in unit_models(m).items()}
    for psi in (V[:, 0], rng.normal(size=s.dim)):   # any vector: <psi|dH|psi>/<psi|psi>
        g = hf_gradient(s, psi)
        nrm = psi @ psi
        for key, D in dH.items():
            ref = psi @ (D @ psi) / nrm
            got = g.level[key[1]] if key[0] == "level" else (g.v[key[1], key[2]] if key[0] == "v" else g.u)
            assert abs(got - ref) < 1e-12 * max(1.0, abs(ref)), key
    s.close()


def e0_gpu(gpu, theta, spec):
    from ed_solver_amd.grad import kanamori_theta
    m = kanamori_theta(theta, spec["nc"], spec["nb"], spec["J"], spec["eimp"])
    with gpu.Sector(m, *spec["sector"]) as s:
        return s.lanczos(nitermax=400, neigen=1, tolerance=1e-14).evals[0]


def theta_of(m):
    nb, nc = m.nb, m.nc
    return np.concatenate([np.diag(m.eb[0]), [m.vbc[0, k, k % nc] for k in range(nb)], [m.u[0, 0]]])


def spec_of(m, sector, name):
    J = m.jj[0, 1] if m.nc > 1 else 0.0
    return dict(nc=m.nc, nb=m.nb, J=J, eimp=m.ec[0, 0, 0], sector=sector, nitermax=400, tolerance=1e-14)


def test_hf_gradient_vs_fd_workload_b(gpu):
    """dE0/d(eps_k, V_k, U) by Hellmann-Feynman against central differences of the GPU
    ground-state energy (h = 1e-5), workload B (dim 63,504)."""
    import torch
    from ed_solver_amd.grad import GroundStateEnergy
    m, sec = golden("B")
    spec = spec_of(m, sec, "B")
    th = theta_of(m)
    t = torch.tensor(th, dtype=torch.float64, requires_grad=True)
    e = GroundStateEnergy.apply(t, spec)
    e.backward()
    g = t.grad.numpy()
    h = 1e-5
    for i in range(len(th)):
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        fd = (e0_gpu(gpu, tp, spec) - e0_gpu(gpu, tm, spec)) / (2 * h)
        print(f"theta[{i}] HF {g[i]: .10f} FD {fd: .10f}", flush=True)
        assert abs(g[i] - fd) < 2e-6, i


def test_hf_gradient_workload_e(gpu):
    """Workload E (2 orbitals + 12 bath sites, (7,7), dim 11,778,624): the Hellmann-Feynman
    gradient from the device ground state against central differences (h = 1e-5) for the
    first bath level, the first coupling and U."""
    from ed_solver_amd.grad import hf_gradient
    m, sec = golden("E")
    spec = spec_of(m, sec, "E")
    th = theta_of(m)
    with gpu.Sector(m, *sec) as s:
        res = s.lanczos(nitermax=400, neigen=1, tolerance=1e-14)
        psi = s.ground_state(res, 0)
        g = hf_gradient(s, psi)
    nb, nc = m.nb, m.nc
    gt = np.concatenate([g.level[:nb], [g.v[k, k % nc] for k in range(nb)], [g.u]])
    h = 1e-5
    for i in (0, nb, 2 * nb):
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        fd = (e0_gpu(gpu, tp, spec) - e0_gpu(gpu, tm, spec)) / (2 * h)
        print(f"E theta[{i}] HF {gt[i]: .10f} FD {fd: .10f}", flush=True)
        assert abs(gt[i] - fd) < 5e-6, i


def bath_case(nc=2, nb=12, nw=4096, beta=200.0, seed=9):
    rng = np.random.default_rng(seed)
    wn = np.pi / beta * (2 * np.arange(1, nw + 1) - 1)
    eps = rng.uniform(-2, 2, nb)
    vbc = np.zeros((nb, nc))
    for k in range(nb):
        vbc[k, k % nc] = rng.uniform(0.1, 0.5)
    eps_t = eps + rng.normal(scale=0.2, size=nb)
    v_t = vbc * (1 + rng.normal(scale=0.2, size=(nb, 1)))
    d_t = obath.delta(eps_t, v_t, wn)
    tgt = np.stack([d_t, d_t * (1 + 0.05j)], axis=1)   # spin down slightly different
    return eps, vbc, wn, tgt


@pytest.mark.parametrize("opts", [(2.0, 0.0, 0.0), (1.0, 1.0, 0.0), (4.0, 0.5, 0.3)])
def test_bath_distance_and_gradient(gpu, opts):
    from ed_solver_amd.grad import bath_distance
    p, q, s = opts
    eps, vbc, wn, tgt = bath_case()
    d, ge, gv = bath_distance(eps, vbc, wn, tgt, p, q, s)
    ref = obath.distance(eps, vbc, wn, tgt, p, q, s)
    assert abs(d - ref) < 1e-12 * ref
    nb, nc = vbc.shape
    mask = vbc != 0
    x0 = np.concatenate([eps, vbc[mask]])

    def f(x):
        v = vbc.copy()
        v[mask] = x[nb:]
        return obath.distance(x[:nb], v, wn, tgt, p, q, s)

    g = np.concatenate([ge, gv[mask]])
    central = obath.fd_gradient_civelli(f, x0, 1e-5, idiff=2)
    forward = obath.fd_gradient_civelli(f, x0, 1e-7, idiff=1)
    scale = np.max(np.abs(g))
    print(f"p={p}: max|g - central| {np.max(np.abs(g - central)) / scale:.2e}, "
          f"max|g - forward| {np.max(np.abs(g - forward)) / scale:.2e}", flush=True)
    assert np.max(np.abs(g - central)) < 1e-6 * scale
    assert np.max(np.abs(g - forward)) < 1e-4 * scale


def test_autograd_harness_gradcheck(gpu):
    """torch.autograd.gradcheck (finite differences inside torch) on both wrappers."""
    import torch
    from ed_solver_amd.grad import BathDistance, GroundStateEnergy
    m, sec = golden("A")
    spec = spec_of(m, sec, "A")
    t = torch.tensor(theta_of(m), dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(lambda x: GroundStateEnergy.apply(x, spec), (t,), eps=1e-5, atol=1e-6, rtol=1e-5)
    eps, vbc, wn, tgt = bath_case(nc=1, nb=5, nw=1024, beta=50.0)
    bspec = dict(nc=1, nb=5, wn=wn, target=tgt, weight_expo=2.0, fit_weight_power=0.0, fit_shift=0.0)
    tb = torch.tensor(np.concatenate([eps, vbc[:, 0]]), dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(lambda x: BathDistance.apply(x, bspec), (tb,), eps=1e-6, atol=1e-8, rtol=1e-5)
