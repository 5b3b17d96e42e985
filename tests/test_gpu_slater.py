"""The Slater 4-index impurity interaction (flag_slater_int) on the GPU: H.v against the
brute-force Hamiltonian for random tensors, and the Kanamori workloads B and C written as
a Slater tensor on a model with U = J = 0 against the models themselves (Lanczos E0,
and G(i w_n) on a small model)."""
import json
import os

import numpy as np
import pytest

import oracle.slater as osl
from ed_solver_amd.aim import AIMParams, kanamori, slater_tensor

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "models.json")


def golden(name):
    with open(GOLD) as f:
        d = json.load(f)[name]
    return AIMParams.from_json(d["model"]), tuple(d["sector"])


def without_interaction(m):
    z = np.zeros((m.nc, m.nc))
    return AIMParams(m.nc, m.nb, m.ec, m.eb, m.vbc, z, z, 0.0, m.cutoff)


@pytest.mark.parametrize("nc,nb,sec,jh", [(2, 3, (3, 2), 0.0), (2, 4, (3, 3), 0.4), (3, 2, (3, 2), 0.3)])
def test_slater_hmult_vs_brute_force(gpu, nc, nb, sec, jh):
    r = np.random.default_rng(nc * 5 + nb)
    m = kanamori(nc, nb, 2.0, jh, r.uniform(-1, 1, nb), r.uniform(0.2, 0.6, nb))
    u4 = r.uniform(-0.5, 0.5, (nc,) * 4)
    s = gpu.Sector(m, *sec, slater=u4)
    H = osl.sector_hamiltonian(m, u4, *sec)
    print(f"Slater H.v nc={nc} nb={nb} {sec}: dim {s.dim}", flush=True)
    for _ in range(2):
        v = r.standard_normal(s.dim)
        ref = H @ v
        assert np.max(np.abs(s.hmult(v) - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))
    s.close()
    # Hermitian: U(s1, s2, s3, s4) = U(s4, s3, s2, s1) (the operator string's adjoint)
    u4 = (u4 + u4.transpose(3, 2, 1, 0)) / 2
    s = gpu.Sector(m, *sec, slater=u4)
    H = osl.sector_hamiltonian(m, u4, *sec)
    assert np.max(np.abs(H - H.T)) < 1e-14
    res = s.lanczos(nitermax=300, neigen=1, tolerance=1e-13)
    assert abs(res.evals[0] - np.linalg.eigvalsh(H)[0]) < 1e-10
    s.close()


@pytest.mark.parametrize("name", ["B", "C"])
def test_slater_kanamori_workload_matches_model(gpu, name):
    """The workload's U, U' and J as a Slater tensor on the model without interaction:
    the same Lanczos E0 and the same alpha / beta to rounding (the matrices are equal)."""
    m, sec = golden(name)
    sr = gpu.Sector(m, *sec)
    ss = gpu.Sector(without_interaction(m), *sec, slater=slater_tensor(m))
    print(f"{name}: dim {sr.dim}", flush=True)
    rr = sr.lanczos(nitermax=300, neigen=1, tolerance=1e-12)
    rs = ss.lanczos(nitermax=300, neigen=1, tolerance=1e-12)
    assert abs(rr.evals[0] - rs.evals[0]) < 1e-10
    ar, br = sr.lanczos_run(40)
    as_, bs = ss.lanczos_run(40)
    assert np.max(np.abs(ar - as_)) < 1e-9 and np.max(np.abs(br - bs)) < 1e-9
    sr.close(); ss.close()


def test_slater_green_matches_model(gpu):
    from ed_solver_amd.solver import compute_green_matrix, get_eigen_energies
    r = np.random.default_rng(9)
    m = kanamori(2, 2, 2.2, 0.35, r.uniform(-1, 1, 2), r.uniform(0.2, 0.6, 2))
    m0, u4 = without_interaction(m), slater_tensor(m)
    beta = 8.0
    wn = np.pi / beta * (2 * np.arange(1, 17) - 1)
    st = get_eigen_energies(m, neigen=36)
    st0 = get_eigen_energies(m0, neigen=36, slater=u4)
    assert np.max(np.abs(np.sort([x.energy for x in st]) - np.sort([x.energy for x in st0]))) < 1e-11
    G = compute_green_matrix(m, st, beta, wn)
    G0 = compute_green_matrix(m0, st0, beta, wn, slater=u4)
    assert np.max(np.abs(G[0, 0])) > 1e-2
    assert np.max(np.abs(G - G0)) < 1e-9
