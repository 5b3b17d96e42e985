"""Block Lanczos (Block_Lanczos_diagonalize, Block_Lanczos.f90:37-355) and its FP64 MFMA
contractions on the GPU, through the C-ABI.

Bars: the Gram (matrix cores) and update (vector unit) products against NumPy to 1e-13
relative (this also pins the matrix-core lane layout the Gram kernel assumes); the block Lanczos eigenvalues against
LAPACK / ARPACK on the oracle's H.v (oracle.lowest_eigs) to 1e-10, residuals
||H x - E x|| below 1e-6, Ritz vectors orthonormal to 1e-10."""
import ctypes
import json
import os

import numpy as np
import pytest

import oracle
from ed_solver_amd.aim import AIMParams, kanamori

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "models.json")


def golden(name):
    with open(GOLD) as f:
        d = json.load(f)[name]
    return AIMParams.from_json(d["model"]), tuple(d["sector"])


def dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def rel(a, b):
    return np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b)))


@pytest.mark.parametrize("n,p,q", [(1000, 37, 8), (4096, 16, 16), (777, 5, 3), (70001, 130, 8), (64, 1, 1)])
def test_mfma_gram_and_update(gpu, n, p, q):
    from hipmem import DevBuf
    from ed_solver_amd._lib import check, lib
    rng = np.random.default_rng(n + p + q)
    ld = (n + 63) // 64 * 64
    X = rng.uniform(-1, 1, (p, ld))
    Y = rng.uniform(-1, 1, (q, ld))
    X[:, n:] = np.nan   # padding rows must never be read
    Y[:, n:] = np.nan
    print(f"gram/update n={n} p={p} q={q}", flush=True)
    dX, dY = DevBuf.from_array(X), DevBuf.from_array(Y)
    G = np.zeros(p * q)
    check(lib().ed_gram_f64_dev(n, p, q, dX.addr, ld, dY.addr, ld, dptr(G), None))
    assert rel(G.reshape(p, q), X[:, :n] @ Y[:, :n].T) < 1e-13
    C = rng.uniform(-1, 1, (p, q))
    check(lib().ed_xc_f64_dev(n, p, q, dX.addr, ld, dptr(C), -0.5, 1.0, dY.addr, ld, None))
    Yn = dY.to_array(np.float64, q * ld).reshape(q, ld)
    assert rel(Yn[:, :n], Y[:, :n] - 0.5 * (X[:, :n].T @ C).T) < 1e-13
    assert np.all(np.isnan(Yn[:, n:]))
    # overwrite (beta = 0) and in place (Y = X, p = q)
    check(lib().ed_xc_f64_dev(n, p, q, dX.addr, ld, dptr(C), 2.0, 0.0, dY.addr, ld, None))
    Yn = dY.to_array(np.float64, q * ld).reshape(q, ld)
    assert rel(Yn[:, :n], 2.0 * (X[:, :n].T @ C).T) < 1e-13
    if p == q:
        check(lib().ed_xc_f64_dev(n, p, q, dX.addr, ld, dptr(C), 1.0, 0.0, dX.addr, ld, None))
        Xn = dX.to_array(np.float64, p * ld).reshape(p, ld)
        assert rel(Xn[:, :n], (X[:, :n].T @ C).T) < 1e-13
    dX.free(); dY.free()


def check_block(gpu, m, nu, nd, nev=8, nblock=8, tol_e=1e-10, reorth="partial"):
    s = gpu.Sector(m, nu, nd)
    o = oracle.OracleSector(m, nu, nd, nthreads=16)
    print(f"block Lanczos sector ({nu},{nd}) dim {s.dim} reorth {reorth}", flush=True)
    ev, vec = gpu.block_lanczos(s, nblock=nblock, neig=nev, nitermax=200, tolerance=1e-12, reorth=reorth)
    r = gpu.block_lanczos.last
    w_ref, _ = oracle.lowest_eigs(o, nev)
    print(f"  steps {r.steps} full passes {r.full_passes} nconv {r.nconv} max|dE| {np.max(np.abs(ev - w_ref)):.2e} "
          f"max resid {r.resid.max():.2e}", flush=True)
    if reorth == "full":
        assert r.full_passes == r.steps
    assert np.max(np.abs(ev - w_ref)) < tol_e
    assert r.resid.max() < 1e-6
    assert r.nconv == nev
    assert np.max(np.abs(vec @ vec.T - np.eye(nev))) < 1e-10
    # the device residual agrees with the oracle H.v
    for k in range(nev):
        assert np.linalg.norm(o.hmult(vec[k]) - ev[k] * vec[k]) < 1e-6
    s.close()
    o.close()
    return ev


@pytest.mark.parametrize("reorth", ["partial", "full"])
def test_block_lanczos_small_sector(gpu, reorth):
    m, (nu, nd) = golden("A")   # dim 400
    check_block(gpu, m, nu, nd, reorth=reorth)


@pytest.mark.parametrize("reorth", ["partial", "full"])
def test_block_lanczos_workload_b(gpu, reorth):
    m, (nu, nd) = golden("B")   # dim 63,504
    check_block(gpu, m, nu, nd, reorth=reorth)
    if reorth == "partial":   # the omega-recurrence skipped most full projections
        r = gpu.block_lanczos.last
        assert r.full_passes < r.steps


@pytest.mark.parametrize("reorth", ["partial", "full"])
def test_block_lanczos_finds_degenerate_pairs(gpu, reorth):
    """Two orbitals with identical baths: levels come in orbital-exchange pairs, which a
    single-vector Lanczos cannot resolve and a block of 8 does."""
    eps = np.array([-0.7, -0.7, 0.4, 0.4, 1.1, 1.1])
    v = np.array([0.35, 0.35, 0.5, 0.5, 0.25, 0.25])
    m = kanamori(2, 6, 3.0, 0.5, eps, v)   # Ns=8
    ev = check_block(gpu, m, 4, 4, reorth=reorth)
    w_all = np.sort(ev)
    assert np.min(np.diff(w_all)) < 1e-9   # at least one exactly degenerate pair found


def test_block_lanczos_window_copies_the_leading_vectors(gpu):
    """ed_block_lanczos_window (the dEmax selection of Block_Lanczos.f90:289-300): the same
    values as ed_block_lanczos, the vectors of the states within dEmax of the lowest (a
    leading run) copied, the other rows left as they were."""
    m, (nu, nd) = golden("B")
    s = gpu.Sector(m, nu, nd)
    ev, vec = gpu.block_lanczos(s, nblock=8, neig=8, nitermax=200, tolerance=1e-12)
    d = 0.5 * ((ev[2] - ev[0]) + (ev[3] - ev[0]))   # between the third and the fourth state
    print(f"window {d:.6f} over E - E0 = {ev - ev[0]}", flush=True)
    ev_w, vec_w = gpu.block_lanczos(s, nblock=8, neig=8, nitermax=200, tolerance=1e-12, dEmax=d)
    full = gpu.block_lanczos.last
    keep = int(np.sum(ev - ev[0] <= d))
    assert 0 < keep < 8 and len(ev_w) == keep and vec_w.shape == (keep, s.dim)
    assert np.max(np.abs(ev_w - ev[:keep])) < 1e-12
    assert np.max(np.abs(full.evals - ev)) < 1e-12
    for k in range(keep):   # the same vectors up to sign
        assert min(np.max(np.abs(vec_w[k] - vec[k])), np.max(np.abs(vec_w[k] + vec[k]))) < 1e-10
    assert not np.any(full.vectors[keep:])
    s.close()


def test_block_lanczos_ritz_threads_do_not_change_the_result(gpu, monkeypatch):
    """The Ritz tests run on worker threads beside the device steps (ED_BLOCK_RITZ_THREADS,
    4 by default) or on the driving thread (0).  The driver takes the scheduled tests in
    order either way, so the stopping step, the full projections, the values and the vectors
    are the same; only the number of device steps run past the stopping one differs."""
    m, (nu, nd) = golden("B")
    s = gpu.Sector(m, nu, nd)
    out = {}
    for th in ("0", "1", "4"):
        monkeypatch.setenv("ED_BLOCK_RITZ_THREADS", th)
        ev, vec = gpu.block_lanczos(s, nblock=8, neig=8, nitermax=200, tolerance=1e-12)
        r = gpu.block_lanczos.last
        out[th] = (ev.copy(), vec.copy(), r.steps, r.full_passes)
        print(f"Ritz threads {th}: steps {r.steps} full {r.full_passes} E0 {ev[0]:.15f}", flush=True)
    ev0, vec0, steps0, full0 = out["0"]
    for th in ("1", "4"):
        ev, vec, steps, full = out[th]
        assert steps == steps0 and full == full0
        assert np.max(np.abs(ev - ev0)) < 1e-13
        assert np.max(np.abs(vec - vec0)) < 1e-10
    s.close()


def test_solver_block_path(gpu):
    """get_eigen_energies with block_size=8 (solver.h:152: N=0, 1, Ns sectors keep plain
    Lanczos) against the oracle's eigenvalues, sector by sector."""
    from ed_solver_amd.solver import get_eigen_energies
    m, _ = golden("B")
    sectors = [(5, 5), (4, 5), (1, 5)]
    states = get_eigen_energies(m, sectors=sectors, neigen=8, block_size=8, nitermax=200)
    for sec in sectors:
        got = np.sort([st.energy for st in states if st.sector == sec])
        o = oracle.OracleSector(m, *sec, nthreads=16)
        ref, _ = oracle.lowest_eigs(o, 8)
        o.close()
        if sec == (1, 5):   # plain Lanczos: the ground state
            assert abs(got[0] - ref[0]) < 1e-10
        else:
            assert np.max(np.abs(got - ref[:len(got)])) < 1e-10
