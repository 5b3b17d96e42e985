"""Complex sectors (Hmultc) on the GPU: tables bit-exact against the oracle restatement,
H.v against the brute-force complex Hamiltonian, and at the sizes of workloads B and C
the Lanczos ground state and G(i w_n) of a phase-transformed real model against the
real sector (same spectrum, same impurity Green's function)."""
import json
import os

import numpy as np
import pytest

import oracle.brute as bf
import oracle.cplx as oc
from ed_solver_amd.aim import AIMParams

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "models.json")
CASES = [(2, 2, 2, 2, 0.4), (1, 3, 2, 1, 0.0), (2, 3, 3, 2, 0.3), (2, 4, 3, 3, 0.0)]


def golden(name):
    with open(GOLD) as f:
        d = json.load(f)[name]
    return AIMParams.from_json(d["model"]), tuple(d["sector"])


def csector(gpu, m, im, nup, ndn):
    return gpu.ComplexSector(m, nup, ndn, im["ec"], im["eb"], im["vbc"])


@pytest.mark.parametrize("case", CASES)
def test_cplx_tables_bit_exact_and_hmult(gpu, case):
    nc, nb, nup, ndn, jh = case
    m, im = oc.random_model(nc, nb, seed=nc * 10 + nb, jhund=jh)
    s = csector(gpu, m, im, nup, ndn)
    print(f"complex case {case}: dim {s.dim}", flush=True)
    for spin, n in ((0, nup), (1, ndn)):
        t, ti, ref = s.tables(spin), s.tables_im(spin), oc.tab_c(m, im, spin, n)
        assert np.array_equal(t["states"], ref["states"]) and np.array_equal(t["noff"], ref["noff"])
        assert np.array_equal(t["diag"], ref["diag"]) and np.array_equal(t["rankoff"], ref["rankoff"])
        assert np.array_equal(t["offdiag"], ref["offdiag"].real) and np.array_equal(ti, ref["offdiag"].imag)
    H = oc.sector_hamiltonian_c(m, im, nup, ndn)
    r = np.random.default_rng(1)
    for _ in range(2):
        v = r.standard_normal(s.dim) + 1j * r.standard_normal(s.dim)
        w = s.hmult(v)
        ref = H @ v
        assert np.max(np.abs(w - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))
        # the planar form of the real-typed entry points: Re <v, H v> / |v|^2
        assert abs(s.rayleigh(v) - np.vdot(v, ref).real / np.vdot(v, v).real) < 1e-12
    s.close()


def test_cplx_hmult_dev(gpu):
    """Device COMPLEX(8) vectors (interleaved), out of place and in place."""
    from hipmem import DevBuf
    m, im = oc.random_model(2, 3, seed=4, jhund=0.3)
    s = csector(gpu, m, im, 3, 2)
    r = np.random.default_rng(2)
    v = r.standard_normal(s.dim) + 1j * r.standard_normal(s.dim)
    x, y = DevBuf.from_array(v), DevBuf(16 * s.dim)
    s.hmult_c128_dev(y.addr, x.addr)
    s.synchronize()
    ref = s.hmult(v)
    assert np.max(np.abs(y.to_array(np.complex128, s.dim) - ref)) < 1e-13
    s.hmult_c128_dev(x.addr, x.addr)
    s.synchronize()
    assert np.max(np.abs(x.to_array(np.complex128, s.dim) - ref)) < 1e-13
    x.free(); y.free()
    s.close()


def test_cplx_lanczos_small_vs_dense(gpu):
    m, im = oc.random_model(2, 4, seed=11, jhund=0.0)
    s = csector(gpu, m, im, 3, 3)
    H = oc.sector_hamiltonian_c(m, im, 3, 3).toarray()
    e = np.linalg.eigvalsh(H)
    res = s.lanczos(nitermax=200, neigen=2, tolerance=1e-13)
    assert res.precision == "f64"
    assert np.max(np.abs(res.evals - e[:2])) < 1e-10
    psi = s.ground_state(res)
    assert psi.dtype == np.complex128 and abs(np.linalg.norm(psi) - 1.0) < 1e-12
    assert np.linalg.norm(H @ psi - e[0] * psi) < 1e-6
    s.close()


def state_phases(sec, theta):
    """exp(-i sum_{a occupied} theta[s, a]) per basis state iup + ido d_up: the diagonal
    unitary W with H' = W^+ H W for the model of oracle.cplx.phase_model."""
    su, sd = sec.tables(0)["states"], sec.tables(1)["states"]
    bits = lambda st, th: np.array([sum(th[a] for a in range(len(th)) if (int(x) >> a) & 1) for x in st])
    ph_up, ph_dn = bits(su, theta[0]), bits(sd, theta[1])
    return np.exp(-1j * (ph_up[:, None] + ph_dn[None, :])).T.reshape(-1)


@pytest.mark.parametrize("name", ["B", "C"])
def test_cplx_phase_transformed_workload_matches_real(gpu, name):
    """A phase on every bath orbital (impurity phases 0, so the pair hopping of J stays
    invariant): H' = W^+ H W with W diagonal in the Fock basis.  The complex sector's own
    Lanczos E0 equals the real one; the real ground state carried over by W is an
    eigenvector of the complex H to the same residual; and c / c+ of the impurity orbital
    and G(i w_n) from it equal the real ones up to rounding."""
    m, (nu, nd) = golden(name)
    theta = np.random.default_rng(3).uniform(0, 2 * np.pi, (2, m.ns))
    theta[:, m.nb:] = 0.0
    m2, im2 = oc.phase_model(m, theta)
    sr = gpu.Sector(m, nu, nd)
    sc = csector(gpu, m2, im2, nu, nd)
    print(f"{name}: dim {sr.dim}", flush=True)
    rr = sr.lanczos(nitermax=300, neigen=1, tolerance=1e-13)
    rc = sc.lanczos(nitermax=300, neigen=1, tolerance=1e-13)
    assert abs(rr.evals[0] - rc.evals[0]) < 1e-10
    e0 = rr.evals[0]
    psr = sr.ground_state(rr)
    w = state_phases(sr, theta)
    psc = w * psr
    res_r = np.linalg.norm(sr.hmult(psr) - e0 * psr)
    res_c = np.linalg.norm(sc.hmult(psc) - e0 * psc)
    assert abs(res_r - res_c) < 1e-10 * max(1.0, res_r), (res_r, res_c)
    wn = gpu.matsubara(100.0, 64)
    for pm, dst in ((+1, (nu + 1, nd)), (-1, (nu - 1, nd))):
        dr, dc = gpu.Sector(m, *dst), csector(gpu, m2, im2, *dst)
        phr = gpu.apply_C(sr, dr, 0, 0, pm, psr)
        phc = gpu.apply_C(sc, dc, 0, 0, pm, psc)
        assert np.max(np.abs(phc - state_phases(dr, theta) * phr)) < 1e-14
        iph = 1 if pm > 0 else 2
        gr, ar, br = gpu.compute_dynamic(dr, iph, e0, phr, 36, wn)
        gc, ac, bc = gpu.compute_dynamic(dc, iph, e0, phc, 36, wn)
        assert np.max(np.abs(ar - ac)) < 1e-9 and np.max(np.abs(br - bc)) < 1e-9
        assert np.max(np.abs(gr - gc)) < 1e-10 * np.max(np.abs(gr))
        dr.close(); dc.close()
    sr.close(); sc.close()


def test_cplx_rejections(gpu):
    from ed_solver_amd import EDError
    from ed_solver_amd.block import block_lanczos
    m, im = oc.random_model(2, 2, seed=1, jhund=0.4)
    s = csector(gpu, m, im, 2, 2)
    sr = gpu.Sector(m, 3, 2)
    with pytest.raises(EDError):
        s.lanczos(precision="f32")
    with pytest.raises(EDError):
        block_lanczos(s, nblock=2, neig=2)
    with pytest.raises(EDError):
        s.set_schedule(8)
    with pytest.raises(EDError):   # one complex and one real sector
        gpu.apply_C(s, sr, 0, 0, +1, np.ones(s.dim, complex))
    s.close(); sr.close()


def test_cplx_full_ed_vs_dense(gpu):
    """Hermitian dense ED (from the real embedding): eigenvalues against eigvalsh of the
    brute-force complex H, residuals, and orthonormal complex eigenvectors, also inside
    the degenerate multiplets of a phase-transformed model of two identical impurity +
    bath pairs."""
    m, im = oc.random_model(2, 3, seed=21, jhund=0.3)
    # two identical impurity + bath pairs without inter-orbital terms: exact degeneracies,
    # kept by a phase on every orbital (J = 0)
    diag2 = lambda a, b: np.stack([np.diag([a, b])] * 2)
    mdeg = AIMParams(2, 2, diag2(-1.0, -1.0), diag2(0.3, 0.3), diag2(0.5, 0.5), np.diag([2.5, 2.5]),
                     np.zeros((2, 2)), 0.0, 1e-8)
    theta = np.random.default_rng(8).uniform(0, 2 * np.pi, (2, mdeg.ns))
    m2, im2 = oc.phase_model(mdeg, theta)
    for (mm, ii, sec) in ((m, im, (2, 2)), (m, im, (3, 2)), (m2, im2, (2, 2)), (m2, im2, (3, 2))):
        s = csector(gpu, mm, ii, *sec)
        H = oc.sector_hamiltonian_c(mm, ii, *sec).toarray()
        e_ref = np.linalg.eigvalsh(H)
        ev, vec = s.full_ed(s.dim)
        print(f"complex full ED {sec}: dim {s.dim}, min gap {np.min(np.diff(e_ref)):.2e}", flush=True)
        assert vec.dtype == np.complex128 and vec.shape == (s.dim, s.dim)
        assert np.max(np.abs(ev - e_ref)) < 1e-12
        assert np.max(np.abs(H @ vec.T - vec.T * ev)) < 1e-11
        assert np.max(np.abs(vec.conj() @ vec.T - np.eye(s.dim))) < 1e-11
        ev3, vec3 = s.full_ed(3)
        assert np.max(np.abs(ev3 - e_ref[:3])) < 1e-12
        s.close()
    assert np.min(np.diff(np.linalg.eigvalsh(oc.sector_hamiltonian_c(m2, im2, 2, 2).toarray()))) < 1e-10


def test_cplx_green_matrix_vs_lehmann(gpu):
    """The solver on a complex model (im): every sector by dense ED, G_ij = <<c_i; c_j+>>
    from the (c_i + c_j) and (c_i - i c_j) runs against the exact Lehmann sum over the
    full Fock space (G_ij != G_ji), through the Lehmann path and through excitation
    Lanczos (full_ed_switch = 0, 36 steps >= every sector dimension)."""
    from ed_solver_amd.solver import compute_green, compute_green_matrix, get_eigen_energies
    m, im = oc.random_model(2, 2, seed=31, jhund=0.0)
    beta = 8.0
    wn = np.pi / beta * (2 * np.arange(1, 25) - 1)
    states = get_eigen_energies(m, neigen=36, im=im)
    assert len(states) == 4 ** 4
    G = compute_green_matrix(m, states, beta, wn, nitergreen=36, im=im)
    G_cf = compute_green_matrix(m, states, beta, wn, nitergreen=36, full_ed_switch=0, im=im)
    g_diag = compute_green(m, states, beta, wn, nitergreen=36, im=im)
    ns, nc = m.ns, m.nc
    spec = {(nu, nd): np.linalg.eigh(oc.sector_hamiltonian_c(m, im, nu, nd).toarray())
            for nu in range(ns + 1) for nd in range(ns + 1)}
    e0 = min(v[0].min() for v in spec.values())
    Z = sum(np.exp(-beta * (v[0] - e0)).sum() for v in spec.values())
    Gref = np.zeros((nc, nc, len(wn)), complex)
    for (nu, nd), (em, vm) in spec.items():
        if nu == ns:
            continue
        en, vn = spec[(nu + 1, nd)]
        for a in range(len(em)):
            amp = [vn.conj().T @ bf.apply_c_full(m, nu, nd, c, 0, +1, vm[:, a]) for c in range(nc)]
            wgt = np.exp(-beta * (em[a] - e0)) + np.exp(-beta * (en - e0))
            den = 1j * wn[None, :] - (en - em[a])[:, None]
            for i in range(nc):
                for j in range(nc):
                    Gref[i, j] += ((np.conj(amp[i]) * amp[j] * wgt)[:, None] / den).sum(0)
    Gref /= Z
    assert np.max(np.abs(Gref[0, 1] - Gref[1, 0])) > 1e-3   # G_ij != G_ji for complex H
    assert np.max(np.abs(G - Gref)) < 1e-9
    assert np.max(np.abs(G_cf - Gref)) < 1e-9
    assert np.max(np.abs(g_diag - np.array([Gref[c, c] for c in range(nc)]))) < 1e-9
    with pytest.raises(ValueError):
        get_eigen_energies(m, block_size=4, im=im)
