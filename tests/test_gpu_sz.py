"""S_z (Nambu) sectors on the GPU against the oracle restatement (tables bit-exact, H.v)
and against the physical spectrum with pairing (dense ED, Lanczos, block Lanczos)."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

import oracle.sz as osz
from ed_solver_amd.aim import AIMParams

pytestmark = pytest.mark.gpu


def sc_model(nc, nb, U, seed, u_offdiag=None):
    r = np.random.default_rng(seed)
    ec = np.stack([np.eye(nc) * (-U / 2 + 0.05), np.eye(nc) * (-U / 2 - 0.05)])
    if nc > 1:
        ec[:, 0, 1] = ec[:, 1, 0] = 0.15
    eps = r.uniform(-1, 1, nb)
    eb = np.stack([np.diag(eps), np.diag(eps + 0.1)])
    if nb > 1:
        eb[:, 0, 1] = eb[:, 1, 0] = 0.2
    vbc = np.zeros((2, nb, nc))
    for k in range(nb):
        vbc[:, k, k % nc] = r.uniform(0.2, 0.6)
    u = np.full((nc, nc), u_offdiag if u_offdiag is not None else 0.0)
    np.fill_diagonal(u, U)
    m = AIMParams(nc, nb, ec, eb, vbc, u, np.zeros((nc, nc)), 0.0, 1e-10)
    return m, r.uniform(-0.4, 0.4, (nb, nb)), r.uniform(-0.3, 0.3, (2, nb, nc))


CASES = [(1, 3, 4, False, None), (1, 3, 3, True, None), (2, 2, 4, True, 1.2), (2, 3, 5, False, 1.0),
         (1, 5, 6, True, None)]


def sparse_sz(tab):
    d = len(tab["states"])
    ptr = np.concatenate([[0], np.cumsum(tab["noff"])])
    rows = np.repeat(np.arange(d), np.diff(ptr))
    H = sp.coo_matrix((tab["offdiag"], (rows, tab["rankoff"])), shape=(d, d)).tocsr()
    return H + sp.diags(tab["quart"] + tab["diag"])


@pytest.mark.parametrize("case", CASES)
def test_sz_tables_bit_exact_and_hmult(gpu, case):
    nc, nb, npart, pair_ib, u_off = case
    m, pb, pvbc = sc_model(nc, nb, 2.0, seed=20 + nb + npart, u_offdiag=u_off)
    print(f"S_z sector nc={nc} nb={nb} npart={npart} pairing_imp_to_bath={pair_ib}", flush=True)
    s = gpu.SzSector(m, npart, pb, pvbc, pair_ib)
    tab = osz.tab_sz(m, npart, pb, pvbc, pair_ib)
    assert s.dim == len(tab["states"]) == osz.sz_dim(m.ns, npart)
    t = s.tables(0)
    assert np.array_equal(t["states"], tab["states"])
    assert np.array_equal(t["noff"], tab["noff"])
    assert np.array_equal(t["rankoff"], tab["rankoff"])
    assert np.array_equal(t["diag"], tab["diag"])
    assert np.array_equal(t["offdiag"], tab["offdiag"])
    assert np.array_equal(s.quart(), tab["quart"])
    assert s.energy_shift == osz.energy_shift(m)
    v = np.random.default_rng(3).uniform(-1, 1, s.dim)
    w = s.hmult(v)
    ref = osz.hmult_sz(tab, v)
    assert np.max(np.abs(w - ref)) / np.max(np.abs(ref)) < 1e-13
    from hipmem import DevBuf
    x32 = DevBuf.from_array(v.astype(np.float32))
    y32 = DevBuf(4 * s.dim)
    s.hmult_dev(y32.addr, x32.addr, "f32")
    s.synchronize()
    w32 = y32.to_array(np.float32, s.dim).astype(np.float64)
    assert np.max(np.abs(w32 - ref)) / np.max(np.abs(ref)) < 1e-5
    # two vectors at a padded leading dimension (the block H.v of the block Lanczos)
    ld = s.dim + 5
    Xh = np.zeros(2 * ld)
    Xh[:s.dim] = v
    Xh[ld:ld + s.dim] = v[::-1]
    X = DevBuf.from_array(Xh)
    Y = DevBuf(8 * 2 * ld)
    s.hmult_block_dev(Y.addr, X.addr, 2, ld)
    s.synchronize()
    Yh = Y.to_array(np.float64, 2 * ld)
    assert np.max(np.abs(Yh[:s.dim] - ref)) / np.max(np.abs(ref)) < 1e-13
    ref2 = osz.hmult_sz(tab, v[::-1].copy())
    assert np.max(np.abs(Yh[ld:ld + s.dim] - ref2)) / np.max(np.abs(ref2)) < 1e-13
    for buf in (x32, y32, X, Y):
        buf.free()
    s.close()


@pytest.mark.parametrize("case", CASES[:4])
def test_sz_full_ed_matches_physical_spectrum(gpu, case):
    nc, nb, npart, pair_ib, u_off = case
    m, pb, pvbc = sc_model(nc, nb, 2.5, seed=40 + npart, u_offdiag=u_off)
    s = gpu.SzSector(m, npart, pb, pvbc, pair_ib)
    if s.dim > 1024:
        pytest.skip("dense ED limit")
    ev, vec = s.full_ed(s.dim)
    e_p = osz.physical_spectrum(m, npart, pb, pvbc, pair_ib)
    assert np.max(np.abs(ev + s.energy_shift - e_p)) < 1e-11
    s.close()


def test_sz_lanczos_and_block_lanczos(gpu):
    """Ns = 8 (2 x 8 Nambu bits), npart = 8: dimension 12,870."""
    from ed_solver_amd.block import block_lanczos
    m, pb, pvbc = sc_model(2, 6, 2.0, seed=7, u_offdiag=1.2)
    s = gpu.SzSector(m, 8, pb, pvbc, True)
    print(f"S_z sector dim {s.dim}", flush=True)
    tab = osz.tab_sz(m, 8, pb, pvbc, True)
    H = sparse_sz(tab)
    e_ref = np.sort(sla.eigsh(H, k=4, which="SA", tol=1e-12)[0])
    r = s.lanczos(nitermax=300, neigen=1, tolerance=1e-12)
    assert abs(r.evals[0] - e_ref[0]) < 1e-10
    psi = s.ground_state(r, 0)
    assert np.linalg.norm(H @ psi - e_ref[0] * psi) < 1e-5
    for prec in ("f64", "f32"):
        a, b = s.lanczos_run(30, precision=prec)
        assert np.all(np.isfinite(a)) and np.all(b[1:] > 0)
    ev, vec = block_lanczos(s, nblock=4, neig=4, nitermax=200, tolerance=1e-12)
    assert np.max(np.abs(np.sort(ev)[:4] - e_ref)) < 1e-9
    s.close()


def with_hund(m, J):
    jj = np.full((m.nc, m.nc), J)
    np.fill_diagonal(jj, 0.0)
    return AIMParams(m.nc, m.nb, m.ec, m.eb, m.vbc, m.u, jj, J, m.cutoff)


@pytest.mark.parametrize("case", [(2, 2, 4, True), (3, 2, 5, True), (2, 3, 5, False), (3, 3, 6, True)])
def test_sz_hund_tables_hmult_and_spectrum(gpu, case):
    """S_z sectors with Hund coupling (the terms of HAIMsz_multc): QUART_INT_SZ and the Hund
    ELL bit-exact against the oracle, H.v to 1e-12, and the dense spectrum + shift against
    the physical spectrum of the Kanamori impurity with pairing."""
    nc, nb, npart, pair_ib = case
    m0, pb, pvbc = sc_model(nc, nb, 3.0, seed=30 + nc + nb, u_offdiag=1.6)
    m = with_hund(m0, 0.6)
    s = gpu.SzSector(m, npart, pb, pvbc, pair_ib)
    print(f"S_z with J {case}: dim {s.dim}", flush=True)
    tab = osz.tab_sz(m, npart, pb, pvbc, pair_ib)
    assert np.array_equal(s.quart(), tab["quart"])
    hund = osz.hund_sz(m, npart)
    rk, val = s.hund()
    rows = [[] for _ in range(s.dim)]
    for i, j, h in hund:
        rows[i].append((j, h))
    assert rk.shape[0] == max(len(r) for r in rows)
    for i, r in enumerate(rows):
        for k in range(rk.shape[0]):
            want = r[k] if k < len(r) else (i, 0.0)
            assert rk[k, i] == want[0] and val[k, i] == want[1], (i, k)
    v = np.random.default_rng(3).uniform(-1, 1, s.dim)
    ref = osz.hmult_sz(tab, v, hund)
    assert np.max(np.abs(s.hmult(v) - ref)) < 1e-12 * np.max(np.abs(ref))
    if s.dim <= 1024:
        ev, _ = s.full_ed(min(s.dim, 8))
        ep = np.sort(osz.physical_spectrum(m, npart, pb, pvbc if pair_ib else None, pair_ib))
        assert np.max(np.abs(np.sort(ev) + s.energy_shift - ep[:len(ev)])) < 1e-10
    r = s.lanczos(nitermax=300, neigen=1, tolerance=1e-12)
    e_ref = np.sort(sla.eigsh(sparse_sz(tab) + sparse_hund(hund, s.dim), k=2, which="SA", tol=1e-12)[0])[0] \
        if s.dim > 64 else np.linalg.eigvalsh(osz.dense_sz(tab, hund))[0]
    assert abs(r.evals[0] - e_ref) < 1e-10
    s.close()


def sparse_hund(hund, d):
    if not hund:
        return sp.csr_matrix((d, d))
    i, j, h = zip(*hund)
    return sp.coo_matrix((h, (i, j)), shape=(d, d)).tocsr()


def test_sz_rejections(gpu):
    from ed_solver_amd import EDError, apply_C
    m, pb, pvbc = sc_model(2, 2, 2.0, seed=1, u_offdiag=1.0)
    s = gpu.SzSector(m, 4, pb)
    s2 = gpu.SzSector(m, 5, pb)
    with pytest.raises(EDError):   # npart 4 -> 5 is c+, not c
        apply_C(s, s2, 0, 0, -1, np.ones(s.dim))
    with pytest.raises(EDError):   # one S_z and one (N_up, N_dn) sector
        apply_C(s, gpu.Sector(m, 2, 2), 0, 0, +1, np.ones(s.dim))
    with pytest.raises(EDError):
        s.set_schedule(8)
    s.close(); s2.close()


def test_sz_sector_above_two_million_matches_updo_sectors(gpu):
    """Ns = 12 without pairing, npart = 12: dimension C(24, 12) = 2,704,156, above the
    8192 x 256 rows of one H.v launch's partial sums (the kernel strides over the rest).
    Without pairing the block N_up - N_dn = 0 is the sum of the (k, k) sectors, so
    E0 + energy_shift = min_k E0(k, k)."""
    m, _, _ = sc_model(3, 9, 2.0, seed=21, u_offdiag=1.0)
    s = gpu.SzSector(m, 12)
    assert s.dim == 2704156
    print(f"S_z sector dim {s.dim}", flush=True)
    r = s.lanczos(nitermax=400, neigen=1, tolerance=1e-12)
    e_sz = r.evals[0] + s.energy_shift
    s.close()
    e_updo = []
    for k in range(13):
        sk = gpu.Sector(m, k, k)
        if sk.dim == 1:
            e_updo.append(float(sk.full_ed(1)[0][0]))
        else:
            e_updo.append(float(sk.lanczos(nitermax=400, neigen=1, tolerance=1e-12).evals[0]))
        sk.close()
    assert abs(e_sz - min(e_updo)) < 1e-9, (e_sz, min(e_updo))


@pytest.mark.parametrize("case", [(2, 3, 5, True, 0.6), (3, 3, 6, True, 0.0), (2, 6, 8, False, 0.5)])
def test_sz_on_the_fly_matches_table(gpu, case):
    """HAIMsz_mult_fly: the entries generated in the H.v give the table path's bits (H.v in
    FP64, fixed-length Lanczos runs in FP64 and FP32)."""
    nc, nb, npart, pair_ib, J = case
    m0, pb, pvbc = sc_model(nc, nb, 3.0, seed=40 + nb, u_offdiag=1.6)
    m = with_hund(m0, J) if J else m0
    s = gpu.SzSector(m, npart, pb, pvbc, pair_ib)
    print(f"S_z fly {case}: dim {s.dim}", flush=True)
    v = np.random.default_rng(8).uniform(-1, 1, s.dim)
    w0 = s.hmult(v)
    a0 = [s.lanczos_run(40, precision=p) for p in ("f64", "f32")]
    s.set_fly(True)
    assert np.array_equal(s.hmult(v), w0)
    a1 = [s.lanczos_run(40, precision=p) for p in ("f64", "f32")]
    for (x0, y0), (x1, y1) in zip(a0, a1):
        assert np.array_equal(x0, x1) and np.array_equal(y0, y1)
    s.set_fly(False)
    assert np.array_equal(s.hmult(v), w0)
    s.close()


def test_sz_on_the_fly_large_sector_rate(gpu):
    """The 2.7M-state S_z sector of Ns = 12: table and on-the-fly H.v agree bitwise; both
    times printed."""
    m, _, _ = sc_model(3, 9, 2.0, seed=21, u_offdiag=1.0)
    s = gpu.SzSector(m, 12)
    v = np.random.default_rng(9).uniform(-1, 1, s.dim)
    w0 = s.hmult(v)
    t0, _ = s.hmult_benchmark("f64", 20)
    s.set_fly(True)
    assert np.array_equal(s.hmult(v), w0)
    t1, _ = s.hmult_benchmark("f64", 20)
    print(f"S_z dim {s.dim}: H.v table {t0 * 1e3:.1f} us, on the fly {t1 * 1e3:.1f} us", flush=True)
    s.close()


def with_hund(m, jh):
    jj = np.full((m.nc, m.nc), jh)
    np.fill_diagonal(jj, 0.0)
    return AIMParams(m.nc, m.nb, m.ec, m.eb, m.vbc, m.u, jj, jh, m.cutoff)


@pytest.mark.parametrize("case", [(1, 3, 4, False, None), (2, 2, 4, True, 1.2), (2, 3, 5, False, 1.0)])
def test_sz_apply_c_bit_exact(gpu, case):
    """c+ / c of every Nambu impurity orbital (apply_Csz) against the oracle, bit for bit."""
    nc, nb, npart, pair, uoff = case
    m, pb, pvbc = sc_model(nc, nb, 2.0, seed=nc + nb, u_offdiag=uoff)
    src = gpu.SzSector(m, npart, pb, pvbc, pair)
    r = np.random.default_rng(npart)
    psi = r.standard_normal(src.dim)
    for pm in (+1, -1):
        dst = gpu.SzSector(m, npart + pm, pb, pvbc, pair)
        for site in range(nc):
            for spin in (0, 1):
                got = gpu.apply_C(src, dst, site, spin, pm, psi)
                ref = osz.apply_c_sz(m, npart, site, spin, pm, psi)
                assert np.array_equal(got, ref), (site, spin, pm)
        dst.close()
    src.close()


@pytest.mark.parametrize("case", [(1, 2, 0.0, False), (2, 2, 0.0, True), (2, 2, 0.4, False)])
def test_sz_nambu_green_vs_physical(gpu, case):
    """The solver in the S_z basis: every npart sector's states (dense ED, all kept), then
    the Nambu G^N (normal up, <<c_dn+ ; c_dn>>, anomalous <<c_up ; c_dn>>) against the
    Lehmann sum of the physical Hamiltonian with pairing over the whole Fock space,
    through the Lehmann path and through excitation Lanczos (full_ed_switch = 0, as many
    steps as the largest sector dimension)."""
    from ed_solver_amd.solver import compute_green_nambu, get_eigen_energies_sz
    nc, nb, jh, pair = case
    m, pb, pvbc = sc_model(nc, nb, 2.0, seed=5 + nc + nb, u_offdiag=1.0 if nc > 1 else None)
    if jh:
        m = with_hund(m, jh)
    beta = 6.0
    wn = np.pi / beta * (2 * np.arange(1, 17) - 1)
    dmax = max(osz.sz_dim(m.ns, n) for n in range(2 * m.ns + 1))
    states = get_eigen_energies_sz(m, pb, pvbc, pair, full_ed_switch=dmax, full_ed_green=True)
    assert len(states) == 4 ** m.ns
    print(f"Nambu G case {case}: {len(states)} states, largest sector {dmax}", flush=True)
    G = compute_green_nambu(m, states, beta, wn, pb, pvbc, pair, full_ed_switch=dmax)
    G_cf = compute_green_nambu(m, states, beta, wn, pb, pvbc, pair, nitergreen=dmax, full_ed_switch=0)
    ref = osz.nambu_green_physical(m, beta, wn, pb=pb, pvbc=pvbc, pairing_imp_to_bath=pair)
    assert np.max(np.abs(ref[:nc, nc:])) > 1e-3   # the anomalous part is there
    assert np.max(np.abs(G - ref)) < 1e-10
    assert np.max(np.abs(G_cf - ref)) < 1e-9


def test_sz_apply_c_n_vs_restatement(gpu):
    """ed_apply_c_n (apply_Csz_N_updo) on every Nambu impurity orbital, both signs, against
    the restatement oracle/sz.py apply_c_n_sz: bit-exact; (N_up, N_dn) sectors are refused
    as the reference STOPs."""
    m, pb, pvbc = sc_model(2, 2, 2.0, seed=11, u_offdiag=1.0)
    npart = m.ns
    src = gpu.SzSector(m, npart, pb, pvbc, True)
    res = src.lanczos(nitermax=60, neigen=1, tolerance=1e-12)
    psi = src.ground_state(res)
    for pm in (+1, -1):
        dst = gpu.SzSector(m, npart + pm, pb, pvbc, True)
        for spin in (0, 1):
            for site in range(m.nc):
                got = gpu.apply_C_n(src, dst, site, spin, pm, psi)
                ref = osz.apply_c_n_sz(m, npart, site, spin, pm, psi)
                assert np.array_equal(got, ref), (site, spin, pm)
                assert np.any(got != 0) and not np.array_equal(got, gpu.apply_C(src, dst, site, spin, pm, psi))
        dst.close()
    src.close()
    mu, (nu, nd) = gpu.synthetic("A")
    a, b = gpu.Sector(mu, nu, nd), gpu.Sector(mu, nu + 1, nd)
    with pytest.raises(gpu.EDError):
        gpu.apply_C_n(a, b, 0, 0, +1, np.ones(a.dim))
    a.close(); b.close()


@pytest.mark.parametrize("case", [(1, 2, 0.0, True), (2, 2, 0.4, False)])
def test_sz_ncup_green_vs_physical(gpu, case):
    """FLAG_NCUP Green's functions of the S_z solver (compute_green_ncup): GN = <<D ; D+>>
    with D = n_dn c_up, n_up c+_dn and the cross GNF = <<D ; d+>> against the Lehmann sums
    of those physical operator strings over the whole Fock space (oracle/sz.py
    ncup_green_physical), through the Lehmann path and through excitation Lanczos (the
    correlated terms built on the device with ED_SITE_NCOND)."""
    from ed_solver_amd.solver import compute_green_ncup, get_eigen_energies_sz
    nc, nb, jh, pair = case
    m, pb, pvbc = sc_model(nc, nb, 2.0, seed=21 + nc + nb, u_offdiag=1.0 if nc > 1 else None)
    if jh:
        m = with_hund(m, jh)
    beta = 6.0
    wn = np.pi / beta * (2 * np.arange(1, 17) - 1)
    dmax = max(osz.sz_dim(m.ns, n) for n in range(2 * m.ns + 1))
    states = get_eigen_energies_sz(m, pb, pvbc, pair, full_ed_switch=dmax, full_ed_green=True)
    print(f"NCUP G case {case}: {len(states)} states, largest sector {dmax}", flush=True)
    gn_ref, gnf_ref = osz.ncup_green_physical(m, beta, wn, pb=pb, pvbc=pvbc, pairing_imp_to_bath=pair)
    G_ref = osz.nambu_green_physical(m, beta, wn, pb=pb, pvbc=pvbc, pairing_imp_to_bath=pair)
    for kw in (dict(full_ed_switch=dmax), dict(nitergreen=dmax, full_ed_switch=0)):
        GN, GNF, G = compute_green_ncup(m, states, beta, wn, pb, pvbc, pair, **kw)
        tol = 1e-10 if kw["full_ed_switch"] else 1e-9
        assert np.max(np.abs(GN - gn_ref)) < tol, kw
        assert np.max(np.abs(GNF - gnf_ref)) < tol, kw
        assert np.array_equal(GNF[0], GNF[1])   # real H: one symmetric combination
        assert np.max(np.abs(G - np.array([G_ref[a, a] for a in range(2 * nc)]))) < tol, kw



def complex_sc_model(nc, nb, seed, jhund=0.0):
    """Complex Ec, Eb, Vbc (oracle.cplx.random_model) with complex Pb and PVbc."""
    import oracle.cplx as oc
    m, im = oc.random_model(nc, nb, seed, jhund=jhund, u_off=1.1 if nc > 1 else 1.0)
    r = np.random.default_rng(seed + 100)
    pb = r.uniform(-0.4, 0.4, (nb, nb)) + 1j * r.uniform(-0.4, 0.4, (nb, nb))
    pvbc = r.uniform(-0.3, 0.3, (2, nb, nc)) + 1j * r.uniform(-0.3, 0.3, (2, nb, nc))
    return m, im, pb, pvbc


@pytest.mark.parametrize("case", [(1, 3, 4, True, 0.0), (2, 2, 4, False, 0.4), (2, 3, 5, True, 0.3)])
def test_complex_sz_tables_hmult_spectrum(gpu, case):
    """Complex S_z sector (ed_sector_create_sz_c, HAIMsz_multc): the complex tables (real and
    imaginary parts) bit for bit against oracle/sz.py tab_sz with im; QUART and the Hund
    ELL as the restatement; ed_hmult_c128 against the restated dense H (conj(h) gather)
    <= 1e-12 relative; the on-the-fly H.v (HAIMsz_mult_fly_c) bitwise equal to the table;
    dense ED + shift against the physical spectrum of the complex Hamiltonian with complex
    pairing <= 1e-10; Lanczos E0 against eigvalsh of the restated H <= 1e-10."""
    nc, nb, npart, pair, jh = case
    m, im, pb, pvbc = complex_sc_model(nc, nb, seed=7 + nb + npart, jhund=jh)
    s = gpu.ComplexSzSector(m, npart, im, pb, pvbc, pair)
    tab = osz.tab_sz(m, npart, pb, pvbc, pair, im=im)
    print(f"complex S_z case {case}: dim {s.dim}, nnz {len(tab['rankoff'])}", flush=True)
    t = s.tables(0)
    assert np.array_equal(t["states"], tab["states"]) and np.array_equal(t["noff"], tab["noff"])
    assert np.array_equal(t["rankoff"], tab["rankoff"])
    assert np.array_equal(t["offdiag"], tab["offdiag"].real)
    assert np.array_equal(s.tables_im(0), tab["offdiag"].imag)
    assert np.array_equal(t["diag"], tab["diag"])
    assert np.array_equal(s.quart(), tab["quart"])
    hund = osz.hund_sz(m, npart)
    H = osz.dense_sz(tab, hund)
    v = np.random.default_rng(npart).standard_normal(s.dim) + 1j * np.random.default_rng(npart + 1).standard_normal(s.dim)
    w = s.hmult(v)
    assert np.max(np.abs(w - H @ v)) < 1e-12 * np.max(np.abs(H @ v))
    s.set_fly(True)
    assert np.array_equal(s.hmult(v), w)
    s.set_fly(False)
    ev, vec = s.full_ed(min(6, s.dim))
    e_phys = osz.physical_spectrum(m, npart, pb, pvbc, pair, im=im)
    assert np.max(np.abs(ev + s.energy_shift - e_phys[:len(ev)])) < 1e-10
    for k in range(len(ev)):
        assert np.linalg.norm(s.hmult(vec[k]) - ev[k] * vec[k]) < 1e-10
    r = s.lanczos(nitermax=min(s.dim - 1, 200), neigen=1, tolerance=1e-13)
    w_ref = np.linalg.eigvalsh(H)
    assert abs(r.evals[0] - w_ref[0]) < 1e-10
    psi = s.ground_state(r)
    assert np.linalg.norm(H @ psi - w_ref[0] * psi) < 1e-6
    s.close()


def test_complex_sz_apply_c_and_green(gpu):
    """Complex S_z solver: c / c+ and the correlated excitation of every Nambu orbital on
    planar vectors bit for bit against the restatement; the Nambu G^N (G_ij != G_ji) and the
    FLAG_NCUP GN, GNF (both orders) against the physical Lehmann sums of the complex
    Hamiltonian with complex pairing, through the Lehmann path and excitation Lanczos."""
    from ed_solver_amd.solver import compute_green_nambu, compute_green_ncup, get_eigen_energies_sz
    m, im, pb, pvbc = complex_sc_model(2, 1, seed=41, jhund=0.3)
    pair = True
    npart = m.ns
    src = gpu.ComplexSzSector(m, npart, im, pb, pvbc, pair)
    r = src.lanczos(nitermax=src.dim - 1, neigen=1, tolerance=1e-13)
    psi = src.ground_state(r)
    for pm in (+1, -1):
        dst = gpu.ComplexSzSector(m, npart + pm, im, pb, pvbc, pair)
        for spin in (0, 1):
            for site in range(m.nc):
                assert np.array_equal(gpu.apply_C(src, dst, site, spin, pm, psi),
                                      osz.apply_c_sz(m, npart, site, spin, pm, psi))
                assert np.array_equal(gpu.apply_C_n(src, dst, site, spin, pm, psi),
                                      osz.apply_c_n_sz(m, npart, site, spin, pm, psi))
        dst.close()
    src.close()
    beta = 5.0
    wn = np.pi / beta * (2 * np.arange(1, 13) - 1)
    dmax = max(osz.sz_dim(m.ns, n) for n in range(2 * m.ns + 1))
    states = get_eigen_energies_sz(m, pb, pvbc, pair, full_ed_switch=dmax, full_ed_green=True, im=im)
    assert len(states) == 4 ** m.ns
    G_ref = osz.nambu_green_physical(m, beta, wn, pb=pb, pvbc=pvbc, pairing_imp_to_bath=pair, im=im)
    gn_ref, gnf_ref = osz.ncup_green_physical(m, beta, wn, pb=pb, pvbc=pvbc, pairing_imp_to_bath=pair, im=im)
    assert np.max(np.abs(G_ref[0, 1] - G_ref[1, 0])) > 1e-4
    for kw in (dict(full_ed_switch=dmax), dict(nitergreen=dmax, full_ed_switch=0)):
        tol = 1e-10 if kw["full_ed_switch"] else 1e-9
        G = compute_green_nambu(m, states, beta, wn, pb, pvbc, pair, im=im, **kw)
        assert np.max(np.abs(G - G_ref)) < tol, kw
        GN, GNF, Gd = compute_green_ncup(m, states, beta, wn, pb, pvbc, pair, im=im, **kw)
        assert np.max(np.abs(GN - gn_ref)) < tol, kw
        assert np.max(np.abs(GNF - gnf_ref)) < tol, kw
