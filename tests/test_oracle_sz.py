"""The S_z (Nambu) restatement of the oracle against the physical Hamiltonian with pairing.

The particle-hole transformation is unitary: eigenvalues of the tabulated Nambu sector
plus tr(Ec_dn) + tr(Eb_dn) must equal the physical spectrum of the block
N_up - N_dn = npart - Ns.  A dropped pairing term, a wrong sign on a Nambu block, a
transposed matrix or a missing shift changes the spectrum and fails here.
"""
import numpy as np
import pytest

import oracle.sz as osz
from ed_solver_amd.aim import AIMParams, kanamori
import oracle.brute as bf


def sc_model(nc, nb, U, seed, u_offdiag=None, spin_split=0.0):
    r = np.random.default_rng(seed)
    ec = np.stack([np.eye(nc) * (-U / 2 + spin_split), np.eye(nc) * (-U / 2 - spin_split)])
    if nc > 1:
        ec[:, 0, 1] = ec[:, 1, 0] = 0.15
    eps = r.uniform(-1, 1, nb)
    eb = np.stack([np.diag(eps), np.diag(eps + 0.1)])
    eb[:, 0, 1] = eb[:, 1, 0] = 0.2
    vbc = np.zeros((2, nb, nc))
    for k in range(nb):
        vbc[:, k, k % nc] = r.uniform(0.2, 0.6)
    u = np.full((nc, nc), u_offdiag if u_offdiag is not None else 0.0)
    np.fill_diagonal(u, U)
    m = AIMParams(nc, nb, ec, eb, vbc, u, np.zeros((nc, nc)), 0.0, 1e-10)
    pb = r.uniform(-0.4, 0.4, (nb, nb))
    pvbc = r.uniform(-0.3, 0.3, (2, nb, nc))
    return m, pb, pvbc


@pytest.mark.parametrize("nc,nb,npart,pair_ib,u_off", [(1, 3, 4, False, None), (1, 3, 5, True, None),
                                                      (2, 2, 4, False, 1.3), (2, 2, 3, True, 1.1),
                                                      (1, 4, 6, True, None)])
def test_nambu_spectrum_equals_physical(nc, nb, npart, pair_ib, u_off):
    m, pb, pvbc = sc_model(nc, nb, 2.0, seed=11 + nb + npart, u_offdiag=u_off, spin_split=0.07)
    tab = osz.tab_sz(m, npart, pb, pvbc, pair_ib)
    H = osz.dense_sz(tab)
    assert np.allclose(H, H.T, atol=1e-14)
    e_n = np.linalg.eigvalsh(H) + osz.energy_shift(m)
    e_p = osz.physical_spectrum(m, npart, pb, pvbc, pair_ib)
    assert len(e_n) == len(e_p) == osz.sz_dim(m.ns, npart)
    assert np.max(np.abs(e_n - e_p)) < 1e-11


def test_plant_a_mistake_is_caught():
    """A Nambu bath down block without the minus (a plausible slip) breaks the spectral
    equality the test above relies on.  (Flipping the sign of a whole hybridisation
    block would not: that is a gauge transformation of the down orbitals.)"""
    m, pb, pvbc = sc_model(2, 2, 2.0, seed=5, u_offdiag=1.2, spin_split=0.05)
    good = osz.nambu_matrices

    def wrong(model, pb=None, pvbc=None, pairing_imp_to_bath=False, im=None):
        ecn, ebn, vbn = good(model, pb, pvbc, pairing_imp_to_bath, im)
        ebn[model.nb:, model.nb:] *= -1.0
        return ecn, ebn, vbn

    osz.nambu_matrices = wrong
    try:
        H = osz.dense_sz(osz.tab_sz(m, 4, pb, pvbc, False))
    finally:
        osz.nambu_matrices = good
    e_n = np.linalg.eigvalsh(H) + osz.energy_shift(m)
    e_p = osz.physical_spectrum(m, 4, pb, pvbc, False)
    assert np.max(np.abs(e_n - e_p)) > 1e-6


def test_no_pairing_reduces_to_spin_sectors():
    """Without pairing the S_z block is the direct sum of (N_up, N_dn) sectors with
    N_up - N_dn fixed: its spectrum is the union of those sectors' spectra."""
    import oracle.brute as bf
    m, _, _ = sc_model(1, 3, 2.5, seed=3)
    npart = 4   # Ns = 4: N_up - N_dn = 0
    e_n = np.sort(np.linalg.eigvalsh(osz.dense_sz(osz.tab_sz(m, npart))) + osz.energy_shift(m))
    parts = [np.linalg.eigvalsh(bf.sector_hamiltonian(m, n, n).toarray()) for n in range(m.ns + 1)]
    assert np.max(np.abs(e_n - np.sort(np.concatenate(parts)))) < 1e-11


@pytest.mark.parametrize("nc,nb,npart,pair_ib", [(2, 2, 4, False), (2, 2, 4, True), (3, 2, 4, True), (2, 3, 5, True),
                                                 (3, 2, 5, False)])
def test_hund_terms_match_physical_spectrum(nc, nb, npart, pair_ib):
    """The spin-flip and pair-hopping terms of HAIMsz_multc and the -J parallel-spin terms
    of QUART_INT_SZ: with them the Nambu spectrum + shift equals the physical spectrum of
    a Kanamori impurity (U, U' = U - 2J, J) with pairing; without them, or with their sign
    flipped, it does not."""
    m0, pb, pvbc = sc_model(nc, nb, 3.0, seed=7 + nb + npart, u_offdiag=1.6, spin_split=0.05)
    jj = np.full((nc, nc), 0.6)
    np.fill_diagonal(jj, 0.0)
    m = AIMParams(nc, nb, m0.ec, m0.eb, m0.vbc, m0.u, jj, 0.6, m0.cutoff)
    P = pb
    tab = osz.tab_sz(m, npart, P, pvbc if pair_ib else None, pair_ib)
    hund = osz.hund_sz(m, npart)
    assert len(hund) > 0
    ep = np.sort(osz.physical_spectrum(m, npart, P, pvbc if pair_ib else None, pair_ib))
    e = np.sort(np.linalg.eigvalsh(osz.dense_sz(tab, hund))) + osz.energy_shift(m)
    assert np.max(np.abs(e - ep)) < 1e-12
    for bad in ([], [(i, j, -h) for i, j, h in hund]):
        eb = np.sort(np.linalg.eigvalsh(osz.dense_sz(tab, bad))) + osz.energy_shift(m)
        assert np.max(np.abs(eb - ep)) > 1e-3


NAMBU_G_CASES = [(1, 2, 0.0, False), (2, 1, 0.5, False), (1, 2, 0.0, True), (2, 1, 0.4, True)]


@pytest.mark.parametrize("nc,nb,jh,pair_ib", NAMBU_G_CASES)
def test_nambu_green_restated_equals_physical(nc, nb, jh, pair_ib):
    """The Nambu G^N_ij = <<d_i ; d_j+>> from the restated S_z sectors (apply_Csz on every
    npart sector) equals the same Lehmann sum of the physical Hamiltonian on the whole
    Fock space: normal blocks, and the anomalous <<c_up ; c_dn>> of the pairing."""
    r = np.random.default_rng(40 + nc * 3 + nb)
    m = kanamori(nc, nb, 2.2, jh, r.uniform(-1, 1, nb), r.uniform(0.2, 0.6, nb))
    pb = r.uniform(-0.3, 0.3, (nb, nb)); pb = (pb + pb.T) / 2
    pvbc = r.uniform(-0.3, 0.3, (2, nb, nc)) if pair_ib else None
    beta = 6.0
    wn = np.pi / beta * (2 * np.arange(1, 9) - 1)
    Gr = osz.nambu_green_restated(m, beta, wn, pb=pb, pvbc=pvbc, pairing_imp_to_bath=pair_ib)
    Gp = osz.nambu_green_physical(m, beta, wn, pb=pb, pvbc=pvbc, pairing_imp_to_bath=pair_ib)
    assert np.max(np.abs(Gp[:nc, nc:])) > 1e-3   # an anomalous part is present
    assert np.max(np.abs(Gr - Gp)) < 1e-12


def test_nambu_green_without_pairing_is_spin_green():
    """No pairing: the up block is G_up of the (N_up, N_dn) route, the down block is
    <<c_dn+ ; c_dn>>(i w) = -conj(G_dn(i w)), the anomalous block vanishes; and a planted
    mistake in apply_c_sz (the sign counted below the bit) changes G."""
    r = np.random.default_rng(7)
    m = kanamori(1, 2, 2.0, 0.0, r.uniform(-1, 1, 2), r.uniform(0.2, 0.6, 2))
    beta = 5.0
    wn = np.pi / beta * (2 * np.arange(1, 9) - 1)
    Gr = osz.nambu_green_restated(m, beta, wn)
    ns = m.ns
    spec = {(u, d): np.linalg.eigh(bf.sector_hamiltonian(m, u, d).toarray()) for u in range(ns + 1)
            for d in range(ns + 1)}
    e0 = min(v[0].min() for v in spec.values())
    Z = sum(np.exp(-beta * (v[0] - e0)).sum() for v in spec.values())
    g = {0: np.zeros(len(wn), complex), 1: np.zeros(len(wn), complex)}
    for spin in (0, 1):
        for (u, d), (em, vm) in spec.items():
            tu, td = (u + 1, d) if spin == 0 else (u, d + 1)
            if tu > ns or td > ns:
                continue
            en, vn = spec[(tu, td)]
            for a in range(len(em)):
                amp = vn.T @ bf.apply_c_full(m, u, d, 0, spin, +1, vm[:, a])
                w = np.exp(-beta * (em[a] - e0)) + np.exp(-beta * (en - e0))
                g[spin] += (amp ** 2 * w)[:, None].__truediv__(1j * wn[None, :] - (en - em[a])[:, None]).sum(0)
    assert np.max(np.abs(Gr[0, 0] - g[0] / Z)) < 1e-12
    assert np.max(np.abs(Gr[1, 1] + np.conj(g[1] / Z))) < 1e-12
    assert np.max(np.abs(Gr[0, 1])) < 1e-14
    # the planted sign reaches apply_c_sz only; the tables keep their own hop signs
    good, good_hop = osz._sign_above, osz._hop

    def hop_with_good_sign(s, to, frm):
        planted, osz._sign_above = osz._sign_above, good
        try:
            return good_hop(s, to, frm)
        finally:
            osz._sign_above = planted
    try:
        # the (N_up, N_dn) convention: only the orbitals above the bit in its own spin string
        top = lambda bit: m.ns if bit < m.ns else 2 * m.ns
        osz._sign_above = lambda s, bit: -1 if bin((s >> (bit + 1)) & ((1 << (top(bit) - bit - 1)) - 1)).count(
            "1") & 1 else 1
        osz._hop = hop_with_good_sign
        bad = osz.nambu_green_restated(m, beta, wn, pb=np.full((2, 2), 0.2))
    finally:
        osz._sign_above, osz._hop = good, good_hop
    ok = osz.nambu_green_restated(m, beta, wn, pb=np.full((2, 2), 0.2))
    assert np.max(np.abs(bad - ok)) > 1e-4


@pytest.mark.parametrize("nc,nb,jh,pair_ib", [(1, 2, 0.0, False), (2, 1, 0.5, True), (1, 3, 0.0, True)])
def test_ncup_green_restated_equals_physical(nc, nb, jh, pair_ib):
    """The correlated-hopping functions of FLAG_NCUP from the restated S_z route
    (apply_Csz_N_updo's condition on the other Nambu orbital of the site) equal the
    Lehmann sums of n_dn c_up and n_up c+_dn written as physical operator strings on the
    whole Fock space: GN = <<D ; D+>> and the cross GNF = <<D ; d+>>.  A planted mistake
    (the condition inverted) fails."""
    r = np.random.default_rng(90 + nc * 5 + nb)
    m = kanamori(nc, nb, 2.4, jh, r.uniform(-1, 1, nb), r.uniform(0.2, 0.6, nb))
    pb = r.uniform(-0.3, 0.3, (nb, nb)); pb = (pb + pb.T) / 2
    pvbc = r.uniform(-0.3, 0.3, (2, nb, nc)) if pair_ib else None
    beta = 7.0
    wn = np.pi / beta * (2 * np.arange(1, 9) - 1)
    gn_r, gnf_r = osz.ncup_green_restated(m, beta, wn, pb=pb, pvbc=pvbc, pairing_imp_to_bath=pair_ib)
    gn_p, gnf_p = osz.ncup_green_physical(m, beta, wn, pb=pb, pvbc=pvbc, pairing_imp_to_bath=pair_ib)
    assert np.max(np.abs(gn_r - gn_p)) < 1e-12
    assert np.max(np.abs(gnf_r - gnf_p)) < 1e-12
    G = osz.nambu_green_physical(m, beta, wn, pb=pb, pvbc=pvbc, pairing_imp_to_bath=pair_ib)
    for a in range(2 * nc):   # the projector matters: GN differs from G and from GNF
        assert np.max(np.abs(gn_p[a] - G[a, a])) > 1e-3
        assert np.max(np.abs(gn_p[a] - gnf_p[0, a])) > 1e-4
    assert np.max(np.abs(gnf_p[0] - gnf_p[1])) < 1e-13   # real H: <<D ; d+>> = <<d ; D+>>
    good = osz.apply_c_n_sz
    try:
        def inverted(model, npart, site, spin, pm, psi):
            full = osz.apply_c_sz(model, npart, site, spin, pm, psi)
            return full - good(model, npart, site, spin, pm, psi)
        osz.apply_c_n_sz = inverted
        bad, _ = osz.ncup_green_restated(m, beta, wn, pb=pb, pvbc=pvbc, pairing_imp_to_bath=pair_ib)
    finally:
        osz.apply_c_n_sz = good
    assert np.max(np.abs(bad - gn_p)) > 1e-3


def complex_sc_model(nc, nb, seed, jhund=0.0):
    """A complex model (Hermitian Ec, Eb, complex Vbc; oracle.cplx.random_model) with complex
    bath pairing Pb and impurity-bath pairing PVbc."""
    import oracle.cplx as oc
    m, im = oc.random_model(nc, nb, seed, jhund=jhund, u_off=1.1 if nc > 1 else 1.0)
    r = np.random.default_rng(seed + 100)
    pb = r.uniform(-0.4, 0.4, (nb, nb)) + 1j * r.uniform(-0.4, 0.4, (nb, nb))
    pvbc = r.uniform(-0.3, 0.3, (2, nb, nc)) + 1j * r.uniform(-0.3, 0.3, (2, nb, nc))
    return m, im, pb, pvbc


@pytest.mark.parametrize("nc,nb,npart,pair_ib,jh", [(1, 3, 4, False, 0.0), (1, 3, 5, True, 0.0),
                                                     (2, 2, 4, True, 0.0), (2, 2, 3, False, 0.4),
                                                     (2, 2, 4, True, 0.3)])
def test_complex_nambu_spectrum_equals_physical(nc, nb, npart, pair_ib, jh):
    """Complex S_z sectors (HAIMsz_multc with COMPLEX(8) tables): the restated Nambu matrices
    of the _complex branches, the complex tab_HAIM2 over them and the conj(h) gather
    give a Hermitian H whose spectrum + shift is the physical spectrum of the complex
    Hamiltonian with complex pairing, written as operator strings on the full Fock space."""
    m, im, pb, pvbc = complex_sc_model(nc, nb, seed=3 + nb + npart, jhund=jh)
    tab = osz.tab_sz(m, npart, pb, pvbc, pair_ib, im=im)
    assert np.iscomplexobj(tab["offdiag"]) and np.max(np.abs(tab["offdiag"].imag)) > 1e-3
    H = osz.dense_sz(tab, osz.hund_sz(m, npart))
    assert np.allclose(H, np.conj(H.T), atol=1e-14)
    e_n = np.linalg.eigvalsh(H) + osz.energy_shift(m)
    e_p = osz.physical_spectrum(m, npart, pb, pvbc, pair_ib, im=im)
    assert np.max(np.abs(e_n - e_p)) < 1e-11
    v = np.random.default_rng(npart).standard_normal(len(e_n)) * (1 + 0.5j)
    assert np.max(np.abs(osz.hmult_sz(tab, v, osz.hund_sz(m, npart)) - H @ v)) < 1e-12


@pytest.mark.parametrize("plant", ["no_conj_pairing", "conj_dn_ec", "no_conj_vbc_dn"])
def test_complex_nambu_planted_mistakes(plant):
    """Plausible slips in the complex Nambu matrices change the spectrum: Pb^T below the
    diagonal instead of conj(Pb)^T, conj(-Ec_dn^T) instead of -Ec_dn^T, -Vbc_dn
    instead of -conj(Vbc_dn)."""
    m, im, pb, pvbc = complex_sc_model(2, 2, seed=17)
    good = osz.nambu_matrices
    nc, nb = m.nc, m.nb
    import oracle.cplx as oc

    def wrong(model, pb=None, pvbc=None, pairing_imp_to_bath=False, im=None):
        ecn, ebn, vbn = good(model, pb, pvbc, pairing_imp_to_bath, im)
        if plant == "no_conj_pairing":
            ebn[nb:, :nb] = np.asarray(pb).T
        elif plant == "conj_dn_ec":
            ecn[nc:, nc:] = np.conj(ecn[nc:, nc:])
        else:
            vbn[nb:, nc:] = -oc.complex_matrices(model, im, 1)[2]
        return ecn, ebn, vbn

    e_p = osz.physical_spectrum(m, 4, pb, pvbc, True, im=im)
    try:
        osz.nambu_matrices = wrong
        tab = osz.tab_sz(m, 4, pb, pvbc, True, im=im)
    finally:
        osz.nambu_matrices = good
    H = osz.dense_sz(tab)
    if np.allclose(H, np.conj(H.T)):
        e_n = np.linalg.eigvalsh(H) + osz.energy_shift(m)
        assert np.max(np.abs(e_n - e_p)) > 1e-4


@pytest.mark.parametrize("nc,nb,jh,pair_ib", [(1, 2, 0.0, True), (2, 1, 0.4, False)])
def test_complex_nambu_and_ncup_green_restated_equal_physical(nc, nb, jh, pair_ib):
    """Complex model with complex pairing: the Nambu G^N and the FLAG_NCUP GN / GNF from the
    restated complex S_z sectors equal the Lehmann sums of the physical complex Hamiltonian;
    G_ij != G_ji and <<D ; d+>> != <<d ; D+>> there, so both orders are needed."""
    m, im, pb, pvbc = complex_sc_model(nc, nb, seed=31 + nc + nb, jhund=jh)
    beta = 5.0
    wn = np.pi / beta * (2 * np.arange(1, 9) - 1)
    Gr = osz.nambu_green_restated(m, beta, wn, pb=pb, pvbc=pvbc, pairing_imp_to_bath=pair_ib, im=im)
    Gp = osz.nambu_green_physical(m, beta, wn, pb=pb, pvbc=pvbc, pairing_imp_to_bath=pair_ib, im=im)
    assert np.max(np.abs(Gr - Gp)) < 1e-12
    assert np.max(np.abs(Gp[0, 1] - Gp[1, 0])) > 1e-4
    gn_r, gnf_r = osz.ncup_green_restated(m, beta, wn, pb=pb, pvbc=pvbc, pairing_imp_to_bath=pair_ib, im=im)
    gn_p, gnf_p = osz.ncup_green_physical(m, beta, wn, pb=pb, pvbc=pvbc, pairing_imp_to_bath=pair_ib, im=im)
    assert np.max(np.abs(gn_r - gn_p)) < 1e-12
    assert np.max(np.abs(gnf_r - gnf_p)) < 1e-12
    assert np.max(np.abs(gnf_p[0] - gnf_p[1])) > 1e-5
