"""The complex-sector oracle (CPU): the restated tab_HAIM2 + HAIMupdo_multc against the
brute-force complex Hamiltonian, entry by entry, and the phase transformation used to
build complex models with a known spectrum."""
import numpy as np
import pytest

import oracle.brute as bf
import oracle.cplx as oc
from ed_solver_amd.aim import AIMParams


random_complex_model = oc.random_model


def dense_restated(m, im, nup, ndn):
    tabs = (oc.tab_c(m, im, 0, nup), oc.tab_c(m, im, 1, ndn))
    hint = bf.sector_hamiltonian(oc.interaction_model(m), nup, ndn)
    D = len(tabs[0]["states"]) * len(tabs[1]["states"])
    return np.column_stack([oc.hmult_c(m, im, nup, ndn, np.eye(D)[:, i], tabs, hint) for i in range(D)]), tabs, hint


@pytest.mark.parametrize("nc,nb,nup,ndn,jh", [(2, 2, 2, 2, 0.4), (1, 3, 2, 1, 0.0), (2, 3, 3, 2, 0.3)])
def test_restated_hmult_matches_brute_force(nc, nb, nup, ndn, jh):
    m, im = random_complex_model(nc, nb, seed=nc * 10 + nb, jhund=jh)
    H, tabs, hint = dense_restated(m, im, nup, ndn)
    Hb = oc.sector_hamiltonian_c(m, im, nup, ndn).toarray()
    assert np.max(np.abs(H - Hb)) < 1e-13
    assert np.max(np.abs(H - H.conj().T)) < 1e-13
    assert np.max(np.abs(H.imag)) > 0.05   # genuinely complex
    # a planted mistake (the conj of HAIMupdo_multc dropped: H^T instead of H) is caught
    bad = tuple(dict(t, offdiag=np.conj(t["offdiag"])) for t in tabs)
    D = H.shape[0]
    Hbad = np.column_stack([oc.hmult_c(m, im, nup, ndn, np.eye(D)[:, i], bad, hint) for i in range(D)])
    assert np.max(np.abs(Hbad - Hb)) > 1e-3


def test_real_limit_equals_real_model():
    m, im = random_complex_model(2, 2, seed=3, jhund=0.4)
    zero = {k: np.zeros_like(v) for k, v in im.items()}
    H, _, _ = dense_restated(m, zero, 2, 1)
    Hr = bf.sector_hamiltonian(m, 2, 1).toarray()
    assert np.max(np.abs(H - Hr)) < 1e-14


def test_phase_transform_keeps_spectrum():
    """Phases on every orbital (J = 0) and on the bath only (J != 0, pair hopping present)."""
    r = np.random.default_rng(5)
    for jh, imp_phase in ((0.0, True), (0.4, False)):
        m, _ = random_complex_model(2, 3, seed=9, jhund=jh)
        theta = r.uniform(0, 2 * np.pi, (2, m.ns))
        if not imp_phase:
            theta[:, m.nb:] = 0.0
        m2, im2 = oc.phase_model(m, theta)
        assert np.max(np.abs(im2["vbc"])) > 0.05
        for nup, ndn in ((2, 2), (3, 2)):
            e_r = np.linalg.eigvalsh(bf.sector_hamiltonian(m, nup, ndn).toarray())
            e_c = np.linalg.eigvalsh(oc.sector_hamiltonian_c(m2, im2, nup, ndn).toarray())
            assert np.max(np.abs(e_r - e_c)) < 1e-12
