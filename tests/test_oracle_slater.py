"""The Slater 4-index interaction oracle (CPU): the restated gather loops of
flag_slater_int against brute-force second quantisation for random tensors, and the
Kanamori model written as a Slater tensor against the model's own U / J terms."""
import numpy as np
import pytest

import oracle.brute as bf
import oracle.slater as osl
from ed_solver_amd.aim import AIMParams, kanamori, slater_tensor


@pytest.mark.parametrize("nc,nb,sec", [(2, 2, (2, 2)), (3, 1, (2, 2)), (2, 3, (3, 2)), (3, 2, (3, 1))])
def test_restated_slater_matches_brute_force(nc, nb, sec):
    r = np.random.default_rng(nc * 7 + nb)
    m = kanamori(nc, nb, 2.0, 0.3 if nc > 1 else 0.0, r.uniform(-1, 1, nb), r.uniform(0.2, 0.6, nb))
    u4 = r.uniform(-0.5, 0.5, (nc,) * 4)   # no symmetry: every index and sign is seen
    H = osl.sector_hamiltonian(m, u4, *sec)
    D = H.shape[0]
    Hs = np.column_stack([osl.hmult_slater(m, u4, *sec, np.eye(D)[:, i]) for i in range(D)])
    H0 = bf.sector_hamiltonian(m, *sec).toarray()
    assert np.max(np.abs(Hs)) > 0.5
    assert np.max(np.abs(H - H0 - Hs)) < 1e-13


def test_kanamori_as_slater_tensor():
    """U = J = 0 plus slater_tensor(model) has the model's spectrum (density, spin flip,
    pair hopping); a planted mistake (the pair-hopping entries dropped) changes it."""
    r = np.random.default_rng(2)
    for nc, nb in ((2, 2), (3, 1)):
        m = kanamori(nc, nb, 2.5, 0.4, r.uniform(-1, 1, nb), r.uniform(0.2, 0.6, nb))
        m0 = AIMParams(nc, nb, m.ec, m.eb, m.vbc, np.zeros((nc, nc)), np.zeros((nc, nc)), 0.0, m.cutoff)
        u4 = slater_tensor(m)
        bad = u4.copy()
        for a in range(nc):
            for b in range(nc):
                if a != b:
                    bad[b, b, a, a] = 0.0
        for sec in ((2, 2), (2, 1)):
            e = np.linalg.eigvalsh(bf.sector_hamiltonian(m, *sec).toarray())
            assert np.max(np.abs(e - np.linalg.eigvalsh(osl.sector_hamiltonian(m0, u4, *sec)))) < 1e-12
            assert np.max(np.abs(e - np.linalg.eigvalsh(osl.sector_hamiltonian(m0, bad, *sec)))) > 1e-3
