"""Brute-force second quantisation for small Anderson-impurity sectors (oracle, tests only).

Independent of the C restatement: every Hamiltonian term is written as a product
of creation / annihilation operators and applied to full Fock states with ONE
global Jordan-Wigner convention (mode m = bit m of s = s_up + 2^Ns s_dn, sign
(-1)^{# occupied modes above m}).  The per-spin sign bookkeeping of the reference
(separate up and down strings, "between" parities for the Hund terms) is therefore
not reused here; agreement with the C restatement pins it.

Operator content follows the reference's Hamiltonian (ed_solver/HAIMupdo.f90:677-1195,
ed_solver/HAIM_offdiag_arrays.f90:419-527):

  H = sum_s sum_ab h^s_ab c+_as c_bs                          (one-body, cutoff on a != b)
    + sum_i U_ii n_iu n_id
    + sum_{i<j, U_ij != 0} [U_ij n_i n_j - J_ij (n_iu n_ju + n_id n_jd)]
    + [|Jhund| > 1e-5 and any U_ij != 0]:
        sum_{i<j, U_ij != 0} J_ij (c+_ju c_iu c+_id c_jd + h.c.)   (spin flip)
        sum_{i != j, U_ij != 0} J_ij c+_ju c_iu c+_jd c_id          (pair hopping)
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from ed_solver_amd.aim import AIMParams

_POP8 = np.array([bin(i).count("1") for i in range(256)], dtype=np.int64)


def popcount(x: np.ndarray) -> np.ndarray:
    x = x.astype(np.int64)
    c = np.zeros_like(x)
    for sh in range(0, 40, 8):
        c += _POP8[(x >> sh) & 0xFF]
    return c


def one_body(model: AIMParams, spin: int) -> np.ndarray:
    """h^spin on the Ns orbitals (bath k -> k, impurity c -> nb + c)."""
    nc, nb = model.nc, model.nb
    h = np.zeros((model.ns, model.ns))
    h[nb:, nb:] = model.ec[spin]
    h[:nb, :nb] = model.eb[spin]
    h[:nb, nb:] = model.vbc[spin]
    h[nb:, :nb] = model.vbc[spin].T
    off = ~np.eye(model.ns, dtype=bool)
    h[off & (np.abs(h) <= model.cutoff)] = 0.0
    return h


def terms(model: AIMParams):
    """List of (coef, ops) with ops = [(mode, is_creation), ...] applied right to left."""
    ns, nb, nc = model.ns, model.nb, model.nc
    T = []
    for s in (0, 1):
        h = one_body(model, s)
        for a in range(ns):
            for b in range(ns):
                if h[a, b] != 0.0:
                    T.append((h[a, b], [(a + s * ns, True), (b + s * ns, False)]))
    U, JJ = model.u, model.jj
    up = lambda c: nb + c
    dn = lambda c: nb + c + ns
    for i in range(nc):
        if U[i, i] != 0.0:
            T.append((U[i, i], [(up(i), True), (up(i), False), (dn(i), True), (dn(i), False)]))
    offc = any(U[i, j] != 0.0 for i in range(nc) for j in range(nc) if i != j)
    for i in range(nc):
        for j in range(i + 1, nc):
            if U[i, j] == 0.0:
                continue
            for a in (up(i), dn(i)):
                for b in (up(j), dn(j)):
                    T.append((U[i, j], [(a, True), (a, False), (b, True), (b, False)]))
            T.append((-JJ[i, j], [(up(i), True), (up(i), False), (up(j), True), (up(j), False)]))
            T.append((-JJ[i, j], [(dn(i), True), (dn(i), False), (dn(j), True), (dn(j), False)]))
    if abs(model.jhund) > 1e-5 and offc:
        for i in range(nc):
            for j in range(i + 1, nc):
                if U[i, j] == 0.0:
                    continue
                T.append((JJ[i, j], [(up(j), True), (up(i), False), (dn(i), True), (dn(j), False)]))
                T.append((JJ[i, j], [(up(i), True), (up(j), False), (dn(j), True), (dn(i), False)]))
        for i in range(nc):
            for j in range(nc):
                if i == j or U[i, j] == 0.0:
                    continue
                T.append((JJ[i, j], [(up(j), True), (up(i), False), (dn(j), True), (dn(i), False)]))
    return T


def apply_ops(states: np.ndarray, ops):
    """Apply an operator string (right to left) to full Fock states; returns (out, sign, ok)."""
    s = states.astype(np.int64).copy()
    sign = np.ones(len(s), dtype=np.int64)
    ok = np.ones(len(s), dtype=bool)
    for mode, cre in reversed(ops):
        occ = ((s >> mode) & 1).astype(bool)
        ok &= (~occ) if cre else occ
        above = popcount(s >> (mode + 1))
        sign *= np.where(above % 2 == 1, -1, 1)
        s = np.where(ok, s ^ (1 << mode), s)
    return s, sign, ok


def sector_basis(ns: int, nup: int, ndn: int):
    allst = np.arange(1 << ns)
    pc = popcount(allst)
    ups = allst[pc == nup]
    dns = allst[pc == ndn]
    return ups, dns


def sector_hamiltonian(model: AIMParams, nup: int, ndn: int) -> sp.csr_matrix:
    """Sparse H in the (nup, ndn) sector, index iup + ido*d_up."""
    ns = model.ns
    ups, dns = sector_basis(ns, nup, ndn)
    du, dd = len(ups), len(dns)
    rank = np.full(1 << ns, -1, dtype=np.int64)
    rank[ups] = np.arange(du)
    rank_dn = np.full(1 << ns, -1, dtype=np.int64)
    rank_dn[dns] = np.arange(dd)
    iu, idn = np.meshgrid(np.arange(du), np.arange(dd), indexing="ij")
    iu = iu.ravel(order="F"); idn = idn.ravel(order="F")
    col = iu + idn * du
    full = ups[iu] + (dns[idn] << ns)
    rows, cols, vals = [], [], []
    for coef, ops in terms(model):
        out, sg, ok = apply_ops(full, ops)
        if not ok.any():
            continue
        o = out[ok]
        ou = o & ((1 << ns) - 1)
        od = o >> ns
        r = rank[ou] + rank_dn[od] * du
        assert (rank[ou] >= 0).all() and (rank_dn[od] >= 0).all()
        rows.append(r); cols.append(col[ok]); vals.append(coef * sg[ok])
    D = du * dd
    if not rows:
        return sp.csr_matrix((D, D))
    H = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(D, D))
    return H.tocsr()


def apply_c_full(model: AIMParams, nup: int, ndn: int, site: int, spin: int, pm: int, psi: np.ndarray):
    """c+ (pm=+1) or c (pm=-1) of impurity orbital `site`, spin (0 up / 1 down) on psi.

    The reference signs the operator inside one spin string only (apply_C.f90:564-566:
    the ket is the spin-resolved state), so the sign here counts occupied modes above
    the target inside that spin's string."""
    ns = model.ns
    ups, dns = sector_basis(ns, nup, ndn)
    du = len(ups)
    n2 = (nup + pm, ndn) if spin == 0 else (nup, ndn + pm)
    ups2, dns2 = sector_basis(ns, *n2)
    du2 = len(ups2)
    r_up2 = np.full(1 << ns, -1); r_up2[ups2] = np.arange(du2)
    r_dn2 = np.full(1 << ns, -1); r_dn2[dns2] = np.arange(len(dns2))
    phi = np.zeros(du2 * len(dns2), np.result_type(psi, np.float64))
    mode = model.nb + site
    src = ups if spin == 0 else dns
    out, sg, ok = apply_ops(src, [(mode, pm > 0)])
    for i in np.nonzero(ok)[0]:
        j = out[i]
        if spin == 0:
            for ido in range(len(dns)):
                phi[r_up2[j] + ido * du2] += sg[i] * psi[i + ido * du]
        else:
            for iup in range(du):
                phi[iup + r_dn2[j] * du2] += sg[i] * psi[iup + i * du]
    return phi
