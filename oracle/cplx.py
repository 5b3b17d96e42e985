"""Oracle for complex (N_up, N_dn) sectors -- test infrastructure only.

Two independent routes, as for the real sectors:

* a restatement of the reference's complex path in NumPy/Python loops for small sectors:
  tab_HAIM2 with COMPLEX(8) coefficients (ed_solver/HAIM_offdiag_arrays.f90:419-527:
  Ec(a,b) sign, Eb(a,b) sign, Vbc(b,c) sign for impurity -> bath, conj(Vbc(b,c)) sign for
  bath -> impurity, each kept when nonzero and |t| > cutoff) and the gather H.v of
  HAIMupdo_multc (ed_solver/HAIMupdo.f90:150-656, USE_CC: w_i += conj(h) v_j for the
  up and the down hops); the interaction terms are real and are those of the real
  model (the same operator strings);
* the brute-force second quantisation of oracle.brute with complex one-body
  coefficients h_ab c+_a c_b (h Hermitian, Vbc on b+_k c_c).

Agreement of the two on every matrix element pins the restatement (a missing conj,
a swapped index or a wrong sign shows as a non-Hermitian or different matrix).  A
phase transformation c_a -> exp(i theta_a) c_a of a real model gives a complex model
with the same spectrum and, for phases equal on the impurity orbitals or a model
without pair hopping, the same diagonal impurity Green's function: that pins the
device results at sizes the dense matrix cannot reach.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from ed_solver_amd.aim import AIMParams

from . import brute as bf
from .sz import _hop


def complex_matrices(model: AIMParams, im: dict, spin: int):
    """(Ec, Eb, Vbc) of one spin as complex arrays."""
    ec = model.ec[spin] + 1j * np.asarray(im["ec"])[spin]
    eb = model.eb[spin] + 1j * np.asarray(im.get("eb", np.zeros((2, model.nb, model.nb))))[spin]
    vbc = model.vbc[spin] + 1j * np.asarray(im.get("vbc", np.zeros((2, model.nb, model.nc))))[spin]
    return ec, eb, vbc


def _keep(t: complex, cut: float) -> bool:
    return t != 0 and abs(t) > cut


def tab_c(model: AIMParams, im: dict, spin: int, n: int) -> dict:
    """tab_HAIM2 of one spin with complex entries, every state of popcount n in order."""
    nc, nb, ns = model.nc, model.nb, model.ns
    ec, eb, vbc = complex_matrices(model, im, spin)
    states = bf.popcount(np.arange(1 << ns))
    states = np.arange(1 << ns)[states == n]
    rank = {int(s): i for i, s in enumerate(states)}
    ibit = [nb + c for c in range(nc)]
    bbit = list(range(nb))
    cut = model.cutoff
    diag = np.zeros(len(states)); noff = np.zeros(len(states), np.int32)
    rk, off = [], []
    for i, s in enumerate(states):
        s = int(s)
        d = 0.0
        for c in range(nc):
            if (s >> ibit[c]) & 1:
                d += ec[c, c].real
        for b in range(nb):
            if (s >> bbit[b]) & 1:
                d += eb[b, b].real
        diag[i] = d
        n0 = len(rk)

        def add(t, to, frm):
            r = _hop(s, to, frm)
            if r is not None:
                rk.append(rank[r[0]]); off.append(t * r[1])

        for a in range(nc):
            for b in range(nc):
                if a != b and _keep(ec[a, b], cut):
                    add(ec[a, b], ibit[a], ibit[b])
        for a in range(nb):
            for b in range(nb):
                if a != b and _keep(eb[a, b], cut):
                    add(eb[a, b], bbit[a], bbit[b])
        for b in range(nb):
            for c in range(nc):
                t = vbc[b, c]
                if _keep(t, cut):
                    add(t, bbit[b], ibit[c])
                    add(np.conj(t), ibit[c], bbit[b])
        noff[i] = len(rk) - n0
    return {"states": states.astype(np.int32), "diag": diag, "noff": noff, "rankoff": np.array(rk, np.int32),
            "offdiag": np.array(off, np.complex128)}


def interaction_model(model: AIMParams) -> AIMParams:
    """The model with every one-body term removed (interaction and Hund terms only)."""
    z = lambda a: np.zeros_like(a)
    return AIMParams(model.nc, model.nb, z(model.ec), z(model.eb), z(model.vbc), model.u, model.jj, model.jhund,
                     model.cutoff)


def hmult_c(model: AIMParams, im: dict, nup: int, ndn: int, v: np.ndarray, tabs=None, hint=None) -> np.ndarray:
    """HAIMupdo_multc in gather form on the column-major d_up x d_dn amplitude matrix:
    w = H_int v + (diag_up + diag_dn) v + sum_up conj(h) v(jup, ido) + sum_dn conj(h) v(iup, jdo)."""
    tu, td = tabs if tabs is not None else (tab_c(model, im, 0, nup), tab_c(model, im, 1, ndn))
    du, dd = len(tu["states"]), len(td["states"])
    V = np.asarray(v, np.complex128).reshape(dd, du).T
    W = (tu["diag"][:, None] + td["diag"][None, :]) * V
    p = np.concatenate([[0], np.cumsum(tu["noff"])])
    for i in range(du):
        for k in range(p[i], p[i + 1]):
            W[i, :] += np.conj(tu["offdiag"][k]) * V[tu["rankoff"][k], :]
    p = np.concatenate([[0], np.cumsum(td["noff"])])
    for i in range(dd):
        for k in range(p[i], p[i + 1]):
            W[:, i] += np.conj(td["offdiag"][k]) * V[:, td["rankoff"][k]]
    if hint is None:
        hint = bf.sector_hamiltonian(interaction_model(model), nup, ndn)
    return W.T.reshape(-1) + hint @ np.asarray(v, np.complex128)


def one_body_c(model: AIMParams, im: dict, spin: int) -> np.ndarray:
    """Complex h^spin on the Ns orbitals: h[k, c] = Vbc (b+_k c_c), h[c, k] = conj(Vbc)."""
    nb = model.nb
    ec, eb, vbc = complex_matrices(model, im, spin)
    h = np.zeros((model.ns, model.ns), np.complex128)
    h[nb:, nb:] = ec
    h[:nb, :nb] = eb
    h[:nb, nb:] = vbc
    h[nb:, :nb] = np.conj(vbc).T
    off = ~np.eye(model.ns, dtype=bool)
    h[off & (np.abs(h) <= model.cutoff)] = 0.0
    return h


def sector_hamiltonian_c(model: AIMParams, im: dict, nup: int, ndn: int) -> sp.csr_matrix:
    """Brute-force complex H in the (nup, ndn) sector, index iup + ido d_up."""
    ns = model.ns
    T = []
    for s in (0, 1):
        h = one_body_c(model, im, s)
        for a in range(ns):
            for b in range(ns):
                if h[a, b] != 0:
                    T.append((h[a, b], [(a + s * ns, True), (b + s * ns, False)]))
    T += [t for t in bf.terms(interaction_model(model))]
    ups, dns = bf.sector_basis(ns, nup, ndn)
    du, dd = len(ups), len(dns)
    rank = np.full(1 << ns, -1, dtype=np.int64); rank[ups] = np.arange(du)
    rank_dn = np.full(1 << ns, -1, dtype=np.int64); rank_dn[dns] = np.arange(dd)
    iu, idn = np.meshgrid(np.arange(du), np.arange(dd), indexing="ij")
    iu = iu.ravel(order="F"); idn = idn.ravel(order="F")
    col = iu + idn * du
    full = ups[iu] + (dns[idn] << ns)
    rows, cols, vals = [], [], []
    for coef, ops in T:
        out, sg, ok = bf.apply_ops(full, ops)
        if not ok.any():
            continue
        o = out[ok]
        r = rank[o & ((1 << ns) - 1)] + rank_dn[o >> ns] * du
        rows.append(r); cols.append(col[ok]); vals.append(coef * sg[ok])
    D = du * dd
    if not vals:   # the empty and the full sector hold no one-body or interaction matrix element
        return sp.csr_matrix((D, D), dtype=np.complex128)
    H = sp.coo_matrix((np.concatenate(vals).astype(np.complex128), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(D, D))
    return H.tocsr()


def phase_model(model: AIMParams, theta: np.ndarray):
    """c_a,s -> exp(i theta[s, a]) c_a,s on the Ns orbitals (bath k -> k, impurity c -> nb + c):
    h' = P^+ h P.  Returns (real-part model, im dict) of the transformed complex model."""
    nc, nb = model.nc, model.nb
    ec, eb, vbc, ec_i, eb_i, vbc_i = [], [], [], [], [], []
    for s in (0, 1):
        h = np.zeros((model.ns, model.ns), np.complex128)
        h[nb:, nb:] = model.ec[s]
        h[:nb, :nb] = model.eb[s]
        h[:nb, nb:] = model.vbc[s]
        h[nb:, :nb] = model.vbc[s].T
        ph = np.exp(1j * np.asarray(theta[s], float))
        h2 = np.conj(ph)[:, None] * h * ph[None, :]
        ec.append(h2[nb:, nb:].real); ec_i.append(h2[nb:, nb:].imag)
        eb.append(h2[:nb, :nb].real); eb_i.append(h2[:nb, :nb].imag)
        vbc.append(h2[:nb, nb:].real); vbc_i.append(h2[:nb, nb:].imag)
    m2 = AIMParams(nc, nb, np.array(ec), np.array(eb).reshape(2, nb, nb), np.array(vbc).reshape(2, nb, nc), model.u,
                   model.jj, model.jhund, model.cutoff)
    return m2, {"ec": np.array(ec_i), "eb": np.array(eb_i).reshape(2, nb, nb), "vbc": np.array(vbc_i).reshape(2, nb, nc)}


def random_model(nc: int, nb: int, seed: int, jhund: float = 0.0, u_off: float = 1.0):
    """A complex model for tests: Hermitian Ec, Eb with random complex entries, complex
    Vbc, Kanamori-style U (2.5 on the diagonal, u_off off it) and J = jhund.
    Returns (real-part model, im dict)."""
    r = np.random.default_rng(seed)

    def herm(n):
        a = r.uniform(-0.5, 0.5, (2, n, n)) + 1j * r.uniform(-0.5, 0.5, (2, n, n))
        return (a + np.conj(np.transpose(a, (0, 2, 1)))) / 2

    ec, eb = herm(nc), herm(nb)
    vbc = r.uniform(-0.6, 0.6, (2, nb, nc)) + 1j * r.uniform(-0.6, 0.6, (2, nb, nc))
    u = np.full((nc, nc), u_off)
    np.fill_diagonal(u, 2.5)
    jj = np.full((nc, nc), jhund)
    np.fill_diagonal(jj, 0.0)
    m = AIMParams(nc, nb, ec.real, eb.real, vbc.real, u, jj, jhund, 1e-8)
    return m, {"ec": ec.imag, "eb": eb.imag, "vbc": vbc.imag}

