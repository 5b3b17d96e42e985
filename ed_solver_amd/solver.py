"""Sector scan, Boltzmann weights and the Green's function over many states.

Host orchestration that mirrors the reference's solver layer on top of the device
entry points (every H.v, Lanczos step, excitation and continued fraction runs in
libed_hip.so):

* ``get_eigen_energies``  solver.h:98-237 -- per sector: dense ED when the dimension
  is <= FULL_ED_SWITCH (Global_Variables.f90:23), else plain Lanczos, ARPACK or dense ED
  by ``which_lanczos`` (or block Lanczos when ``block_size`` > 0); then the per-sector
  window filter [E_s, E_s + dEmax] (filter_eigen) and the global one [E0, E0 + dEmax];
* ``partition``           Z and the Boltzmann weights e^{-beta (E - E0)} / Z
  (eigen_sector_class.f90:431-444, green_class_computeAA.h:107);
* ``compute_green``       compute_greenAA (green_class_computeAA.f90:40-185,
  green_class_computeAA.h:101-302) for the diagonal impurity Green's function:
  G(i w_n) = sum_states boltz * (particle + hole), each state's excitation vectors
  built on the device from one copy of the state and its runs batched.
"""
from __future__ import annotations

import dataclasses
from math import comb

import numpy as np

from .aim import AIMParams
from .hclass import (ComplexSector, Sector, apply_C, apply_C_n, compute_dynamic_batched, dynamic_from_states,
                     dynamic_full_ed)

FULL_ED_SWITCH = 40   # Global_Variables.f90:23


@dataclasses.dataclass
class EigenState:
    sector: tuple[int, int]
    energy: float
    vector: np.ndarray


# Least up-spin dimension the multi-run Lanczos kernel takes (ed_lanczos_batched runs
# smaller ones one by one).
_RUNS_MIN_DUP = 64


def spin_reversed(model: AIMParams) -> AIMParams:
    """The model with the spin labels exchanged (Ec, Eb, Vbc of spin up and down swapped;
    U, J and Jhund are spin-symmetric): its sector (N_dn, N_up) has the Hamiltonian of the
    original's (N_up, N_dn) in a relabelled basis."""
    return AIMParams(model.nc, model.nb, model.ec[::-1].copy(), model.eb[::-1].copy(), model.vbc[::-1].copy(),
                     model.u, model.jj, model.jhund, model.cutoff)


def all_sectors(ns: int) -> list[tuple[int, int]]:
    return [(nu, nd) for nu in range(ns + 1) for nd in range(ns + 1)]


def make_sector(model: AIMParams, nu: int, nd: int, im: dict | None = None, slater=None) -> Sector:
    """The (N_up, N_dn) sector: real (H_class) or, when im holds the imaginary parts of
    Ec / Eb / Vbc ({"ec": [2, nc, nc], "eb": [2, nb, nb], "vbc": [2, nb, nc]}), complex
    (Hmultc, HAIMupdo.f90:150-656); slater adds the 4-index impurity interaction
    (flag_slater_int) to a real sector."""
    if im is None:
        return Sector(model, nu, nd, slater)
    if slater is not None:
        raise ValueError("the Slater interaction is built for real sectors")
    return ComplexSector(model, nu, nd, im["ec"], im.get("eb"), im.get("vbc"))


def get_eigen_energies(model: AIMParams, sectors=None, neigen: int = 1, dEmax: float = np.inf, nitermax: int = 128,
                       tolerance: float = 1e-12, block_size: int = 0, full_ed_switch: int = FULL_ED_SWITCH,
                       precision="f64", which_lanczos: str = "GPU", neigen_arpack: int | None = None,
                       arpack_ncv: int | None = None, full_ed_green: bool = False, im: dict | None = None,
                       slater=None, on_block=None):
    """Lowest states of every sector (solver.h:98-237), filtered to [E0, E0 + dEmax].

    Per sector (solver.h:150-200): block Lanczos when block_size > 0, unless N_up or N_dn
    is 0 or 1 or both are Ns; otherwise dense ED at dimension <= full_ed_switch, else
    which_lanczos = 'NORMAL' / 'GPU' (the device Lanczos), 'FULL_ED' or 'ARPACK'.
    full_ed_green (FLAG_FULL_ED_GREEN): dense-ED sectors keep all their states
    (ED_ARPACK.f90:652-665) and neither window filter runs (solver.h:211,
    solver.f90:93), so compute_green's Lehmann sums over small sectors are complete.
    im (imaginary parts of Ec / Eb / Vbc, see make_sector): complex sectors, where dense
    ED and the plain device Lanczos run; block Lanczos and ARPACK are real only.  slater:
    the 4-index impurity interaction U(s1, s2, s3, s4) added to every sector.  on_block(sector,
    result) is called after each block Lanczos solve (its steps and full projections)."""
    sectors = all_sectors(model.ns) if sectors is None else sectors
    if im is not None and (block_size > 0 or which_lanczos == "ARPACK"):
        raise ValueError("complex sectors: block Lanczos and ARPACK are real only")
    states: list[EigenState] = []
    ns = model.ns
    for (nu, nd) in sectors:
        dim = comb(ns, nu) * comb(ns, nd)
        plain = block_size == 0 or nu in (0, 1) or nd in (0, 1) or (nu == ns and nd == ns)
        with make_sector(model, nu, nd, im, slater) as sec:
            if dim <= full_ed_switch or (plain and which_lanczos == "FULL_ED"):
                ev, vec = sec.full_ed(dim if full_ed_green else min(neigen, dim))
                found = [(ev[k], vec[k]) for k in range(len(ev))]
            elif not plain:
                from .block import block_lanczos
                ev, vec = block_lanczos(sec, nblock=block_size, neig=neigen, nitermax=nitermax, tolerance=tolerance,
                                        dEmax=None if np.isinf(dEmax) else dEmax)
                found = [(ev[k], vec[k]) for k in range(len(ev))]
                if on_block is not None:
                    on_block((nu, nd), block_lanczos.last)
            elif which_lanczos == "ARPACK":
                from .arpack import arpack_diago
                r = arpack_diago(sec, neigen_arpack or max(2 * neigen, 2) + 1, tolerance, dEmax, ncv=arpack_ncv)
                found = [(r.evals[k], r.vectors[k]) for k in range(r.nconv)]
            elif which_lanczos in ("GPU", "NORMAL"):
                res = sec.lanczos(nitermax=nitermax, neigen=neigen, tolerance=tolerance, precision=precision)
                found = [(res.evals[k], sec.ground_state(res, k)) for k in range(len(res.evals))]
            else:
                raise ValueError(f"which_lanczos {which_lanczos!r}: expected NORMAL, GPU, FULL_ED or ARPACK")
        if not found:
            continue
        e_min = min(e for e, _ in found)
        for e, v in found:   # filter_eigen(lowest, [dE, dE + dEmax]) per sector
            if full_ed_green or e <= e_min + dEmax:
                states.append(EigenState((nu, nd), float(e), v))
    if not states:
        return []
    if full_ed_green:
        return states
    e0 = min(s.energy for s in states)
    return [s for s in states if s.energy <= e0 + dEmax]


def sector_scan(model: AIMParams, beta: float, weight_cutoff: float = 1e-6, block_size: int = 8, neigen: int = 8,
                nitermax: int = 128, tolerance: float = 1e-12, full_ed_switch: int = FULL_ED_SWITCH,
                batched: bool = True):
    """Every (N_up, N_dn) sector with Boltzmann weight above weight_cutoff, neigen lowest
    states each (solver.h:98-237 over the whole Fock space).

    Pass 1: the lowest energy of every sector: dense ED at dimension <= full_ed_switch;
    the others' Lanczos_fast_diagonalize all at once when `batched` (min(Nitermax, dim - 1)
    fixed steps of every sector in one batched device call, ed_lanczos_batched, then the
    reference's convergence test replayed on each sector's coefficients,
    ed_lanczos_converged: the same iteration and Ritz value as the sector's own loop up to
    rounding), else sector by sector.  In the batched form a sector whose up spin has fewer
    than 64 states (the multi-run kernel's minimum) and whose down spin has more runs as the
    spin-reversed sector (N_dn, N_up) of the spin-reversed model: the same Hamiltonian up to
    a relabelling of the basis (the spins' strings are signed independently and every
    interaction term is symmetric under the exchange), so the same spectrum, reached from
    another start vector.  Pass 2: the sectors with e^{-beta (E_s - E0)} >
    weight_cutoff get `neigen` states by block Lanczos of width `block_size` (plain Lanczos
    in the sectors solver.h:152 excludes); states below the weight cutoff are dropped.
    Returns (states, info)."""
    import time
    from .hclass import lanczos_batched, lanczos_converged
    ns = model.ns
    t0 = time.perf_counter()
    e_low = {}
    lanczos_keys = []
    reversed_model = None
    for (nu, nd) in all_sectors(ns):
        dim = comb(ns, nu) * comb(ns, nd)
        if dim <= full_ed_switch:
            with Sector(model, nu, nd) as sec:
                e_low[(nu, nd)] = float(sec.full_ed(1)[0][0])
        elif batched:
            lanczos_keys.append((nu, nd))
        else:
            with Sector(model, nu, nd) as sec:
                e_low[(nu, nd)] = float(sec.lanczos(nitermax=nitermax, neigen=1, tolerance=tolerance).evals[0])
    # batched pass 1, grouped by run length Niter = min(Nitermax, dimen - 1)
    groups: dict[int, list] = {}
    for key in lanczos_keys:
        groups.setdefault(min(nitermax, comb(ns, key[0]) * comb(ns, key[1]) - 1), []).append(key)
    def pass1_sector(k):
        nonlocal reversed_model
        nu, nd = k
        if comb(ns, nu) < _RUNS_MIN_DUP <= comb(ns, nd):
            if reversed_model is None:
                reversed_model = spin_reversed(model)
            return Sector(reversed_model, nd, nu)
        return Sector(model, nu, nd)

    t_parts = {"dense_s": time.perf_counter() - t0, "create_s": 0.0, "batched_s": 0.0, "release_s": 0.0,
               "converged_s": 0.0}
    for niter, keys in groups.items():
        t1 = time.perf_counter()
        secs = [pass1_sector(k) for k in keys]
        t2 = time.perf_counter()
        try:
            a, b, steps = lanczos_batched(secs, niter)
        finally:
            t3 = time.perf_counter()
            for sec in secs:
                sec.close()
        t4 = time.perf_counter()
        for i, key in enumerate(keys):
            _, ev = lanczos_converged(a[i, :steps[i]], b[i, :steps[i]], 1, tolerance, nitermax)
            e_low[key] = float(ev[0])
        for name, dt in (("create_s", t2 - t1), ("batched_s", t3 - t2), ("release_s", t4 - t3),
                         ("converged_s", time.perf_counter() - t4)):
            t_parts[name] += dt
    t_pass1 = time.perf_counter() - t0
    e0 = min(e_low.values())
    kept = sorted(s for s, e in e_low.items() if np.exp(-beta * (e - e0)) > weight_cutoff)
    t0 = time.perf_counter()
    steps, fulls, ms, ritz = [], [], [], []

    def count(_sector, result):
        steps.append(result.steps)
        fulls.append(result.full_passes)
        ms.append(result.ms)
        ritz.append(result.ritz_ms)

    # a state is kept when e^{-beta (E - E0)} > weight_cutoff, i.e. E - E0 < ln(1 / weight_cutoff) / beta:
    # that window, relative to each sector's lowest state and to E0, is get_eigen_energies'
    # dEmax, so only those states' vectors are copied back from the device
    states = get_eigen_energies(model, sectors=kept, neigen=neigen, block_size=block_size, nitermax=nitermax,
                                tolerance=tolerance, full_ed_switch=full_ed_switch, on_block=count,
                                dEmax=max(0.0, float(np.log(1.0 / weight_cutoff) / beta)))
    t_pass2 = time.perf_counter() - t0
    states = [st for st in states if np.exp(-beta * (st.energy - e0)) > weight_cutoff]
    info = {"sectors": len(e_low), "kept_sectors": kept, "states": len(states), "E0": e0,
            "t_pass1_s": t_pass1, "t_pass1_parts": t_parts, "t_pass2_s": t_pass2, "e_low": e_low, "block_steps": steps,
            "block_full_passes": fulls, "block_ms": ms, "block_ritz_ms": ritz}
    return states, info


def partition(states: list[EigenState], beta: float) -> np.ndarray:
    """Boltzmann weights e^{-beta (E - E0)} / Z of the kept states."""
    e = np.array([s.energy for s in states])
    w = np.exp(-beta * (e - e.min()))
    return w / w.sum()


def _excite(src: Sector, dst: Sector, op, pm: int, psi: np.ndarray) -> np.ndarray:
    """A+|psi> (pm = +1) or A|psi> (pm = -1) for A = sum coef c_(site, spin) over the
    entries (site, spin, coef[, correlated]) of op: the creation side carries conj(coef); an
    entry marked correlated takes the excitation of apply_Csz_N_updo (apply_C_n)."""
    phi = None
    for c, sp, coef, *corr in op:
        fn = apply_C_n if corr and corr[0] else apply_C
        t = (np.conj(coef) if pm > 0 else coef) * fn(src, dst, c, sp, pm, psi)
        phi = t if phi is None else phi + t
    return phi


def _green_ops(states: list[EigenState], beta: float, wn: np.ndarray, ops, nitergreen: int, weight_cutoff: float,
               cutoff_dynamic: float, precision, full_ed_switch: int, make, target, device_terms) -> np.ndarray:
    """sum over states of boltz * (particle + hole) for each operator A = sum coef c_(site, spin)
    of `ops` (a list of [(site, spin, coef), ...]): the particle run from A+|psi>, the hole
    run from A|psi> (green_class_computeAA.h:101-302 with the vectors of
    symmetric_combineAA, green_class_compute_symmetric_combination.f90:39-137).
    make(key) opens the sector of a state's key, target(key, spin, pm) gives the key of
    the sector A+ (pm = +1) or A (pm = -1) leads to, or None.  Small target sectors take
    the Lehmann sum over the held states, every other run is an excitation Lanczos.  With
    device_terms (real sectors: op -> (spin, [(site index, coef), ...]) of
    ed_dynamic_from_state) a state's excitation vectors are built on the device from one
    copy of the state, and the runs of all states advance together in one batched device
    call (ed_dynamic_from_states); otherwise (complex sectors) apply_C builds them per
    operator term and a state's runs advance together."""
    w = partition(states, beta)
    G = np.zeros((len(ops), len(wn)), complex)
    sectors: dict[tuple, Sector] = {}

    def sector(key):
        if key not in sectors:
            sectors[key] = make(key)
        return sectors[key]

    held: dict[tuple, tuple[np.ndarray, np.ndarray]] = {}

    def held_states(key):
        if key not in held:
            sel = [x for x in states if x.sector == key]
            held[key] = (np.array([x.energy for x in sel]), np.array([x.vector for x in sel]).reshape(len(sel), -1))
        return held[key]

    dev_states, dev_jobs, dev_meta = [], [], []   # the device-built runs of every state
    dev_spin = 0
    try:
        for st, wt in zip(states, w):
            if wt <= weight_cutoff:
                continue
            src = sector(st.sector)
            jobs = []   # (operator index, target sector, iph, operator) of the excitation runs
            for io, op in enumerate(ops):
                for iph, pm in ((1, +1), (2, -1)):
                    tkey = target(st.sector, op[0][1], pm)
                    if tkey is None:
                        continue
                    dst = sector(tkey)
                    if dst.dim <= full_ed_switch:
                        phi = _excite(src, dst, op, pm, st.vector)
                        ev, vec = held_states(tkey)
                        G[io] += wt * dynamic_full_ed(iph, st.energy, phi, ev, vec, wn, cutoff_dynamic)
                    else:
                        jobs.append((io, dst, iph, op))
            if jobs and device_terms is not None:
                # the state's vector goes to the device once and its excitation vectors are
                # built there; the runs of every state go in one call below
                q = len(dev_states)
                dev_states.append((src, st.vector, st.energy))
                for io, dst, iph, op in jobs:
                    dev_spin, terms = device_terms(op)
                    dev_jobs.append((q, dst, iph, terms))
                    dev_meta.append((io, wt))
            elif jobs:
                phis = [_excite(src, dst, op, +1 if iph == 1 else -1, st.vector) for _, dst, iph, op in jobs]
                dyn, _, _ = compute_dynamic_batched([j[1] for j in jobs], phis, [j[2] for j in jobs],
                                                    [st.energy] * len(jobs), nitergreen, wn, cutoff_dynamic,
                                                    precision)
                for k, (io, *_rest) in enumerate(jobs):
                    G[io] += wt * dyn[k]
        if dev_jobs:
            # every excitation run of every state advances together (ed_dynamic_from_states)
            dyn = dynamic_from_states(dev_states, dev_jobs, dev_spin, nitergreen, wn, cutoff_dynamic, precision)
            for k, (io, wt) in enumerate(dev_meta):
                G[io] += wt * dyn[k]
    finally:
        for s in sectors.values():
            s.close()
    return G


def _updo_ops(model: AIMParams, spin: int, im: dict | None, slater=None):
    """make / target / device_terms of _green_ops for the (N_up, N_dn) sectors."""
    def target(key, sp, pm):
        tn = (key[0] + pm, key[1]) if sp == 0 else (key[0], key[1] + pm)
        return tn if 0 <= tn[0] <= model.ns and 0 <= tn[1] <= model.ns else None
    terms = (lambda op: (op[0][1], [(c, coef) for c, _, coef in op])) if im is None else None
    return (lambda key: make_sector(model, key[0], key[1], im, slater)), target, terms


def compute_green(model: AIMParams, states: list[EigenState], beta: float, wn: np.ndarray, nitergreen: int = 36,
                  spin: int = 0, orbitals=None, weight_cutoff: float = 0.0, cutoff_dynamic: float = 1e-8,
                  precision="f64", full_ed_switch: int = FULL_ED_SWITCH, im: dict | None = None,
                  slater=None) -> np.ndarray:
    """Diagonal impurity G_cc(i w_n), one row per orbital c (green_class_computeAA.h:101-302).

    For every state with Boltzmann weight above weight_cutoff and every orbital: the
    particle run from c+|psi> in (N+1) and the hole run from c|psi> in (N-1).  As
    compute_dynamic dispatches (green_class_compute_dynamic.f90:44-62): a target sector
    of dimension <= full_ed_switch takes the Lehmann sum over the states `states` holds
    for it (compute_dynamic_full_ed, :77-165; none held gives zero), every other run is
    an excitation Lanczos, and those all advance together in one batched device call.
    im / slater: the states are of that model (get_eigen_energies with the same im / slater)."""
    orbitals = list(range(model.nc)) if orbitals is None else list(orbitals)
    return _green_ops(states, beta, wn, [[(c, spin, 1.0)] for c in orbitals], nitergreen, weight_cutoff,
                      cutoff_dynamic, precision, full_ed_switch, *_updo_ops(model, spin, im, slater))


def compute_green_matrix(model: AIMParams, states: list[EigenState], beta: float, wn: np.ndarray,
                         nitergreen: int = 36, spin: int = 0, orbitals=None, weight_cutoff: float = 0.0,
                         cutoff_dynamic: float = 1e-8, precision="f64",
                         full_ed_switch: int = FULL_ED_SWITCH, im: dict | None = None,
                         slater=None) -> np.ndarray:
    """Impurity G_ij(i w_n) = <<c_i ; c_j+>> of one spin, all orbital pairs, shape (norb, norb, nw).

    symmetric_combineAA (green_class_compute_symmetric_combination.f90:39-137) and
    reshuffle_vecAA (green_class_compute_symmetric_combination_AA.h:10-57): besides the
    diagonal runs from c_i, one run per pair i < j from A = c_i + c_j (particle run from
    A+|psi>, hole run from A|psi>), whose G_A = G_ii + G_jj + G_ij + G_ji.  Real H:
    G_ij = G_ji = (G_A - G_ii - G_jj) / 2.  Complex H (im given, states of the complex
    model): one more run per pair from B = c_i - i c_j, G_B = G_ii + G_jj + i (G_ij - G_ji),
    so G_ij, G_ji = (S +- D) / 2 with S = G_A - G_ii - G_jj, D = -i (G_B - G_ii - G_jj).
    Every run of every pair shares the one batched device call per state."""
    orbitals = list(range(model.nc)) if orbitals is None else list(orbitals)
    n = len(orbitals)
    pairs = [(a, b) for a in range(n) for b in range(a + 1, n)]
    ops = [[(c, spin, 1.0)] for c in orbitals] + [[(orbitals[a], spin, 1.0), (orbitals[b], spin, 1.0)]
                                                  for a, b in pairs]
    if im is not None:
        ops += [[(orbitals[a], spin, 1.0), (orbitals[b], spin, -1j)] for a, b in pairs]
    g = _green_ops(states, beta, wn, ops, nitergreen, weight_cutoff, cutoff_dynamic, precision, full_ed_switch,
                   *_updo_ops(model, spin, im, slater))
    G = np.zeros((n, n, len(wn)), complex)
    for a in range(n):
        G[a, a] = g[a]
    for k, (a, b) in enumerate(pairs):
        S = g[n + k] - g[a] - g[b]
        if im is None:
            G[a, b] = G[b, a] = 0.5 * S
        else:
            D = -1j * (g[n + len(pairs) + k] - g[a] - g[b])
            G[a, b], G[b, a] = 0.5 * (S + D), 0.5 * (S - D)
    return G


# --------------------------------------------------------------------------------
# S_z (Nambu) basis: superconducting baths (HAIMsz; apply_Csz; compute_greenAB with Cup / Cdo)

def _sz_complex(pb, pvbc, im) -> bool:
    return im is not None or np.iscomplexobj(pb) or np.iscomplexobj(pvbc)


def make_sz_sector(model: AIMParams, npart: int, pb=None, pvbc=None, pairing_imp_to_bath: bool = False,
                   im: dict | None = None):
    """new_H's sz branch: the real S_z sector, or the complex one (COMPLEX(8) HAIMsz_multc)
    when im (imaginary parts of Ec, Eb, Vbc) is given or the pairing is complex."""
    from .hclass import ComplexSzSector, SzSector
    if _sz_complex(pb, pvbc, im):
        nc, nb = model.nc, model.nb
        im = im if im is not None else {"ec": np.zeros((2, nc, nc))}
        return ComplexSzSector(model, npart, im, pb, pvbc, pairing_imp_to_bath)
    return SzSector(model, npart, pb, pvbc, pairing_imp_to_bath)


def get_eigen_energies_sz(model: AIMParams, pb=None, pvbc=None, pairing_imp_to_bath: bool = False, sectors=None,
                          neigen: int = 1, dEmax: float = np.inf, nitermax: int = 128, tolerance: float = 1e-12,
                          full_ed_switch: int = FULL_ED_SWITCH, precision="f64", full_ed_green: bool = False,
                          im: dict | None = None):
    """Lowest states of the S_z sectors npart = 0 .. 2 Ns (or `sectors`), filtered to
    [E0, E0 + dEmax] as get_eigen_energies (solver.h:98-237 with the sz branch of new_H):
    dense ED at dimension <= full_ed_switch, else the device Lanczos.  States carry the
    key (npart,) and the eigenvalue of the Nambu H; the physical energy adds the
    sector's energy_shift (the same for every npart, so weights and G do not change).
    A complex model (im, or complex pb / pvbc) runs the complex S_z sectors; its states
    carry complex vectors."""
    sectors = [(n,) for n in range(2 * model.ns + 1)] if sectors is None else [tuple(np.atleast_1d(k)) for k in sectors]
    states: list[EigenState] = []
    for key in sectors:
        with make_sz_sector(model, key[0], pb, pvbc, pairing_imp_to_bath, im) as sec:
            if sec.dim <= full_ed_switch:
                ev, vec = sec.full_ed(sec.dim if full_ed_green else min(neigen, sec.dim))
                found = [(ev[k], vec[k]) for k in range(len(ev))]
            else:
                res = sec.lanczos(nitermax=nitermax, neigen=neigen, tolerance=tolerance, precision=precision)
                found = [(res.evals[k], sec.ground_state(res, k)) for k in range(len(res.evals))]
        if not found:
            continue
        e_min = min(e for e, _ in found)
        states += [EigenState(key, float(e), v) for e, v in found if full_ed_green or e <= e_min + dEmax]
    if not states or full_ed_green:
        return states
    e0 = min(s.energy for s in states)
    return [s for s in states if s.energy <= e0 + dEmax]


def compute_green_nambu(model: AIMParams, states: list[EigenState], beta: float, wn: np.ndarray, pb=None,
                        pvbc=None, pairing_imp_to_bath: bool = False, nitergreen: int = 36, orbitals=None,
                        weight_cutoff: float = 0.0, cutoff_dynamic: float = 1e-8, precision="f64",
                        full_ed_switch: int = FULL_ED_SWITCH, im: dict | None = None) -> np.ndarray:
    """Nambu impurity Green's function G^N_ij(i w_n) = <<d_i ; d_j+>>, shape (2n, 2n, nw), over
    the Nambu orbitals d_(c, 0) = c_(c, up) (index c) and d_(c, 1) = c+_(c, dn) (index n + c)
    of the S_z sectors (states from get_eigen_energies_sz with the same pairing):
    * G^N[c, c'] = G_up,cc'; G^N[n + c, n + c'] = <<c_dn+ ; c_dn>> = -conj(G_dn,cc'(i w_n));
    * G^N[c, n + c'] = <<c_up ; c_dn>>, the anomalous F (compute_greenAB with apply_Cup /
      apply_Cdo, green_class_computeAB.f90:23-197; correlations_dyn_stat.h:115).
    Every d_i+ raises npart by one (apply_Csz, apply_C.f90:277-327: the Nambu orbital of
    the down spin is created by the physical c_dn), so the off-diagonal elements come
    from the symmetric combinations (d_i + d_j) of the real H as in compute_green_matrix.
    A state's excitation vectors are built on the device (ed_dynamic_from_state with Nambu
    orbital indices) and its runs batched.  A complex model (im, or complex pb / pvbc):
    one more run per pair from d_i - i d_j, and G_ij, G_ji = (S +- D) / 2 as in
    compute_green_matrix (the _complex branch of symmetric_combineAA); its excitation
    vectors are built per term by apply_C and the runs batched."""
    orbitals = list(range(model.nc)) if orbitals is None else list(orbitals)
    nambu = [(c, 0) for c in orbitals] + [(c, 1) for c in orbitals]
    n = len(nambu)
    pairs = [(a, b) for a in range(n) for b in range(a + 1, n)]
    ops = [[(c, sp, 1.0)] for c, sp in nambu] + [[(*nambu[a], 1.0), (*nambu[b], 1.0)] for a, b in pairs]
    cx = _sz_complex(pb, pvbc, im)
    if cx:
        ops += [[(*nambu[a], 1.0), (*nambu[b], -1j)] for a, b in pairs]
    g = _green_ops(states, beta, wn, ops, nitergreen, weight_cutoff, cutoff_dynamic, precision, full_ed_switch,
                   *_sz_ops(model, pb, pvbc, pairing_imp_to_bath, im))
    G = np.zeros((n, n, len(wn)), complex)
    for a in range(n):
        G[a, a] = g[a]
    for k, (a, b) in enumerate(pairs):
        S = g[n + k] - g[a] - g[b]
        if not cx:
            G[a, b] = G[b, a] = 0.5 * S
        else:
            D = -1j * (g[n + len(pairs) + k] - g[a] - g[b])
            G[a, b], G[b, a] = 0.5 * (S + D), 0.5 * (S - D)
    return G


def _sz_ops(model: AIMParams, pb, pvbc, pairing_imp_to_bath, im=None):
    """make / target / device_terms of _green_ops for the S_z sectors: every Nambu operator
    moves npart by one; device terms are Nambu orbitals spin * nc + site, with SITE_NCOND
    on the correlated ones (real sectors; a complex model builds its vectors per term)."""
    from .hclass import SITE_NCOND

    def target(key, sp, pm):
        t = key[0] + pm
        return (t,) if 0 <= t <= 2 * model.ns else None

    nc = model.nc

    def terms(op):
        return 0, [((sp * nc + c) | (SITE_NCOND if corr and corr[0] else 0), coef) for c, sp, coef, *corr in op]

    make = lambda key: make_sz_sector(model, key[0], pb, pvbc, pairing_imp_to_bath, im)
    return make, target, (None if _sz_complex(pb, pvbc, im) else terms)


def compute_green_ncup(model: AIMParams, states: list[EigenState], beta: float, wn: np.ndarray, pb=None,
                       pvbc=None, pairing_imp_to_bath: bool = False, nitergreen: int = 36, orbitals=None,
                       weight_cutoff: float = 0.0, cutoff_dynamic: float = 1e-8, precision="f64",
                       full_ed_switch: int = FULL_ED_SWITCH, im: dict | None = None):
    """The correlated-hopping Green's functions of FLAG_NCUP (correlations_dyn_stat.h:69-100)
    on the S_z sectors, per Nambu impurity orbital a = (c, s) (index s n + c over `orbitals`):
    * GN[a] = <<D_a ; D_a+>>: compute_greenAA with apply_N_Cup / apply_N_Cdo, the operator
      D_(c, 0) = n_dn c_up and D_(c, 1) = n_up c+_dn (apply_Csz_N_updo, apply_C.f90:404-483);
    * GNF[0, a] = <<D_a ; d_a+>> and GNF[1, a] = <<d_a ; D_a+>>: compute_greenAB with
      apply_N_Cup and apply_Cup (green_class_computeAB.f90:23-197; its greenAB and greenBA),
      d_(c, 0) = c_up, d_(c, 1) = c+_dn.  A real H has <<D ; d+>> = <<d ; D+>> = (G_(D+d) -
      GN - G) / 2 from the symmetric combination D + d, as the reference's real build; a
      complex model (im, or complex pairing) adds the runs of D - i d and splits the two
      as compute_green_matrix does.
    Returns (GN, GNF, G): GN and G of shape (2n, nw), GNF of shape (2, 2n, nw); G is the
    diagonal Nambu G^N of the same runs.  States from get_eigen_energies_sz with the same
    model; a real state's runs are built on the device and batched (ed_dynamic_from_state
    with ED_SITE_NCOND terms)."""
    orbitals = list(range(model.nc)) if orbitals is None else list(orbitals)
    nambu = [(c, 0) for c in orbitals] + [(c, 1) for c in orbitals]
    n = len(nambu)
    ops = ([[(c, sp, 1.0, True)] for c, sp in nambu] + [[(c, sp, 1.0)] for c, sp in nambu]
           + [[(c, sp, 1.0, True), (c, sp, 1.0)] for c, sp in nambu])
    cx = _sz_complex(pb, pvbc, im)
    if cx:
        ops += [[(c, sp, 1.0, True), (c, sp, -1j)] for c, sp in nambu]
    g = _green_ops(states, beta, wn, ops, nitergreen, weight_cutoff, cutoff_dynamic, precision, full_ed_switch,
                   *_sz_ops(model, pb, pvbc, pairing_imp_to_bath, im))
    GN, G = g[:n], g[n:2 * n]
    S = g[2 * n:3 * n] - GN - G
    if not cx:
        GNF = np.stack([0.5 * S, 0.5 * S])
    else:
        D = -1j * (g[3 * n:] - GN - G)
        GNF = np.stack([0.5 * (S + D), 0.5 * (S - D)])
    return GN, GNF, G
