"""Parity of the HIP path (through the C-ABI) against the oracle, on the GPU.

Bars: integer tables / states / ranks / apply_C bit-exact; FP64 H.v within 1e-13
relative (max |diff| / max |w|); E0 within 1e-10 (the FP64 tolerance of every E0 test here);
G(i w_n) within 1e-10 relative; FP32 H.v within 1e-5 relative.
"""
import json
import os
import sys

import numpy as np
import pytest

import oracle
import oracle.brute as bf
from ed_solver_amd.aim import AIMParams, kanamori, synthetic

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "models.json")


def golden(name):
    with open(GOLD) as f:
        d = json.load(f)[name]
    return AIMParams.from_json(d["model"]), tuple(d["sector"])


def small_models():
    rng = np.random.default_rng(7)
    out = []
    for (nc, nb, U, J, sec) in [(1, 5, 2.0, 0.0, (3, 3)), (2, 3, 3.0, 0.5, (2, 3)), (3, 3, 2.5, 0.4, (3, 3)),
                                (3, 0, 2.5, 0.4, (1, 1)), (2, 4, 3.0, 0.6, (3, 2)), (1, 1, 2.0, 0.0, (1, 1)),
                                (3, 4, 2.5, 0.4, (0, 3)), (2, 5, 3.0, 0.5, (7, 1))]:
        m = kanamori(nc, nb, U, J, rng.uniform(-1, 1, nb), rng.uniform(0.2, 0.6, nb))
        out.append((m, sec))
    # a model with impurity and bath hoppings (non-diagonal Ec / Eb) and spin-dependent levels
    nc, nb = 2, 4
    ec = np.array([[[-1.0, 0.3], [0.3, -0.8]], [[-1.1, 0.3], [0.3, -0.7]]])
    eb = np.stack([np.diag(rng.uniform(-1, 1, nb)) + 0.2 * (np.eye(nb, k=1) + np.eye(nb, k=-1))] * 2)
    vbc = rng.uniform(-0.5, 0.5, (2, nb, nc))
    u = np.array([[3.0, 2.0], [2.0, 3.0]]); jj = np.full((2, 2), 0.5)
    out.append((AIMParams(nc, nb, ec, eb, vbc, u, jj, 0.5, 1e-8), (3, 2)))
    return out


def rel(a, b):
    return np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b)))


def test_start_vector_device_matches_oracle(gpu):
    from ed_solver_amd import start_vector
    assert np.array_equal(start_vector(23456), oracle.start_vector(23456))


@pytest.mark.parametrize("idx", range(9))
def test_tables_bit_exact_small(gpu, idx):
    m, (nu, nd) = small_models()[idx]
    s = gpu.Sector(m, nu, nd)
    o = oracle.OracleSector(m, nu, nd)
    for spin in (0, 1):
        a, b = s.tables(spin), o.tables(spin)
        for key in ("states", "noff", "rankoff"):
            assert np.array_equal(a[key], b[key]), (idx, spin, key)
        for key in ("diag", "offdiag"):
            assert np.array_equal(a[key].view(np.int64), b[key].view(np.int64)), (idx, spin, key)
    s.close()


@pytest.mark.parametrize("name", ["A", "B", "C", "E"])
def test_tables_bit_exact_workloads(gpu, name):
    m, (nu, nd) = golden(name)
    s = gpu.Sector(m, nu, nd)
    o = oracle.OracleSector(m, nu, nd)
    assert s.dim == o.dim
    for spin in (0, 1):
        a, b = s.tables(spin), o.tables(spin)
        assert np.array_equal(a["states"], b["states"])
        assert np.array_equal(a["noff"], b["noff"])
        assert np.array_equal(a["rankoff"], b["rankoff"])
        assert np.array_equal(a["diag"].view(np.int64), b["diag"].view(np.int64))
        assert np.array_equal(a["offdiag"].view(np.int64), b["offdiag"].view(np.int64))
    s.close()


@pytest.mark.parametrize("idx", range(9))
def test_hmult_small_vs_brute_and_oracle(gpu, idx):
    m, (nu, nd) = small_models()[idx]
    s = gpu.Sector(m, nu, nd)
    o = oracle.OracleSector(m, nu, nd)
    H = bf.sector_hamiltonian(m, nu, nd)
    rng = np.random.default_rng(idx)
    v = rng.uniform(-1, 1, s.dim)
    w = s.hmult(v)
    assert rel(w, o.hmult(v)) < 1e-13
    assert rel(w, H @ v) < 1e-13
    s.close()


@pytest.mark.parametrize("name", ["A", "B", "C"])
def test_hmult_workloads_fp64(gpu, name):
    m, (nu, nd) = golden(name)
    s = gpu.Sector(m, nu, nd)
    o = oracle.OracleSector(m, nu, nd, nthreads=8)
    v = np.random.default_rng(3).uniform(-1, 1, s.dim)
    assert rel(s.hmult(v), o.hmult(v)) < 1e-13
    s.close()


def test_shared_spin_tables(gpu):
    """Sectors of one model share the tables of each (spin, N): (5,5), (5,4) and (4,5) of
    workload B hold the N = 5 tables of both spins, and a second model with other bath
    levels gets its own.  Each H.v matches the oracle (1e-13) while the others are alive,
    and after the sectors that built the shared tables are released."""
    m, _ = golden("B")
    m2, _ = golden("B")
    m2.eb = np.array(m2.eb, dtype=np.float64)
    m2.eb[:, 0, 0] += 0.25   # another model: its own tables
    keys = [(5, 5), (5, 4), (4, 5)]
    secs = {k: gpu.Sector(m, *k) for k in keys}
    other = gpu.Sector(m2, 5, 5)
    rng = np.random.default_rng(11)

    def check(sec, mm, k):
        o = oracle.OracleSector(mm, *k, nthreads=8)
        v = rng.uniform(-1, 1, sec.dim)
        assert rel(sec.hmult(v), o.hmult(v)) < 1e-13, k
        o.close()

    for k in keys:
        check(secs[k], m, k)
    check(other, m2, (5, 5))
    secs.pop((5, 5)).close()   # the first builder of both N = 5 tables
    check(secs[(5, 4)], m, (5, 4))
    check(secs[(4, 5)], m, (4, 5))
    again = gpu.Sector(m, 5, 5)   # from the tables still held
    check(again, m, (5, 5))
    for x in list(secs.values()) + [again, other]:
        x.close()


@pytest.mark.parametrize("name", ["B", "C", "E"])
def test_hmult_every_schedule(gpu, name):
    """Every H.v schedule the autotuner can pick gives the oracle's H.v (FP64 1e-13)."""
    m, (nu, nd) = golden(name)
    s = gpu.Sector(m, nu, nd)
    o = oracle.OracleSector(m, nu, nd, nthreads=16)
    v = np.random.default_rng(5).uniform(-1, 1, s.dim)
    ref = o.hmult(v)
    choice, ms = s.tune()
    print(f"{name}: tuned schedule {choice}, ms per schedule {ms}", flush=True)
    for c in (1, 2, 4, 8, 16, 18, 19, 20):
        try:
            s.set_schedule(c)
        except gpu.EDError:
            assert c >= 8
            continue
        assert rel(s.hmult(v), ref) < 1e-13, c
    s.close()


@pytest.mark.parametrize("sector", [(6, 5), (6, 9), (5, 6)])
def test_hmult_and_lanczos_unequal_table_widths(gpu, sector):
    """Sectors of workload C whose down table is wider than the up table ((6,5), (6,9)) or
    narrower ((5,6)): H.v on every schedule and the Lanczos ground state against the
    oracle (the fast kernel unrolls to the wider width and must not read past the up table)."""
    m, _ = golden("C")
    s = gpu.Sector(m, *sector)
    o = oracle.OracleSector(m, *sector, nthreads=16)
    print(f"sector {sector} dim {s.dim}", flush=True)
    v = np.random.default_rng(6).uniform(-1, 1, s.dim)
    ref = o.hmult(v)
    s.tune()
    for c in (1, 2, 4, 8, 16, 18, 19, 20):
        try:
            s.set_schedule(c)
        except gpu.EDError:
            assert c >= 8
            continue
        assert rel(s.hmult(v), ref) < 1e-13, c
        r = s.lanczos(nitermax=128, neigen=1, tolerance=1e-12)
        ro = o.lanczos(nitermax=128, neigen=1, tol=1e-12)
        assert abs(r.evals[0] - ro["evals"][0]) < 1e-10, c
    s.close()


def test_hmult_fp32_and_block(gpu):
    from hipmem import DevBuf
    m, (nu, nd) = golden("B")
    s = gpu.Sector(m, nu, nd)
    o = oracle.OracleSector(m, nu, nd, nthreads=8)
    v = np.random.default_rng(4).uniform(-1, 1, s.dim)
    ref = o.hmult(v)
    x32 = DevBuf.from_array(v.astype(np.float32))
    y32 = DevBuf(4 * s.dim)
    s.hmult_dev(y32.addr, x32.addr, "f32")
    s.synchronize()
    assert rel(y32.to_array(np.float32, s.dim).astype(np.float64), ref) < 1e-5
    # block H.v on 3 vectors with a padded leading dimension
    ld = s.dim + 37
    vs = [np.random.default_rng(10 + k).uniform(-1, 1, s.dim) for k in range(3)]
    Xh = np.zeros(3 * ld)
    for k in range(3):
        Xh[k * ld:k * ld + s.dim] = vs[k]
    X = DevBuf.from_array(Xh)
    Y = DevBuf(8 * 3 * ld)
    s.hmult_block_dev(Y.addr, X.addr, 3, ld)
    s.synchronize()
    Yh = Y.to_array(np.float64, 3 * ld)
    for k in range(3):
        assert rel(Yh[k * ld:k * ld + s.dim], o.hmult(vs[k])) < 1e-13
    for b in (x32, y32, X, Y):
        b.free()
    s.close()


@pytest.mark.parametrize("name", ["A", "B"])
def test_lanczos_alpha_beta_and_e0(gpu, name):
    m, (nu, nd) = golden(name)
    s = gpu.Sector(m, nu, nd)
    o = oracle.OracleSector(m, nu, nd, nthreads=8)
    r = s.lanczos(nitermax=128, neigen=1, tolerance=1e-12)
    ro = o.lanczos(nitermax=128, neigen=1, tol=1e-12)
    # same start vector -> same recurrence; early coefficients agree to round-off
    n = min(20, r.iters, ro["iters"])
    assert np.max(np.abs(r.diag[:n] - ro["alpha"][:n])) < 1e-10
    assert np.max(np.abs(r.subdiag[:n] - ro["beta"][:n])) < 1e-10
    assert abs(r.evals[0] - ro["evals"][0]) < 1e-10
    assert abs(r.iters - ro["iters"]) <= 1
    if s.dim <= 5000:
        e = np.linalg.eigvalsh(bf.sector_hamiltonian(m, nu, nd).toarray())
        assert abs(r.evals[0] - e[0]) < 1e-10
    s.close()


def test_lanczos_e0_and_gs_workload_c(gpu):
    import scipy.sparse.linalg as sla
    m, (nu, nd) = golden("C")
    s = gpu.Sector(m, nu, nd)
    o = oracle.OracleSector(m, nu, nd, nthreads=16)
    r = s.lanczos(nitermax=128, neigen=2, tolerance=1e-12)
    ro = o.lanczos(nitermax=128, neigen=2, tol=1e-12)
    assert abs(r.evals[0] - ro["evals"][0]) < 1e-10
    psi = s.ground_state(r, 0)
    assert abs(np.linalg.norm(psi) - 1) < 1e-12
    # residual of the Ritz pair with the oracle H.v
    hp = o.hmult(psi)
    e_rq = psi @ hp
    assert abs(e_rq - r.evals[0]) < 1e-9
    assert np.linalg.norm(hp - e_rq * psi) < 1e-4
    psi_o = o.ground_state(ro["iters"], ro["vecp"][0])
    assert abs(abs(psi @ psi_o) - 1) < 1e-9
    s.close()


@pytest.mark.parametrize("sched", [8, 16])
def test_lanczos_schedules_workload_c(gpu, sched):
    """The fused Lanczos step with the fast H.v in natural (8) and XCD-grouped (16)
    column order reproduces the oracle recurrence: 24 fixed steps, alpha/beta to 1e-10."""
    m, (nu, nd) = golden("C")
    s = gpu.Sector(m, nu, nd)
    o = oracle.OracleSector(m, nu, nd, nthreads=16)
    s.set_schedule(sched)
    r = s.lanczos(nitermax=24, neigen=1, tolerance=0.0)
    ro = o.lanczos(nitermax=24, neigen=1, tol=0.0)
    assert r.iters == ro["iters"] == 24
    assert np.max(np.abs(r.diag - ro["alpha"][:24])) < 1e-10
    assert np.max(np.abs(r.subdiag - ro["beta"][:24])) < 1e-10
    s.close()


@pytest.mark.parametrize("sched", [8, 16])
def test_lanczos_large_columns_workload_e(gpu, sched):
    """Workload E (d_up = 3432, more rows than a workgroup has threads: each thread loops
    over its rows) in natural (8) and XCD-grouped (16) column order: the Lanczos
    recurrence (16 fixed two-kernel steps) against the oracle's, alpha/beta to 1e-10
    relative in FP64 and 1e-5 with FP32 vectors; the two orders agree to 1e-12."""
    m, (nu, nd) = golden("E")
    s = gpu.Sector(m, nu, nd)
    try:
        s.set_schedule(sched)
    except gpu.EDError:
        pytest.skip(f"schedule {sched} does not apply to workload E")
    o = oracle.OracleSector(m, nu, nd, nthreads=16)
    print(f"E schedule {sched}: dim {s.dim}", flush=True)
    a, b = s.lanczos_run(16)
    ro = o.lanczos(nitermax=16, neigen=1, tol=0.0)
    assert ro["iters"] == 16
    # relative: beta_1 = ||u_1|| is about 2e3 here (a sum of 11.8M squares in another order)
    assert np.max(np.abs(a - ro["alpha"][:16]) / np.maximum(1.0, np.abs(ro["alpha"][:16]))) < 1e-10
    assert np.max(np.abs(b - ro["beta"][:16]) / np.maximum(1.0, np.abs(ro["beta"][:16]))) < 1e-10
    a32, b32 = s.lanczos_run(16, precision="f32")
    assert np.max(np.abs(a32 - ro["alpha"][:16])) < 1e-5 * np.max(np.abs(ro["alpha"][:16]))
    assert np.max(np.abs(b32 - ro["beta"][:16])) < 1e-5 * np.max(np.abs(ro["beta"][:16]))
    s.set_schedule(24 - sched)
    a8, _ = s.lanczos_run(16)
    assert np.max(np.abs(a8 - a)) < 1e-12
    s.close()


def test_lanczos_reorth_b(gpu):
    m, (nu, nd) = golden("B")
    s = gpu.Sector(m, nu, nd)
    o = oracle.OracleSector(m, nu, nd, nthreads=8)
    r = s.lanczos(nitermax=200, neigen=3, tolerance=0.0, reorth=True)   # fixed 200 iterations
    assert r.iters == 200
    ro = o.lanczos(nitermax=128, neigen=3, tol=1e-12)
    assert abs(r.evals[0] - ro["evals"][0]) < 1e-10
    s.close()


@pytest.mark.parametrize("name", ["B", "C"])
def test_mixed_precision_e0(gpu, name):
    """FP32-vector Lanczos + FP64 restart: |E0 - E0_oracle(FP64)| < 1e-8, and the FP32
    phase alone within 1e-4."""
    m, (nu, nd) = golden(name)
    s = gpu.Sector(m, nu, nd)
    o = oracle.OracleSector(m, nu, nd, nthreads=16)
    ro = o.lanczos(nitermax=128, neigen=1, tol=1e-12)
    r = s.lanczos(nitermax=128, neigen=1, tolerance=1e-12, precision="f32")
    assert r.precision == "mixed" and r.iters_f32 > 0
    assert abs(r.evals[0] - ro["evals"][0]) < 1e-8
    psi = s.ground_state(r, 0)
    assert abs(s.rayleigh(psi) - ro["evals"][0]) < 1e-8
    s.close()


def test_hclass_global_surface(gpu):
    from ed_solver_amd import Hmult__, delete_H, dimen_H, new_H, EDError
    m, (nu, nd) = golden("A")
    o = oracle.OracleSector(m, nu, nd)
    new_H(m, nu, nd)
    n = dimen_H()
    v = np.random.default_rng(1).uniform(-1, 1, n)
    out = np.zeros(n)
    Hmult__(n, out, v)
    assert rel(out, o.hmult(v)) < 1e-13
    with pytest.raises(EDError):
        Hmult__(n - 1, out[:-1], v[:-1])
    delete_H()
    with pytest.raises(EDError):
        dimen_H()


def test_apply_c_bit_exact(gpu):
    m, (nu, nd) = golden("A")
    rng = np.random.default_rng(2)
    src = gpu.Sector(m, nu, nd)
    osrc = oracle.OracleSector(m, nu, nd)
    psi = rng.uniform(-1, 1, src.dim)
    for spin in (0, 1):
        for pm in (1, -1):
            dn = (nu + pm, nd) if spin == 0 else (nu, nd + pm)
            dst = gpu.Sector(m, *dn)
            odst = oracle.OracleSector(m, *dn)
            for site in range(m.nc):
                a = gpu.apply_C(src, dst, site, spin, pm, psi)
                b = oracle.apply_c(osrc, odst, site, spin, pm, psi)
                assert np.array_equal(a, b)
                c = bf.apply_c_full(m, nu, nd, site, spin, pm, psi)
                assert np.array_equal(a, c)
            dst.close()


def test_continued_fraction_vs_oracle(gpu):
    rng = np.random.default_rng(5)
    a = rng.uniform(-2, 2, 36); b = rng.uniform(0.1, 1.0, 36)
    wn = oracle.matsubara(50.0, 1024)
    g = gpu.invert_zmtridiag(a, b, wn)
    go = np.array([oracle.cfrac(a, b, 1j * w) for w in wn])
    assert rel(g, go) < 1e-14


def test_dynamic_workload_a(gpu):
    """Config A: GS + G(i w_n) on 1024 Matsubara points at beta = 50, FP64."""
    m, (nu, nd) = golden("A")
    wn = oracle.matsubara(50.0, 1024)
    s = gpu.Sector(m, nu, nd)
    o = oracle.OracleSector(m, nu, nd)
    r = s.lanczos(nitermax=128)
    ro = o.lanczos(nitermax=128)
    psi = s.ground_state(r)
    psi_o = o.ground_state(ro["iters"], ro["vecp"][0])
    psi_o *= np.sign(psi @ psi_o)
    assert np.max(np.abs(psi - psi_o)) < 1e-8
    E0 = r.evals[0]
    for iph, pm in ((1, 1), (2, -1)):
        dst = gpu.Sector(m, nu + pm, nd)
        odst = oracle.OracleSector(m, nu + pm, nd)
        phi = gpu.apply_C(s, dst, 0, 0, pm, psi_o)
        dyn, a, b = gpu.compute_dynamic(dst, iph, E0, phi, 36, wn)
        dyn_o, ao, bo = odst.dynamic(iph, E0, phi, 36, wn)
        assert rel(dyn, dyn_o) < 1e-10
        dst.close()


def test_dynamic_batched_matches_single(gpu):
    m, (nu, nd) = golden("B")
    wn = oracle.matsubara(100.0, 256)
    s = gpu.Sector(m, nu, nd)
    r = s.lanczos(nitermax=128)
    psi = s.ground_state(r)
    secs, phis, iphs, ens = [], [], [], []
    for site in range(m.nc):
        for iph, pm in ((1, 1), (2, -1)):
            dst = gpu.Sector(m, nu + pm, nd)
            secs.append(dst); phis.append(gpu.apply_C(s, dst, site, 0, pm, psi)); iphs.append(iph)
            ens.append(r.evals[0])
    dyn_b, _, _ = gpu.compute_dynamic_batched(secs, phis, iphs, ens, 36, wn)
    for k in range(len(secs)):
        d1, _, _ = gpu.compute_dynamic(secs[k], iphs[k], ens[k], phis[k], 36, wn)
        assert rel(dyn_b[k], d1) < 1e-12


def test_breakdown_is_an_error(gpu):
    """A start vector that is an eigenvector breaks down at step 1 (one-step convergence),
    a zero start vector is the reference's 'starting lanczos with null vector' STOP."""
    from ed_solver_amd import BreakdownError
    m = kanamori(1, 0, 2.0, 0.0, np.zeros(0), np.zeros(0), eimp=-1.0)
    s = gpu.Sector(m, 1, 1)   # dimension 1: no Lanczos step exists
    with pytest.raises(gpu.EDError):
        s.lanczos()
    s.close()
    m, (nu, nd) = golden("A")
    s = gpu.Sector(m, nu, nd)
    with pytest.raises(BreakdownError):
        s.lanczos(start=np.zeros(s.dim))
    s.close()


def test_workload_e_hmult(gpu):
    """Ns = 14 (dim 11,778,624): H.v against the oracle on one vector."""
    m, (nu, nd) = golden("E")
    s = gpu.Sector(m, nu, nd)
    o = oracle.OracleSector(m, nu, nd, nthreads=16)
    v = np.random.default_rng(9).uniform(-1, 1, s.dim)
    assert rel(s.hmult(v), o.hmult(v)) < 1e-13
    s.close()


@pytest.mark.parametrize("name", ["B", "C"])
def test_gs_vector_in_the_reference_call_order(gpu, name):
    """get_GS (solver.h:43-60): new_H, Lanczos, delete_H, new_H again, Lanczos_get_GS_sector.
    On the fresh sector the vector comes from the second pass (the recurrence again from
    the default start vector, Lanczos_plain.f90:156-229): psi overlaps the oracle's psi to
    1e-9 and equals the kept-basis psi of the first sector to 1e-12."""
    from ed_solver_amd.hclass import Lanczos_fast_diagonalize, Lanczos_get_GS_sector, delete_H, new_H
    m, (nu, nd) = golden(name)
    s1 = new_H(m, nu, nd)
    r = Lanczos_fast_diagonalize(neigen=2, nitermax=128, tolerance=1e-12)
    psi_kept = s1.ground_state(r, 0)
    delete_H()
    new_H(m, nu, nd)
    psi = Lanczos_get_GS_sector(r, 0)
    psi1 = Lanczos_get_GS_sector(r, 1)
    delete_H()
    o = oracle.OracleSector(m, nu, nd, nthreads=16)
    ro = o.lanczos(nitermax=128, neigen=2, tol=1e-12)
    assert r.iters == ro["iters"]
    psi_o = o.ground_state(ro["iters"], ro["vecp"][0])
    psi_o /= np.linalg.norm(psi_o)
    o.close()
    print(f"{name}: second pass vs oracle overlap 1 - {1 - abs(psi @ psi_o):.2e}, vs kept basis "
          f"{np.max(np.abs(psi - psi_kept)):.2e}", flush=True)
    assert abs(abs(psi @ psi_o) - 1) < 1e-9
    assert np.max(np.abs(psi - psi_kept)) < 1e-12
    assert abs(psi @ psi1) < 1e-8   # the second Ritz vector of the same pass
