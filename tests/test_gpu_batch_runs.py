"""The multi-run persistent kernel of the batched excitation runs (k_lz_persist_runs): every
compute_dynamic recurrence of a Green's function (green_class_computeAA.h:101-302 ->
green_class_compute_dynamic.f90:279-334) advanced in ONE cooperative launch.

Bars: G(i w_n) within 1e-10 relative of the oracle's orc_dynamic (FP64) where the
oracle runs in seconds, and within 1e-12 relative of the one-run-at-a-time device path and
of the two-launch batch (the same recurrence, partial sums added in another order);
alpha / beta within 1e-12.  Each test checks through ed_dynamic_batched_stats that the
persistent launch ran (or, for the refusal hook, that it was refused and the two-launch
batch ran instead)."""
import json
import os

import numpy as np
import pytest

import oracle
from ed_solver_amd.aim import AIMParams

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "models.json")


def golden(name):
    with open(GOLD) as f:
        d = json.load(f)[name]
    return AIMParams.from_json(d["model"]), tuple(d["sector"])


def rel(a, b):
    return np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b)))


def excitations(gpu, m, nu, nd, both_spins=False):
    """Ground state of (nu, nd) and its c+ / c excitations of every impurity orbital."""
    s = gpu.Sector(m, nu, nd)
    r = s.lanczos(nitermax=128)
    psi = s.ground_state(r)
    secs, phis, iphs = [], [], []
    spins = (0, 1) if both_spins else (0,)
    cache = {}
    for spin in spins:
        for site in range(m.nc):
            for iph, pm in ((1, 1), (2, -1)):
                key = (nu + pm, nd) if spin == 0 else (nu, nd + pm)
                if key not in cache:
                    cache[key] = gpu.Sector(m, *key)
                dst = cache[key]
                secs.append(dst)
                phis.append(gpu.apply_C(s, dst, site, spin, pm, psi))
                iphs.append(iph)
    return s, float(r.evals[0]), secs, phis, iphs, list(cache.values())


def test_runs_kernel_matches_oracle_single_and_two_launch(gpu):
    """Workload B (Ns=10, excited sectors (6,5), (4,5), (5,6), (5,4), dim 52,920): all 8
    runs in one launch against the oracle, the single runs and the two-launch batch."""
    m, (nu, nd) = golden("B")
    wn = oracle.matsubara(100.0, 256)
    s, E0, secs, phis, iphs, owned = excitations(gpu, m, nu, nd, both_spins=True)
    n0, _ = gpu.batched_stats()
    dyn, a, b = gpu.compute_dynamic_batched(secs, phis, iphs, [E0] * len(secs), 36, wn)
    n1, _ = gpu.batched_stats()
    assert n1 == n0 + 1, "the multi-run persistent kernel did not run"
    for k in range(len(secs)):
        o = oracle.OracleSector(m, secs[k].nup, secs[k].ndn)
        dyn_o, ao, bo = o.dynamic(iphs[k], E0, phis[k], 36, wn)
        o.close()
        assert rel(dyn[k], dyn_o) < 1e-10, k
        d1, a1, b1 = gpu.compute_dynamic(secs[k], iphs[k], E0, phis[k], 36, wn)
        assert rel(dyn[k], d1) < 1e-12, k
        assert np.max(np.abs(a[k] - a1)) < 1e-12 and np.max(np.abs(b[k] - b1)) < 1e-12, k
    # the two-launch batch (persistent kernel switched off on the sectors)
    for x in owned:
        x.set_persistent(False)
    dyn2, a2, b2 = gpu.compute_dynamic_batched(secs, phis, iphs, [E0] * len(secs), 36, wn)
    assert gpu.batched_stats()[0] == n1
    assert rel(dyn2, dyn) < 1e-12
    assert np.max(np.abs(a2 - a)) < 1e-12 and np.max(np.abs(b2 - b)) < 1e-12
    for x in owned + [s]:
        x.close()


def test_runs_kernel_uneven_dimensions_fp32(gpu):
    """Runs of uneven dimension in one launch: workload B's excitations (dim 52,920) with
    workload C's (dim 731,808, 3 of them), FP64 and FP32, each against its single run."""
    wn = oracle.matsubara(100.0, 512)
    mb, (bu, bd) = golden("B")
    mc, (cu, cd) = golden("C")
    sb, Eb, secs_b, phis_b, iph_b, own_b = excitations(gpu, mb, bu, bd)
    sc, Ec, secs_c, phis_c, iph_c, own_c = excitations(gpu, mc, cu, cd)
    secs = secs_b[:3] + secs_c[:3]
    phis = phis_b[:3] + phis_c[:3]
    iphs = iph_b[:3] + iph_c[:3]
    ens = [Eb] * 3 + [Ec] * 3
    for prec, tol in (("f64", 1e-12), ("f32", 5e-5)):
        n0, _ = gpu.batched_stats()
        dyn, a, b = gpu.compute_dynamic_batched(secs, phis, iphs, ens, 36, wn, precision=prec)
        assert gpu.batched_stats()[0] == n0 + 1
        for k in range(len(secs)):
            d1, _, _ = gpu.compute_dynamic(secs[k], iphs[k], ens[k], phis[k], 36, wn, precision=prec)
            err = rel(dyn[k], d1)
            print(f"  {prec} run {k} (dim {secs[k].dim}): rel diff to the single run {err:.1e}", flush=True)
            assert err < tol, (prec, k)
    for x in own_b + own_c + [sb, sc]:
        x.close()


def test_runs_kernel_refusal_hook(gpu):
    """Test hook 2 of ed_lanczos_set_persistent: the launch counts as refused, the two-launch
    batch runs, the results are the same."""
    from ed_solver_amd._lib import check
    m, (nu, nd) = golden("B")
    wn = oracle.matsubara(100.0, 128)
    s, E0, secs, phis, iphs, owned = excitations(gpu, m, nu, nd)
    ref, _, _ = gpu.compute_dynamic_batched(secs, phis, iphs, [E0] * len(secs), 36, wn)
    n0, r0 = gpu.batched_stats()
    for x in owned:
        check(gpu.lib().ed_lanczos_set_persistent(x.handle, 2))
    dyn, _, _ = gpu.compute_dynamic_batched(secs, phis, iphs, [E0] * len(secs), 36, wn)
    n1, r1 = gpu.batched_stats()
    assert (n1, r1) == (n0, r0 + 1)
    assert rel(dyn, ref) < 1e-12
    for x in owned + [s]:
        x.close()


def breakdown_model():
    """nc = 1, nb = 9, no hybridisation, one bath hop (sites 0 and 1): a basis state with
    bath sites 0 and 1 both filled or both empty in each spin is an exact eigenvector."""
    nb = 9
    eb1 = np.diag(np.linspace(-1.0, 1.0, nb))
    eb1[0, 1] = eb1[1, 0] = 0.3
    return AIMParams(nc=1, nb=nb, ec=np.full((2, 1, 1), -1.0), eb=np.stack([eb1, eb1]),
                     vbc=np.zeros((2, nb, 1)), u=np.array([[2.0]]), jj=np.zeros((1, 1)), jhund=0.0, cutoff=1e-8)


def test_runs_kernel_breakdown_is_an_error_and_drains(gpu):
    """A run whose start vector is an exact eigenvector breaks down at step 2 (beta_2 = 0,
    Lanczos_plain.h:85-92): the batched call raises BreakdownError, as the single run does,
    the launch ends (the other runs go on to the last barrier), and the next batch is right."""
    from ed_solver_amd import BreakdownError
    m = breakdown_model()
    dst = gpu.Sector(m, 5, 5)            # d_up = d_dn = C(10, 5) = 252
    up = oracle.sector_states(10, 5)
    dn = oracle.sector_states(10, 5)
    iu = int(np.flatnonzero((up & 3) == 3)[0])
    idn = int(np.flatnonzero((dn & 3) == 0)[0])
    phi = np.zeros(dst.dim)
    phi[iu + idn * len(up)] = 1.0
    rng = np.random.default_rng(3)
    good = rng.standard_normal(dst.dim)
    wn = oracle.matsubara(50.0, 64)
    with pytest.raises(BreakdownError):
        gpu.compute_dynamic(dst, 1, 0.0, phi, 20, wn)
    n0, _ = gpu.batched_stats()
    with pytest.raises(BreakdownError):
        gpu.compute_dynamic_batched([dst, dst], [good, phi], [1, 1], [0.0, 0.0], 20, wn)
    assert gpu.batched_stats()[0] == n0 + 1
    dyn, _, _ = gpu.compute_dynamic_batched([dst], [good], [1], [0.0], 20, wn)
    o = oracle.OracleSector(m, 5, 5)
    dyn_o, _, _ = o.dynamic(1, 0.0, good, 20, wn)
    o.close()
    assert rel(dyn[0], dyn_o) < 1e-10
    dst.close()


def test_lanczos_batched_matches_single_runs(gpu):
    """ed_lanczos_batched (Lanczos_GPU_updo for many sectors at once) against each sector's
    own ed_lanczos_run from the same default start vector: sectors of workload C of uneven
    dimension (the multi-run kernel, d_up = 66 included), one with d_up = 12 (run alone, in the
    shared work space); alpha/beta to 1e-10 relative.  The replayed convergence test gives the per-sector Lanczos E0 to 1e-10, and
    on a 66-state sector run for Niter = dim - 1 steps its exact lowest eigenvalue."""
    import json, os
    from ed_solver_amd.aim import AIMParams
    from ed_solver_amd.hclass import lanczos_batched, lanczos_converged
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "models.json")) as f:
        m = AIMParams.from_json(json.load(f)["C"]["model"])
    keys = [(6, 6), (5, 6), (4, 7), (3, 5), (2, 6), (10, 5), (1, 6)]
    secs = [gpu.Sector(m, *k) for k in keys]
    niter = 64
    print(f"batched Lanczos over {[s.dim for s in secs]}", flush=True)
    a, b, steps = lanczos_batched(secs, niter)
    assert list(steps) == [niter] * len(keys)
    for i, s in enumerate(secs):
        a1, b1 = s.lanczos_run(niter)
        scale = np.maximum(1.0, np.abs(a1))
        assert np.max(np.abs(a[i] - a1) / scale) < 1e-10, keys[i]
        assert np.max(np.abs(b[i] - b1) / np.maximum(1.0, np.abs(b1))) < 1e-10, keys[i]
        it, ev = lanczos_converged(a[i], b[i], 1, 1e-12, 128)
        r = s.lanczos(nitermax=niter, neigen=1, tolerance=1e-12)
        assert abs(ev[0] - r.evals[0]) < 1e-10, keys[i]
        assert abs(it - r.iters) <= 1, (keys[i], it, r.iters)
    for s in secs:
        s.close()
    # a sector of dimension 66 with d_up = 1 (run alone) and the reference's Niter = dim - 1
    tiny = gpu.Sector(m, 0, 2)
    a, b, steps = lanczos_batched([tiny], tiny.dim - 1)
    assert steps[0] == tiny.dim - 1
    it, ev = lanczos_converged(a[0, :steps[0]], b[0, :steps[0]], 1, 1e-12, 128)
    assert abs(ev[0] - tiny.full_ed(1)[0][0]) < 1e-10
    tiny.close()


def test_dynamic_from_states_matches_per_state(gpu):
    """ed_dynamic_from_states (the runs of several eigenstates in one call) against one
    ed_dynamic_from_state call per state: two states of workload B in different sectors, all
    c / c+ runs of every orbital and a (c_0 + c_1) combination; G to 1e-12 relative."""
    import json, os
    from ed_solver_amd.aim import AIMParams
    from ed_solver_amd.hclass import dynamic_from_state, dynamic_from_states
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "models.json")) as f:
        m = AIMParams.from_json(json.load(f)["B"]["model"])
    wn = gpu.matsubara(50.0, 256)
    states, jobs = [], []
    opened = []
    for q, (nu, nd) in enumerate([(5, 5), (4, 5)]):
        s = gpu.Sector(m, nu, nd); opened.append(s)
        r = s.lanczos(nitermax=128, neigen=1, tolerance=1e-12)
        states.append((s, s.ground_state(r), float(r.evals[0])))
        sp = gpu.Sector(m, nu + 1, nd); sh = gpu.Sector(m, nu - 1, nd); opened += [sp, sh]
        for c in range(m.nc):
            jobs.append((q, sp, 1, [(c, 1.0)]))
            jobs.append((q, sh, 2, [(c, 1.0)]))
        jobs.append((q, sp, 1, [(0, 1.0), (1, 1.0)]))
    dyn = dynamic_from_states(states, jobs, 0, 36, wn)
    for q, (s, psi, e) in enumerate(states):
        mine = [k for k, j in enumerate(jobs) if j[0] == q]
        ref = dynamic_from_state(s, psi, e, [jobs[k][1:] for k in mine], 0, 36, wn)
        assert np.max(np.abs(dyn[mine] - ref)) < 1e-12 * np.max(np.abs(ref))
    for s in opened:
        s.close()
