"""The persistent Lanczos kernel (one cooperative launch per chunk of steps, grid
barriers between the H.v and the update) against the two-kernel step and the oracle.

With the fast H.v at 256 threads per workgroup (d_up = 252 with one row per thread,
d_up = 924 with four) the two-barrier form and the two-kernel step form every sum in the
same order, so alpha and beta must agree bit for bit; against the oracle the bar is the
FP64 1e-10 of the other Lanczos tests.  The one-barrier form (the default where a thread
owns one row) forms beta from the expansion of the norm and rebuilds the exchanged vector
at every gather: it must agree with the two-barrier form to rounding over the steps
before the Lanczos vectors lose orthogonality, and give the same Ritz values.
"""
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


def run(s, niter, persistent, precision="f64", start=None, form="two-barrier"):
    s.set_persistent(persistent, form)
    return s.lanczos_run(niter, start=start, precision=precision)


def ritz(a, b, k=3):
    t = np.diag(a) + np.diag(b[1:], 1) + np.diag(b[1:], -1)
    return np.linalg.eigvalsh(t)[:k]


@pytest.mark.parametrize("name,sched", [("B", 8), ("B", 16), ("C", 20)])
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_persistent_matches_two_kernel_step_bitwise(gpu, name, sched, precision):
    m, (nu, nd) = golden(name)
    s = gpu.Sector(m, nu, nd)
    s.set_schedule(sched)
    assert s.persistent_active(precision), "persistent kernel expected to apply here"
    print(f"{name} schedule {sched} {precision}: dim {s.dim}", flush=True)
    a1, b1 = run(s, 60, True, precision)
    a0, b0 = run(s, 60, False, precision)
    assert np.array_equal(a1, a0) and np.array_equal(b1, b0)
    s.close()


@pytest.mark.parametrize("name,sched", [("B", 16), ("C", 20)])
def test_persistent_vs_oracle(gpu, name, sched):
    """A fixed 40-step run (one persistent launch) against the oracle recurrence, and the
    Ritz value of its tridiagonal against the oracle's E0 after 40 steps."""
    m, (nu, nd) = golden(name)
    s = gpu.Sector(m, nu, nd)
    o = oracle.OracleSector(m, nu, nd, nthreads=16)
    s.set_schedule(sched)
    s.set_persistent(True)
    assert s.persistent_active("f64")
    print(f"{name}: dim {s.dim}", flush=True)
    a, b = s.lanczos_run(40)
    ro = o.lanczos(nitermax=40, neigen=1, tol=0.0)
    assert ro["iters"] == 40
    assert np.max(np.abs(a - ro["alpha"][:40])) < 1e-10
    assert np.max(np.abs(b - ro["beta"][:40])) < 1e-10
    t = np.diag(a) + np.diag(b[1:], 1) + np.diag(b[1:], -1)
    assert abs(np.linalg.eigvalsh(t)[0] - ro["evals"][0]) < 1e-10
    s.close()


def test_persistent_long_run_and_wide_table(gpu):
    """150 steps in one launch on workload C; the excited sector (7,6) of its G(i w_n) runs
    (d_up 792, ELL width 12, four rows per thread: the instance that reloads its packed
    up-hop entries every step) against the two-kernel step, FP64 and FP32."""
    m, (nu, nd) = golden("C")
    s = gpu.Sector(m, nu, nd)
    s.set_schedule(20)
    print(f"C: dim {s.dim}", flush=True)
    a1, b1 = run(s, 150, True)
    a0, b0 = run(s, 150, False)
    assert np.array_equal(a1, a0) and np.array_equal(b1, b0)
    s.close()
    s = gpu.Sector(m, 7, 6)
    s.set_schedule(20)
    print(f"(7,6): dim {s.dim}", flush=True)
    for prec in ("f64", "f32"):
        s.set_persistent(True)
        assert s.persistent_active(prec)
        a1, b1 = run(s, 60, True, prec)
        a0, b0 = run(s, 60, False, prec)
        assert np.array_equal(a1, a0) and np.array_equal(b1, b0), prec
    s.close()


def test_persistent_zero_start_vector_is_an_error(gpu):
    from ed_solver_amd import BreakdownError
    m, (nu, nd) = golden("B")
    s = gpu.Sector(m, nu, nd)
    s.set_schedule(16)
    for form in ("one-barrier", "two-barrier"):
        s.set_persistent(True, form)
        assert s.persistent_active("f64")
        with pytest.raises(BreakdownError):
            s.lanczos_run(40, start=np.zeros(s.dim))
    # the sector stays usable afterwards, on the persistent path too
    a1, b1 = s.lanczos_run(40)
    s.set_persistent(False)
    a0, b0 = s.lanczos_run(40)
    assert np.array_equal(a1, a0) and np.array_equal(b1, b0)
    s.set_persistent(True, "one-barrier")
    a2, b2 = s.lanczos_run(40)
    assert np.max(np.abs(a2 - a0)) < 1e-12 and np.max(np.abs(b2 - b0)) < 1e-12
    s.close()


@pytest.mark.parametrize("name,sched", [("B", 16), ("C", 20)])
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_monitored_diagonalize_matches_chunked(gpu, name, sched, precision):
    """The ground-state solve with the convergence test: one persistent launch watched
    from the host (stopped a few steps after the iteration that ends the loop) against
    chunks of 8 two-kernel steps with the test between them.  Same scalars, so the same
    iteration count, eigenvalues, Lanczos coefficients and ground-state vector; the mixed
    form runs both of its phases this way."""
    m, (nu, nd) = golden(name)
    s = gpu.Sector(m, nu, nd)
    s.set_schedule(sched)
    assert s.persistent_active(precision)
    print(f"{name} schedule {sched} {precision}: dim {s.dim}", flush=True)
    s.set_persistent(True, "two-barrier")
    r1 = s.lanczos(nitermax=128, neigen=2, tolerance=1e-12, precision=precision)
    psi1 = s.ground_state(r1)
    s.set_persistent(False)
    r0 = s.lanczos(nitermax=128, neigen=2, tolerance=1e-12, precision=precision)
    psi0 = s.ground_state(r0)
    assert r1.iters == r0.iters and r1.iters_f32 == r0.iters_f32
    assert np.array_equal(r1.diag, r0.diag) and np.array_equal(r1.subdiag, r0.subdiag)
    assert np.array_equal(r1.evals, r0.evals) and np.array_equal(r1.vecp, r0.vecp)
    assert np.max(np.abs(psi1 - psi0)) < 1e-13
    # the chunked run stops at the end of the chunk holding the last iteration; the
    # monitored one within a few steps of it
    assert r1.steps_run >= r1.iters + r1.iters_f32
    assert r1.steps_run <= 128 * (2 if precision == "f32" else 1)
    print(f"  iters {r1.iters} (+{r1.iters_f32} FP32), steps run {r1.steps_run} vs chunked {r0.steps_run}", flush=True)
    s.close()


def test_monitored_vs_oracle_and_short_limit(gpu):
    """Monitored solve on workload B against the oracle's Lanczos (E0, iterations), and a
    limit of 20 iterations that ends the loop at it == nitermax."""
    m, (nu, nd) = golden("B")
    s = gpu.Sector(m, nu, nd)
    o = oracle.OracleSector(m, nu, nd, nthreads=16)
    s.set_schedule(16)
    s.set_persistent(True)
    assert s.persistent_active("f64")
    r = s.lanczos(nitermax=128, neigen=1, tolerance=1e-12)
    ro = o.lanczos(nitermax=128, neigen=1, tol=1e-12)
    assert abs(r.iters - ro["iters"]) <= 1   # the test compares Ritz values 1e-12 apart
    assert abs(r.evals[0] - ro["evals"][0]) < 1e-10
    r20 = s.lanczos(nitermax=20, neigen=1, tolerance=1e-12)
    assert r20.iters == 20 and r20.steps_run == 20
    assert np.max(np.abs(r20.diag - ro["alpha"][:20])) < 1e-10
    s.close()


def test_refused_cooperative_launch_takes_the_two_kernel_step(gpu):
    """The refusal path of the persistent kernel (forced by the test hook enable = 2, which
    takes it without launching): the run continues with the two-kernel step and its
    alpha / beta are those of that step; the refusal is counted and the kernel is off for
    the sector afterwards.  The monitored ground-state solve takes the same path."""
    from ed_solver_amd._lib import check, lib
    m, (nu, nd) = golden("C")
    s = gpu.Sector(m, nu, nd)
    s.set_schedule(20)
    a0, b0 = run(s, 40, False)
    r0 = s.lanczos(nitermax=128, neigen=1, tolerance=1e-12)
    check(lib().ed_lanczos_set_persistent(s.handle, 2))
    assert s.persistent_active("f64")
    a1, b1 = s.lanczos_run(40)
    assert s.persistent_refusals() == 1 and not s.persistent_active("f64")
    assert np.array_equal(a1, a0) and np.array_equal(b1, b0)
    check(lib().ed_lanczos_set_persistent(s.handle, 2))   # resets the applicability check
    r1 = s.lanczos(nitermax=128, neigen=1, tolerance=1e-12)
    assert s.persistent_refusals() == 2
    assert r1.iters == r0.iters and np.array_equal(r1.evals, r0.evals)
    s.close()


@pytest.mark.parametrize("sched", [8, 16])
def test_one_barrier_form_against_two_barrier_and_oracle(gpu, sched):
    """Workload B (d_up = 252, one row per thread): the one-barrier form applies by
    default.  128 FP64 steps: alpha / beta within 1e-12 of the two-barrier form over the
    first 80 steps (measured 1e-14), the lowest three Ritz values of T_128 within 1e-11;
    40 steps against the oracle to 1e-10; the hook that takes every beta from the reduced
    norm (one more barrier per step) to the same bars;
    FP32: the lowest Ritz value of T_60 within 1e-5 relative of the two-barrier form's."""
    m, (nu, nd) = golden("B")
    s = gpu.Sector(m, nu, nd)
    o = oracle.OracleSector(m, nu, nd, nthreads=16)
    s.set_schedule(sched)
    print(f"B schedule {sched}: dim {s.dim}", flush=True)
    s.set_persistent(True)
    assert s.persistent_shape()["rows"] == 1 and s.persistent_active("f64")
    a1, b1 = run(s, 128, True, form="one-barrier")
    a0, b0 = run(s, 128, True, form="two-barrier")
    assert np.max(np.abs(a1[:80] - a0[:80])) < 1e-12
    assert np.max(np.abs(b1[:80] - b0[:80]) / b0[:80]) < 1e-12
    assert np.max(np.abs(ritz(a1, b1) - ritz(a0, b0))) < 1e-11
    ae, be = run(s, 128, True, form="exact-beta")
    assert np.max(np.abs(ae[:80] - a0[:80])) < 1e-12 and np.max(np.abs(be[:80] - b0[:80]) / b0[:80]) < 1e-12
    assert np.max(np.abs(ritz(ae, be) - ritz(a0, b0))) < 1e-11
    a40, b40 = run(s, 40, True, form="one-barrier")
    ro = o.lanczos(nitermax=40, neigen=1, tol=0.0)
    assert np.max(np.abs(a40 - ro["alpha"][:40])) < 1e-10
    assert np.max(np.abs(b40 - ro["beta"][:40])) < 1e-10
    f1, g1 = run(s, 60, True, "f32", form="one-barrier")
    f0, g0 = run(s, 60, True, "f32", form="two-barrier")
    e1, e0 = ritz(f1, g1, 1)[0], ritz(f0, g0, 1)[0]
    print(f"  FP32 E0 one-barrier {e1:.9f} two-barrier {e0:.9f}", flush=True)
    assert abs(e1 - e0) < 1e-5 * abs(e0)
    o.close()
    s.close()


def test_one_barrier_monitored_solve(gpu):
    """The monitored ground-state solve through the one-barrier form on workload B: the
    same E0 as the two-barrier form to 1e-11 and the oracle's to 1e-10, iteration counts
    within one (the convergence test compares Ritz values 1e-12 apart), the ground-state
    vector within 1e-9 of the two-barrier form's."""
    m, (nu, nd) = golden("B")
    s = gpu.Sector(m, nu, nd)
    o = oracle.OracleSector(m, nu, nd, nthreads=16)
    s.set_schedule(16)
    s.set_persistent(True, "one-barrier")
    r1 = s.lanczos(nitermax=128, neigen=1, tolerance=1e-12)
    psi1 = s.ground_state(r1)
    s.set_persistent(True, "two-barrier")
    r0 = s.lanczos(nitermax=128, neigen=1, tolerance=1e-12)
    psi0 = s.ground_state(r0)
    ro = o.lanczos(nitermax=128, neigen=1, tol=1e-12)
    print(f"  iters one-barrier {r1.iters} two-barrier {r0.iters} oracle {ro['iters']}", flush=True)
    assert r1.iters < 128 and abs(r1.iters - r0.iters) <= 1 and abs(r1.iters - ro["iters"]) <= 1
    assert abs(r1.evals[0] - r0.evals[0]) < 1e-11
    assert abs(r1.evals[0] - ro["evals"][0]) < 1e-10
    # two eigenvalues (not converged to 1e-12 in 128 steps): the same Ritz values
    q1 = s.lanczos(nitermax=128, neigen=2, tolerance=1e-12)
    s.set_persistent(True, "one-barrier")
    q0 = s.lanczos(nitermax=128, neigen=2, tolerance=1e-12)
    assert np.max(np.abs(q1.evals[:2] - q0.evals[:2])) < 1e-11
    assert min(np.max(np.abs(psi1 - psi0)), np.max(np.abs(psi1 + psi0))) < 1e-9
    o.close()
    s.close()


def test_one_barrier_breakdown_takes_the_exact_norm(gpu):
    """An exact eigenvector as start vector (a basis state the Hamiltonian leaves alone):
    beta_2 expands to N - 2 c C + c^2 E with N = c^2 E, complete cancellation, so the kernel
    reduces the norm of u_2 itself, finds it zero and stops as the reference does
    (Lanczos_plain.h:85-92); the sector then runs a good vector as before."""
    from ed_solver_amd import BreakdownError
    from test_gpu_batch_runs import breakdown_model
    m = breakdown_model()
    s = gpu.Sector(m, 5, 5)
    s.set_persistent(True, "one-barrier")
    assert s.persistent_shape()["rows"] == 1 and s.persistent_active("f64")
    up = oracle.sector_states(10, 5)
    iu = int(np.flatnonzero((up & 3) == 3)[0])
    idn = int(np.flatnonzero((up & 3) == 0)[0])
    phi = np.zeros(s.dim)
    phi[iu + idn * len(up)] = 1.0
    with pytest.raises(BreakdownError):
        s.lanczos_run(20, start=phi)
    assert s.exact_beta_steps() == 1   # beta_2 from the reduced norm, then the stop
    good = np.random.default_rng(3).standard_normal(s.dim)
    a1, b1 = s.lanczos_run(30, start=good)
    print(f"  random start: {s.exact_beta_steps()} of 30 steps took the reduced norm", flush=True)
    s.set_persistent(False)
    a0, b0 = s.lanczos_run(30, start=good)
    assert np.max(np.abs(a1 - a0)) < 1e-12 and np.max(np.abs(b1 - b0) / b0) < 1e-12
    s.close()


@pytest.mark.parametrize("name,sector", [("B", None), ("C", (3, 6))])
def test_tuned_lanczos_schedule_takes_the_persistent_kernel(gpu, name, sector):
    """After the autotuner, the Lanczos recurrence runs the persistent kernel even where the
    row kernel gives the fastest standalone H.v (within 1.5x of a fast-kernel schedule): the
    tuned sector reports the persistent kernel active and its ground-state solve matches the
    oracle (E0 to 1e-10, iterations within one)."""
    m, (nu, nd) = golden(name)
    if sector:
        nu, nd = sector
    s = gpu.Sector(m, nu, nd)
    choice, ms = s.tune()
    print(f"{name} ({nu},{nd}): dim {s.dim}, H.v schedule {choice}, ms {np.round(np.array(ms) * 1e3, 2)}", flush=True)
    assert s.persistent_active("f64")
    o = oracle.OracleSector(m, nu, nd, nthreads=16)
    r = s.lanczos(nitermax=128, neigen=1, tolerance=1e-12)
    ro = o.lanczos(nitermax=128, neigen=1, tol=1e-12)
    assert abs(r.evals[0] - ro["evals"][0]) < 1e-10 and abs(r.iters - ro["iters"]) <= 1
    o.close()
    s.close()
