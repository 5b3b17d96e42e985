"""The persistent Lanczos kernel (one cooperative launch per chunk of steps, grid
barriers between the H.v and the update) against the two-kernel step and the oracle.

With the fast H.v at 256 threads per workgroup (d_up = 252 with one row per thread,
d_up = 924 with four) both paths form every sum in the same order, so alpha and beta
must agree bit for bit; against the oracle the bar is the FP64 1e-10 of the other
Lanczos tests.
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


def run(s, niter, persistent, precision="f64", start=None):
    s.set_persistent(persistent)
    return s.lanczos_run(niter, start=start, precision=precision)


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
    s.set_persistent(True)
    assert s.persistent_active("f64")
    with pytest.raises(BreakdownError):
        s.lanczos_run(40, start=np.zeros(s.dim))
    # the sector stays usable afterwards, on the persistent path too
    a1, b1 = s.lanczos_run(40)
    s.set_persistent(False)
    a0, b0 = s.lanczos_run(40)
    assert np.array_equal(a1, a0) and np.array_equal(b1, b0)
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
    s.set_persistent(True)
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
