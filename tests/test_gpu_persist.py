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


def test_persistent_long_run_and_excluded_width(gpu):
    """150 steps in one launch on workload C; the excited sector (7,6) (d_up 792, ELL
    width 12, four rows per thread) is outside the compiled instances and keeps the
    two-kernel step."""
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
    assert not s.persistent_active("f64") and not s.persistent_active("f32")
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
