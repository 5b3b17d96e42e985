"""The Rayleigh form of the mixed-precision ground state (ED_PREC_MIXED_RQ): FP32
Lanczos, then E0 from one FP64 H.v of its Ritz vector when the Kato-Temple bound allows,
against the FP64 reference solve at the workload sizes B and C; the fallback to the
FP64 restart when the bound does not allow it, and for more than one eigenvalue."""
import json
import os

import numpy as np
import pytest

from ed_solver_amd.aim import AIMParams

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "models.json")


def golden(name):
    with open(GOLD) as f:
        d = json.load(f)[name]
    return AIMParams.from_json(d["model"]), tuple(d["sector"])


@pytest.mark.parametrize("name", ["B", "C"])
def test_rayleigh_form_ground_energy(gpu, name):
    m, sec = golden(name)
    s = gpu.Sector(m, *sec)
    print(f"{name}: dim {s.dim}", flush=True)
    e_ref = float(s.lanczos(nitermax=300, neigen=1, tolerance=1e-13).evals[0])
    r = s.lanczos(nitermax=300, neigen=1, tolerance=1e-9, precision="mixed_rq")
    acc, r2, gap = s.rq_info()
    print(f"{name}: res2 {r2:.3e} gap {gap:.4f} accepted {acc} ({r.iters_f32} FP32 it)", flush=True)
    assert acc and gap > 0.0 and r2 <= 1e-9 * gap / 2
    e = float(r.evals[0])
    # a Rayleigh quotient lies above E0, by at most res2 / (lambda_2 - E)
    assert e >= e_ref - 1e-12 and e - e_ref <= 2 * r2 / gap + 1e-12
    assert abs(e - e_ref) < 1e-8
    assert r.iters == r.iters_f32
    psi = s.ground_state(r)
    assert abs(np.linalg.norm(psi) - 1.0) < 1e-10 and abs(s.rayleigh(psi) - e) < 1e-12
    # a tolerance below what FP32 vectors reach: the FP32 phase stops waiting for the
    # target as soon as its eigenvalue test is met and the FP64 restart runs
    r_strict = s.lanczos(nitermax=300, neigen=1, tolerance=1e-14, precision="mixed_rq")
    acc2, r2b, gapb = s.rq_info()
    r_plain = s.lanczos(nitermax=300, neigen=1, tolerance=1e-14, precision="f32")
    print(f"{name}: accepted res2 {r2:.2e} gap {gap:.3f} ({r.iters_f32} FP32 it); strict tolerance: "
          f"{r_strict.iters_f32} FP32 + {r_strict.iters} FP64 it (mixed form without the Rayleigh test "
          f"{r_plain.iters_f32} + {r_plain.iters})", flush=True)
    assert not acc2
    assert r_strict.iters_f32 == r_plain.iters_f32
    assert abs(float(r_strict.evals[0]) - e_ref) < 1e-11
    # two eigenvalues: the FP64 restart as well
    s.lanczos(nitermax=300, neigen=2, tolerance=1e-9, precision="mixed_rq")
    assert not s.rq_info()[0]
    s.close()


@pytest.mark.parametrize("split", [1e-4, 1e-3, 1e-2])
def test_rayleigh_form_small_gap(gpu, split):
    """A nearly degenerate ground state (two orbitals with identical baths, one bath level
    moved by `split`: lambda_2 - lambda_1 = 1.1e-6, 1.1e-5, 1.1e-4 in the (5, 4) sector).
    Whenever the Rayleigh form accepts its E, |E - E0| <= tolerance, unless the FP64
    reference solve at the same tolerance is itself further off (a pair the Lanczos run
    has not split yet: neither form can see the second state, and both stop on the pair's
    mixture); a result that is not accepted is the FP64 restart's."""
    from ed_solver_amd.aim import kanamori
    import oracle
    eps = np.array([-0.7, -0.7, 0.4, 0.4, 1.1, 1.1])
    eps[1] += split
    v = np.array([0.35, 0.35, 0.5, 0.5, 0.25, 0.25])
    m = kanamori(2, 6, 3.0, 0.5, eps, v)
    o = oracle.OracleSector(m, 5, 4, nthreads=8)
    w, _ = oracle.lowest_eigs(o, 2)
    o.close()
    s = gpu.Sector(m, 5, 4)
    for tol in (1e-9, 1e-11):
        e64 = float(s.lanczos(nitermax=300, neigen=1, tolerance=tol).evals[0])
        r = s.lanczos(nitermax=300, neigen=1, tolerance=tol, precision="mixed_rq")
        acc, r2, gap = s.rq_info()
        err, err64 = abs(float(r.evals[0]) - w[0]), abs(e64 - w[0])
        print(f"split {split:g} (lambda_2 - lambda_1 = {w[1] - w[0]:.2e}) tol {tol:g}: accepted {acc}, "
              f"res2 {r2:.2e}, gap {gap:.2e}, |E - E0| {err:.2e} (FP64 solve {err64:.2e}), FP32 it {r.iters_f32}",
              flush=True)
        if acc:
            assert err <= max(tol, err64 + tol)
        else:
            assert err <= max(10 * tol, 2 * err64 + tol)
    s.close()
