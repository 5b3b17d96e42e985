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
    # a tolerance the FP32 Ritz vector cannot meet: the FP64 restart runs
    r_strict = s.lanczos(nitermax=300, neigen=1, tolerance=1e-14, precision="mixed_rq")
    acc2, r2b, gapb = s.rq_info()
    print(f"{name}: accepted res2 {r2:.2e} gap {gap:.3f} ({r.iters_f32} FP32 it); strict res2 {r2b:.2e}", flush=True)
    assert not acc2 and r2b > 1e-14 * gapb / 2
    assert abs(float(r_strict.evals[0]) - e_ref) < 1e-11
    # two eigenvalues: the FP64 restart as well
    s.lanczos(nitermax=300, neigen=2, tolerance=1e-9, precision="mixed_rq")
    assert not s.rq_info()[0]
    s.close()
