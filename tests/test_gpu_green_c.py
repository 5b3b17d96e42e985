"""G(i w_n) at the metric workload (Ns=12, sector (6,6), dim 853,776) against the oracle.

compute_dynamic (green_class_compute_dynamic.f90:191-407) on the six excitation vectors
c+_(c, up)|psi> and c_(c, up)|psi> of the three impurity orbitals (sectors (7,6) and (5,6),
dim 731,808), 36 Lanczos steps, 2048 Matsubara points at beta = 100, computed by the
device path that the solver uses (ed_dynamic_from_state: the state copied once, the
excitation vectors and all six runs on the device) and by orc_dynamic, the C
restatement, from the same psi.

Bars: FP64 within 1e-10 relative (max |G - G_o| / max |G_o| per run).  FP32 H.v with
FP64 recurrence scalars and continued fraction: the relative error at each w_n,
|G32(i w_n) - G_o(i w_n)| / |G_o(i w_n)|, stays below 5e-5.  It is largest at the lowest
frequency, where G is most sensitive to the pole positions the FP32 vectors shift
(measured 2.0e-5 at w_1 = pi/100, 1e-8 at the median w_n, 5e-9 at w_2048; printed)."""
import json
import os

import numpy as np
import pytest

import oracle
from ed_solver_amd.aim import AIMParams

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "models.json")
F32_PER_W_TOL = 5e-5


def golden(name):
    with open(GOLD) as f:
        d = json.load(f)[name]
    return AIMParams.from_json(d["model"]), tuple(d["sector"])


@pytest.fixture(scope="module")
def workload_c(gpu):
    m, (nu, nd) = golden("C")
    s = gpu.Sector(m, nu, nd)
    r = s.lanczos(nitermax=128, neigen=1, tolerance=1e-12)
    psi = s.ground_state(r)
    E0 = float(r.evals[0])
    wn = oracle.matsubara(100.0, 2048)
    sp, sh = gpu.Sector(m, nu + 1, nd), gpu.Sector(m, nu - 1, nd)
    jobs = []
    for c in range(m.nc):
        jobs.append((sp, 1, [(c, 1.0)]))
        jobs.append((sh, 2, [(c, 1.0)]))
    # oracle: orc_apply_c + orc_dynamic from the same psi
    o = oracle.OracleSector(m, nu, nd, nthreads=16)
    op, oh = oracle.OracleSector(m, nu + 1, nd, nthreads=16), oracle.OracleSector(m, nu - 1, nd, nthreads=16)
    ref = []
    for c in range(m.nc):
        for dst, iph, pm in ((op, 1, +1), (oh, 2, -1)):
            phi = oracle.apply_c(o, dst, c, 0, pm, psi)
            dyn_o, _, _ = dst.dynamic(iph, E0, phi, 36, wn)
            ref.append(dyn_o)
    for x in (o, op, oh):
        x.close()
    print(f"workload C: dim {s.dim}, excited dim {sp.dim}, E0 {E0:.12f}", flush=True)
    yield s, psi, E0, jobs, wn, np.array(ref)
    for x in (s, sp, sh):
        x.close()


def test_green_workload_c_fp64_vs_oracle(gpu, workload_c):
    s, psi, E0, jobs, wn, ref = workload_c
    dyn = gpu.dynamic_from_state(s, psi, E0, jobs, 0, 36, wn)
    for k in range(len(jobs)):
        err = np.max(np.abs(dyn[k] - ref[k])) / np.max(np.abs(ref[k]))
        print(f"  run {k} (orbital {k // 2}, iph {jobs[k][1]}): max rel err {err:.2e}", flush=True)
        assert err < 1e-10
    # the batched call from host excitation vectors agrees too
    phis = [gpu.apply_C(s, j[0], j[2][0][0], 0, +1 if j[1] == 1 else -1, psi) for j in jobs]
    dyn_h, _, _ = gpu.compute_dynamic_batched([j[0] for j in jobs], phis, [j[1] for j in jobs], [E0] * len(jobs), 36,
                                              wn)
    assert np.max(np.abs(dyn_h - ref)) / np.max(np.abs(ref)) < 1e-10


def test_green_workload_c_fp32_per_w_error(gpu, workload_c):
    """FP32 H.v and vectors, FP64 recurrence scalars and continued fraction: the error of
    G(i w_n) against the FP64 oracle, stated per w_n."""
    s, psi, E0, jobs, wn, ref = workload_c
    dyn = gpu.dynamic_from_state(s, psi, E0, jobs, 0, 36, wn, precision="f32")
    per_w = np.max(np.abs(dyn - ref) / np.abs(ref), axis=0)   # worst run at each w_n
    print(f"  FP32 G: per-w relative error max {per_w.max():.2e} (w_1 {per_w[0]:.2e}, median {np.median(per_w):.2e}, "
          f"w_2048 {per_w[-1]:.2e})", flush=True)
    assert per_w.max() < F32_PER_W_TOL
    # the single-run path in FP32 has the same error bound
    phi = gpu.apply_C(s, jobs[0][0], 0, 0, +1, psi)
    d1, _, _ = gpu.compute_dynamic(jobs[0][0], 1, E0, phi, 36, wn, precision="f32")
    assert np.max(np.abs(d1 - ref[0]) / np.abs(ref[0])) < F32_PER_W_TOL
