"""ed_dynamic_noalloc (the allocation-free variant of ed_dynamic, static device scratch)
against ed_dynamic through the C-ABI, on two calls in a row, the second with a smaller
input.

Bars (as every G(i w_n) parity check): FP64 dyn, a, b within 1e-10 relative of
ed_dynamic; FP32 within the stated FP32 per-w_n bar, 5e-5 relative.  The first call
runs on a sector that was never tuned, so it takes the untuned default H.v schedule;
ed_dynamic tunes the sector afterwards and the second sector is tuned first.  Free
device memory (hipMemGetInfo) is the same before and after a call.  Sizes above the
scratch's limits are refused with ED_ERR_ARG."""
import ctypes

import numpy as np
import pytest

from ed_solver_amd._lib import ED_ERR_ARG, ED_PREC_F32, ED_PREC_F64, check, dptr, lib

pytestmark = pytest.mark.gpu


def free_mem():
    from hipmem import hip
    h = hip()
    h.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    assert h.hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)) == 0
    return f.value


def run(fn, sec, iph, energy, phi, niter, prec, wn):
    nw = len(wn)
    n = min(niter, sec.dim)
    dyn = np.zeros(2 * nw); a = np.zeros(n); b = np.zeros(n)
    rc = fn(sec.handle, iph, energy, dptr(phi), niter, 1e-8, prec, nw, dptr(wn), dptr(dyn), dptr(a), dptr(b))
    return rc, dyn[0::2] + 1j * dyn[1::2], a, b


def rel(x, y):
    return np.max(np.abs(x - y)) / np.max(np.abs(y))


def excitation(ed, name):
    model, (nu, nd) = ed.synthetic(name)
    s = ed.Sector(model, nu, nd)
    r = s.lanczos(nitermax=128)
    psi = s.ground_state(r)
    dst = ed.Sector(model, nu + 1, nd)
    phi = ed.apply_C(s, dst, 0, 0, +1, psi)
    s.close()
    return dst, float(r.evals[0]), np.ascontiguousarray(phi)


@pytest.mark.parametrize("prec,tol", [(ED_PREC_F64, 1e-10), (ED_PREC_F32, 5e-5)])
def test_noalloc_matches_ed_dynamic(gpu, prec, tol):
    ed = gpu
    L = lib()
    cases = [("C", 36, 2048), ("B", 30, 300)]   # the second call is the smaller input
    for name, niter, nw in cases:
        dst, e0, phi = excitation(ed, name)
        wn = ed.matsubara(100.0, nw)
        print(f"no-alloc G: workload {name}, sector dim {dst.dim}, {niter} steps, {nw} w_n, prec {prec}", flush=True)
        before = free_mem()
        rc, d0, a0, b0 = run(L.ed_dynamic_noalloc, dst, 1, e0, phi, niter, prec, wn)
        check(rc)
        assert free_mem() == before
        rc, d1, a1, b1 = run(L.ed_dynamic, dst, 1, e0, phi, niter, prec, wn)
        check(rc)
        if prec == ED_PREC_F64:
            assert rel(d0, d1) < tol and rel(a0, a1) < tol and rel(b0, b1) < tol
        else:
            assert np.max(np.abs(d0 - d1) / np.abs(d1)) < tol
        # again, now on the tuned sector: same results
        rc, d2, a2, b2 = run(L.ed_dynamic_noalloc, dst, 1, e0, phi, niter, prec, wn)
        check(rc)
        assert np.max(np.abs(d2 - d1) / np.abs(d1)) < tol
        dst.close()


def test_noalloc_hole_run_and_limits(gpu):
    ed = gpu
    L = lib()
    model, (nu, nd) = ed.synthetic("B")
    s = ed.Sector(model, nu, nd)
    r = s.lanczos(nitermax=128)
    psi = s.ground_state(r)
    dst = ed.Sector(model, nu - 1, nd)
    phi = np.ascontiguousarray(ed.apply_C(s, dst, 1, 0, -1, psi))
    wn = ed.matsubara(50.0, 512)
    rc, d0, _, _ = run(L.ed_dynamic_noalloc, dst, 2, float(r.evals[0]), phi, 36, ED_PREC_F64, wn)
    check(rc)
    rc, d1, _, _ = run(L.ed_dynamic, dst, 2, float(r.evals[0]), phi, 36, ED_PREC_F64, wn)
    check(rc)
    assert rel(d0, d1) < 1e-10
    # nw beyond the scratch
    wn_big = ed.matsubara(50.0, 16385)
    rc, *_ = run(L.ed_dynamic_noalloc, dst, 2, 0.0, phi, 36, ED_PREC_F64, wn_big)
    assert rc == ED_ERR_ARG and b"scratch maximum" in L.ed_last_error()
    s.close(); dst.close()
    # a sector longer than the scratch's vectors (Ns=14 (7,7), dim 11,778,624)
    mE, (nuE, ndE) = ed.synthetic("E")
    sE = ed.Sector(mE, nuE, ndE)
    phiE = np.ones(sE.dim)
    rc, *_ = run(L.ed_dynamic_noalloc, sE, 1, 0.0, phiE, 36, ED_PREC_F64, wn)
    assert rc == ED_ERR_ARG and b"scratch maximum" in L.ed_last_error()
    sE.close()
