"""CPU checks of the C-ABI library: it loads, exports every symbol include/ed_hip.h
declares, and the host-side argument checks raise before any device work."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ed_hip.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ed_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_expected_surface():
    syms = declared_symbols()
    for must in ("ed_sector_create", "ed_sector_destroy", "ed_hmult_f64", "ed_sector_dim", "ed_lanczos_diagonalize",
                 "ed_lanczos_gs_vector", "ed_lanczos_run", "ed_dynamic", "ed_continued_fraction", "ed_apply_c"):
        assert must in syms


def test_library_exports_every_declared_symbol():
    import ed_solver_amd as ed
    L = ed.lib()
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_python_binding_covers_header():
    from ed_solver_amd._lib import SIGNATURES
    assert sorted(SIGNATURES) == declared_symbols()


def test_fortran_binding_covers_header():
    """module ed_hip (ISO_C_BINDING) has one bind(C) interface per C-ABI entry point."""
    src = open(os.path.join(ROOT, "ed_solver_amd", "fortran", "ed_hip_mod.f90")).read()
    bound = set(re.findall(r'bind\(C,\s*name="(ed_[a-z0-9_]+)"\)', src))
    assert sorted(bound) == declared_symbols()


def test_version_and_defaults():
    import ed_solver_amd as ed
    from ed_solver_amd._lib import LanczosOpts
    assert ed.lib().ed_version().startswith(b"ed_hip")
    o = LanczosOpts()
    assert ed.lib().ed_lanczos_default_opts(ctypes.byref(o)) == 0
    assert (o.nitermax, o.neigen, o.tolerance, o.precision) == (128, 1, 1e-12, 0)


def test_start_vector_host_entry_matches_oracle():
    import ed_solver_amd as ed
    import oracle
    for n in (1, 7, 10000, 25001):
        a = ed.start_vector(n)
        b = oracle.start_vector(n)
        assert np.array_equal(a, b)
        assert a.min() > 0 and a.max() <= 1.0


def test_tridiag_eig_vs_scipy():
    import scipy.linalg as la
    import ed_solver_amd as ed
    rng = np.random.default_rng(0)
    for n in (1, 2, 5, 40, 128):
        a = rng.uniform(-3, 3, n)
        b = np.concatenate([[7.0], rng.uniform(0.01, 1.5, n - 1)])   # b[0] is ||v1||, not referenced
        w, z = ed.tridiag_eig(a, b)
        if n == 1:
            assert w[0] == a[0]
            continue
        w_ref, z_ref = la.eigh_tridiagonal(a, b[1:])
        assert np.max(np.abs(w - w_ref)) < 1e-12 * max(1, np.abs(w_ref).max())
        T = np.diag(a) + np.diag(b[1:], 1) + np.diag(b[1:], -1)
        assert np.max(np.abs(T @ z.T - z.T * w)) < 1e-11
        assert np.max(np.abs(z @ z.T - np.eye(n))) < 1e-12


def _lanczos_T(H, n, seed=1):
    """alpha, beta (beta[0] = ||v1||) of n plain Lanczos steps on a dense symmetric H."""
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(H.shape[0])
    a, b = np.zeros(n), np.zeros(n)
    b[0] = np.linalg.norm(v)
    v /= b[0]
    vm = np.zeros_like(v)
    for j in range(n):
        w = H @ v - (b[j] * vm if j else 0.0)
        a[j] = w @ v
        w -= a[j] * v
        if j + 1 < n:
            b[j + 1] = np.linalg.norm(w)
            vm, v = v, w / b[j + 1]
    return a, b


def test_tridiag_lowest_warm_sequence_vs_scipy():
    """The convergence test's eigenvalues: every T_n of a Lanczos run (n = 1..90, Ritz
    values converging and then flat to machine precision) from the values of T_{n-1},
    against scipy's; also with no previous values and with wrong ones (too low, too
    high), which only cost speed."""
    import scipy.linalg as la
    import ed_solver_amd as ed
    rng = np.random.default_rng(3)
    Q, _ = np.linalg.qr(rng.standard_normal((300, 300)))
    lam = np.concatenate([[-14.7, -14.2, -14.2], np.linspace(-13, 9, 297)])
    H = (Q * lam) @ Q.T
    a, b = _lanczos_T(H, 90)
    for k in (1, 3):
        prev = np.full(k, 1e300)
        for n in range(1, 91):
            w = ed.tridiag_lowest(a[:n], b[:n], k, prev)
            ref = la.eigvalsh_tridiagonal(a[:n], b[1:n]) if n > 1 else a[:1]
            m = min(k, n)
            scale = max(1.0, np.abs(ref).max())
            assert np.max(np.abs(w[:m] - ref[:m])) < 16 * np.finfo(float).eps * scale * 10, (k, n)
            assert np.all(w[m:] >= 1e299)
            assert np.array_equal(w, ed.tridiag_lowest(a[:n], b[:n], k, prev))   # deterministic
            cold = ed.tridiag_lowest(a[:n], b[:n], k, None)
            assert np.max(np.abs(w[:m] - cold[:m])) < 16 * np.finfo(float).eps * scale * 10
            if n > k:
                for bad in (ref[:k] - 1.0, ref[:k] + 5.0, np.full(k, -1e6)):
                    wb = ed.tridiag_lowest(a[:n], b[:n], k, bad)
                    assert np.max(np.abs(wb - ref[:k])) < 16 * np.finfo(float).eps * scale * 10, (k, n, bad)
            prev = np.where(np.arange(k) < m, w, 1e300)


def test_tridiag_lowest_degenerate_and_split():
    """Exactly degenerate lowest eigenvalues (zero off-diagonals split T into blocks, a
    repeated level) and a zero pivot on the way (d_0 - x = 0 at the start)."""
    import ed_solver_amd as ed
    d = np.array([1.0, 1.0, 3.0, 1.0, 2.0])
    b = np.array([9.0, 0.0, 0.5, 0.0, 0.0])
    import scipy.linalg as la
    ref = la.eigvalsh_tridiagonal(d, b[1:])
    for prev in (None, np.array([1.0, 1.0, 1.2]), np.array([1.5, 1.5, 1.5])):
        w = ed.tridiag_lowest(d, b, 3, prev)
        assert np.max(np.abs(w - ref[:3])) < 1e-14, (prev, w, ref)
    w = ed.tridiag_lowest(np.array([0.0, 0.0]), np.array([1.0, 1.0]), 1, np.array([0.0]))
    assert abs(w[0] + 1.0) < 1e-15


def test_band_eig_vs_lapack():
    """The host Ritz step of block Lanczos: band reduction + QL + inverse iteration against
    numpy.linalg.eigh, on random band matrices and on a matrix with exactly degenerate
    eigenvalues (two identical diagonal blocks)."""
    import ed_solver_amd as ed
    rng = np.random.default_rng(1)
    cases = []
    for n, b in ((1, 0), (6, 1), (40, 3), (200, 8), (64, 16), (96, 8)):
        A = rng.normal(size=(n, n))
        A = np.triu(np.tril(A + A.T, b), -b)
        cases.append((A, b))
    B = rng.normal(size=(50, 50))
    B = np.triu(np.tril(B + B.T, 2), -2)
    A = np.zeros((100, 100)); A[:50, :50] = B; A[50:, 50:] = B
    cases.append((A, 4))
    for A, b in cases:
        n = A.shape[0]
        nev = min(8, n)
        w, z = ed.band_eig(A, b, nev)
        ref = np.linalg.eigvalsh(A)[:nev]
        scale = max(1.0, np.abs(ref).max())
        assert np.max(np.abs(w - ref)) < 1e-12 * scale
        assert np.max(np.linalg.norm(A @ z.T - z.T * w, axis=0)) < 1e-11 * scale
        assert np.max(np.abs(z @ z.T - np.eye(nev))) < 1e-11
    # degenerate pairs are really there and resolved
    w, _ = ed.band_eig(cases[-1][0], 4, 8)
    assert np.max(np.abs(w[0::2] - w[1::2])) < 1e-12


def test_errors_raise_without_device_work():
    import ed_solver_amd as ed
    from ed_solver_amd._lib import LanczosOpts, check
    with pytest.raises(ed.EDError):
        check(ed.lib().ed_lanczos_default_opts(None))
    with pytest.raises(ed.EDError):
        check(ed.lib().ed_start_vector(0, None))
    with pytest.raises(ed.EDError):
        ed.invert_zmtridiag(np.zeros(0), np.zeros(0), np.zeros(3))
    with pytest.raises(ValueError):
        ed.aim.AIMParams(nc=3, nb=13, ec=np.zeros((2, 3, 3)), eb=np.zeros((2, 13, 13)), vbc=np.zeros((2, 13, 3)),
                         u=np.zeros((3, 3)), jj=np.zeros((3, 3)), jhund=0.0)


def test_package_has_no_cpu_fallback():
    """The product never imports the oracle."""
    pkg = os.path.join(ROOT, "ed_solver_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith(".py"):
                src = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in src and "from oracle" not in src, f
