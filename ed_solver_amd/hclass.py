"""Python host mirror of the reference's Hamiltonian boundary and the solvers that call it.

Names, argument meaning and error behaviour follow the reference:

* ``Sector`` is one (N_up, N_dn) sector resident on the GPU -- the state that
  new_H / delete_H manage in H_class.f90:28-58 (module-global ``sector_h``) and
  HAIMupdo.f90:53-121 (tables), made a handle so several sectors can coexist.
* ``new_H`` / ``delete_H`` / ``Hmult__`` / ``dimen_H`` keep the one-global-sector
  surface of H_class.f90:13-152 for code written against it.
* ``Lanczos_fast_diagonalize`` / ``Lanczos_get_GS_sector`` mirror
  Lanczos_plain.f90:33-229, ``compute_dynamic`` mirrors
  green_class_compute_dynamic.f90:191-407, ``apply_C`` mirrors apply_C.f90:497-598,
  ``invert_zmtridiag`` mirrors matrices_tridiag.f90:163-198.

Where the reference STOPs, these raise ``EDError`` (``BreakdownError`` for a
zero-norm Lanczos vector).  Every call goes to libed_hip.so; nothing here computes
on the CPU beyond argument checks.
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np

from ._lib import (ED_PREC_F32, ED_PREC_F64, ED_PREC_MIXED_RQ, EDError, LanczosOpts, check, dptr, iptr, lib)
from .aim import AIMParams, CModel

__all__ = ["Sector", "LanczosResult", "new_H", "delete_H", "Hmult__", "dimen_H", "Lanczos_fast_diagonalize",
           "Lanczos_get_GS_sector", "apply_C", "compute_dynamic", "compute_dynamic_batched", "invert_zmtridiag",
           "matsubara", "start_vector", "tridiag_eig"]

_PREC = {"f64": ED_PREC_F64, "fp64": ED_PREC_F64, "f32": ED_PREC_F32, "fp32": ED_PREC_F32,
         "mixed_rq": ED_PREC_MIXED_RQ, ED_PREC_F64: ED_PREC_F64, ED_PREC_F32: ED_PREC_F32,
         ED_PREC_MIXED_RQ: ED_PREC_MIXED_RQ}


def _prec(p) -> int:
    try:
        return _PREC[p]
    except KeyError:
        raise ValueError(f"precision must be 'f64', 'f32' or 'mixed_rq', got {p!r}") from None


@dataclasses.dataclass
class LanczosResult:
    iters: int
    evals: np.ndarray        # lowest eigenvalues of T_iters
    vecp: np.ndarray         # (neig, iters) Lanczos coefficients
    diag: np.ndarray         # alpha_1..alpha_iters
    subdiag: np.ndarray      # ||v1||, beta_2..beta_iters
    precision: str
    iters_f32: int = 0       # FP32 phase of the mixed form (0 for FP64)
    steps_run: int = 0       # steps the device ran (>= the iterations: a persistent run stops a few steps late)


class Sector:
    """One (N_up, N_dn) sector of an Anderson impurity model on the current GPU."""

    def __init__(self, model: AIMParams, nup: int, ndn: int, slater: np.ndarray | None = None):
        """slater: the 4-index impurity interaction U(s1, s2, s3, s4) [nc]^4 of
        flag_slater_int (HAIMupdo.f90:905-999), added to the model's U / J terms."""
        self.model = model
        self.nup, self.ndn = int(nup), int(ndn)
        self._cm = CModel.from_params(model)
        h = ctypes.c_void_p()
        if slater is None:
            check(lib().ed_sector_create(ctypes.byref(self._cm), self.nup, self.ndn, ctypes.byref(h)))
        else:
            nc = model.nc
            self._u4 = np.ascontiguousarray(slater, np.float64).reshape(nc, nc, nc, nc)
            check(lib().ed_sector_create_slater(ctypes.byref(self._cm), dptr(self._u4), self.nup, self.ndn,
                                                ctypes.byref(h)))
        self._h = h
        self.dim = int(lib().ed_sector_dim(h))
        du, dd = ctypes.c_int32(), ctypes.c_int32()
        check(lib().ed_sector_spin_dims(h, ctypes.byref(du), ctypes.byref(dd)))
        self.dup, self.ddn = du.value, dd.value

    @property
    def handle(self) -> ctypes.c_void_p:
        if self._h is None:
            raise EDError(-5, "sector already destroyed (delete_H)")
        return self._h

    # host vectors of the Lanczos / apply_C / dynamic entry points: dim doubles here,
    # the planar (Re, Im) form of dim complex amplitudes for a ComplexSector
    @property
    def vlen(self) -> int:
        return self.dim

    def _pack(self, v: np.ndarray, what: str = "vector") -> np.ndarray:
        v = np.ascontiguousarray(v, np.float64)
        if v.shape != (self.dim,):
            raise EDError(-1, f"{what} of length {v.size}, sector dimension {self.dim}")
        return v

    def _unpack(self, v: np.ndarray) -> np.ndarray:
        return v

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            lib().ed_sector_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ----------------------------------------------------------------- tables
    @property
    def energy_shift(self) -> float:
        """Constant the particle-hole transformation removes from an S_z sector's H (0 otherwise)."""
        v = ctypes.c_double()
        check(lib().ed_sector_energy_shift(self.handle, ctypes.byref(v)))
        return v.value

    def nnz(self, spin: int) -> int:
        n = ctypes.c_int64()
        check(lib().ed_sector_nnz(self.handle, spin, ctypes.byref(n)))
        return n.value

    def tables(self, spin: int) -> dict:
        """tab_HAIM2 output of one spin (states, diag, noff, 0-based rankoff, offdiag)."""
        d = self.ddn if spin else self.dup
        nnz = self.nnz(spin)
        st = np.zeros(d, np.int32); dg = np.zeros(d); no = np.zeros(d, np.int32)
        rk = np.zeros(max(nnz, 1), np.int32); of = np.zeros(max(nnz, 1))
        check(lib().ed_sector_tables(self.handle, spin, iptr(st), dptr(dg), iptr(no), iptr(rk), dptr(of)))
        return {"states": st, "diag": dg, "noff": no, "rankoff": rk[:nnz], "offdiag": of[:nnz]}

    # -------------------------------------------------------------------- H.v
    def hmult(self, v: np.ndarray) -> np.ndarray:
        """Hmultr: host vectors in, host vector out (FP64)."""
        v = np.ascontiguousarray(v, dtype=np.float64)
        if v.shape != (self.dim,):
            raise EDError(-1, f"Hmult: vector of length {v.size}, sector dimension {self.dim}")
        out = np.empty_like(v)
        check(lib().ed_hmult_f64(self.handle, self.dim, dptr(out), dptr(v)))
        return out

    def hmult_dev(self, out_ptr: int, in_ptr: int, precision="f64", stream: int | None = None) -> None:
        """H.v on device pointers (e.g. torch tensor.data_ptr()); asynchronous."""
        p = _prec(precision)
        f = lib().ed_hmult_f64_dev if p == ED_PREC_F64 else lib().ed_hmult_f32_dev
        check(f(self.handle, self.dim, ctypes.c_void_p(out_ptr), ctypes.c_void_p(in_ptr),
                ctypes.c_void_p(stream) if stream else None))

    def synchronize(self) -> None:
        check(lib().ed_sector_synchronize(self.handle))

    def hmult_block_dev(self, out_ptr: int, in_ptr: int, nvec: int, ld: int, stream: int | None = None) -> None:
        check(lib().ed_hmult_block_f64_dev(self.handle, self.dim, nvec, ld, ctypes.c_void_p(out_ptr),
                                           ctypes.c_void_p(in_ptr), ctypes.c_void_p(stream) if stream else None))

    def tune(self) -> tuple[int, list]:
        """Runs the H.v schedule autotuner; returns (choice, ms per choice) with choices 1/2/4 (row
        kernel), 8 (fast kernel), 16 (fast kernel, XCD-grouped column order) and 18/19/20 (the
        latter with 2/3/4 rows per thread)."""
        c = ctypes.c_int()
        ms = np.zeros(8)
        check(lib().ed_hmult_tune(self.handle, ctypes.byref(c), dptr(ms)))
        return c.value, ms.tolist()

    def set_schedule(self, choice: int) -> None:
        check(lib().ed_hmult_set_schedule(self.handle, int(choice)))

    def lanczos_benchmark(self, precision="f64", niter: int = 128, reps: int = 5) -> tuple[float, float, bool]:
        """(ms per fixed-length run of niter steps, algorithmic bytes per run, persistent kernel used)."""
        ms, by, pk = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        check(lib().ed_lanczos_benchmark(self.handle, _prec(precision), int(niter), int(reps), ctypes.byref(ms),
                                         ctypes.byref(by), ctypes.byref(pk)))
        return ms.value, by.value, bool(pk.value)

    _PERSIST_FORMS = {"one-barrier": 1, "two-barrier": 4, "one-column": 3, "exact-beta": 5}

    def set_persistent(self, enable: bool, form: str = "one-barrier") -> None:
        """Use the persistent Lanczos kernel where it applies (True, the default) or never.
        form: "one-barrier" (the default: one grid barrier per step, alpha/beta equal to
        the two-kernel step's to rounding), "two-barrier" (the multi-column kernel with the
        two-kernel step's arithmetic, alpha/beta bit for bit), "one-column", or
        "exact-beta" (one-barrier with every beta from the reduced norm, a test hook)."""
        if form not in self._PERSIST_FORMS:
            raise ValueError(f"form must be one of {sorted(self._PERSIST_FORMS)}")
        check(lib().ed_lanczos_set_persistent(self.handle, self._PERSIST_FORMS[form] if enable else 0))

    def persistent_shape(self, precision="f64") -> dict:
        """Launch shape of the persistent kernel: columns per workgroup (0: not used),
        workgroups, rows per thread, table width."""
        v = [ctypes.c_int() for _ in range(4)]
        check(lib().ed_lanczos_persistent_shape(self.handle, _prec(precision), *[ctypes.byref(x) for x in v]))
        return {"cols_per_wg": v[0].value, "workgroups": v[1].value, "rows": v[2].value, "width": v[3].value}

    def persistent_refusals(self) -> int:
        """Cooperative launches of the persistent kernel the runtime refused on this sector."""
        v = ctypes.c_int()
        check(lib().ed_lanczos_persistent_refusals(self.handle, ctypes.byref(v)))
        return v.value

    def exact_beta_steps(self) -> int:
        """Steps of the last one-barrier persistent launch whose beta came from the reduced
        norm (the expansion of the norm had cancelled)."""
        v = ctypes.c_int()
        check(lib().ed_lanczos_exact_beta_steps(self.handle, ctypes.byref(v)))
        return v.value

    def persistent_active(self, precision="f64") -> bool:
        """Whether the next Lanczos run in `precision` uses the persistent kernel."""
        v = ctypes.c_int()
        check(lib().ed_lanczos_persistent_active(self.handle, _prec(precision), ctypes.byref(v)))
        return bool(v.value)

    def hmult_benchmark(self, precision="f64", reps: int = 50) -> tuple[float, float]:
        ms, by = ctypes.c_double(), ctypes.c_double()
        check(lib().ed_hmult_benchmark(self.handle, _prec(precision), reps, ctypes.byref(ms), ctypes.byref(by)))
        return ms.value, by.value

    # ---------------------------------------------------------------- Lanczos
    def lanczos(self, nitermax: int = 128, neigen: int = 1, tolerance: float = 1e-12, precision="f64",
                start: np.ndarray | None = None, check_every: int = 8, reorth: bool = False) -> LanczosResult:
        """Lanczos_fast_diagonalize on the device (convergence test as the reference)."""
        o = LanczosOpts(int(nitermax), int(neigen), float(tolerance), _prec(precision), int(check_every),
                        1 if reorth else 0)
        n = max(1, min(self.dim - 1, nitermax))
        neig_cap = max(1, min(neigen, self.dim, n))
        ev = np.zeros(neig_cap); vecp = np.zeros(neig_cap * n); dg = np.zeros(n); sd = np.zeros(n)
        it, ne = ctypes.c_int(), ctypes.c_int()
        sp = None
        if start is not None:
            start = self._pack(start, "start vector")
            sp = dptr(start)
        check(lib().ed_lanczos_diagonalize(self.handle, ctypes.byref(o), sp, ctypes.byref(it), ctypes.byref(ne),
                                           dptr(ev), dptr(vecp), dptr(dg), dptr(sd)))
        k, m = ne.value, it.value
        i32, nrun = ctypes.c_int(), ctypes.c_int()
        check(lib().ed_lanczos_info(self.handle, ctypes.byref(i32)))
        check(lib().ed_lanczos_steps_run(self.handle, ctypes.byref(nrun)))
        return LanczosResult(m, ev[:k].copy(), vecp[:k * m].reshape(k, m).copy(), dg[:m].copy(), sd[:m].copy(),
                             "f64" if o.precision == ED_PREC_F64 else "mixed", i32.value, nrun.value)

    def rq_info(self) -> tuple[bool, float, float]:
        """(accepted, squared residual, Ritz gap) of the last precision='mixed_rq' solve."""
        a, r2, g = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        check(lib().ed_lanczos_rq_info(self.handle, ctypes.byref(a), ctypes.byref(r2), ctypes.byref(g)))
        return bool(a.value), r2.value, g.value

    def ground_state(self, res: LanczosResult, k: int = 0) -> np.ndarray:
        """Lanczos_get_GS_sector: psi_k = normalised sum_j vecp[k, j] v_j over the basis kept on
        the device, or, on a sector that keeps none (a fresh one, as after the reference's
        second new_H), from the FP64 recurrence run again from the default start vector."""
        psi = np.zeros(self.vlen)
        row = np.ascontiguousarray(res.vecp[k], np.float64)
        check(lib().ed_lanczos_gs_vector(self.handle, res.iters, dptr(row), dptr(psi)))
        return self._unpack(psi)

    def full_ed(self, neig: int = 1):
        """ED_diago: lowest neig eigenpairs of the dense sector H (dim <= 1024; 512 complex)."""
        neig = int(min(neig, self.dim))
        ev = np.zeros(neig); vec = np.zeros(neig * self.vlen)
        check(lib().ed_full_ed(self.handle, neig, dptr(ev), dptr(vec)))
        vec = vec.reshape(neig, self.vlen)
        return ev, np.array([self._unpack(v) for v in vec]).reshape(neig, self.dim)

    def lanczos_run(self, niter: int, start: np.ndarray | None = None, precision="f64"):
        """Lanczos_GPU_updo: exactly niter steps, returns (diag, subdiag)."""
        dg = np.zeros(niter); sd = np.zeros(niter)
        sp = None
        if start is not None:
            start = self._pack(start, "start vector")
            sp = dptr(start)
        check(lib().ed_lanczos_run(self.handle, niter, sp, _prec(precision), dptr(dg), dptr(sd)))
        return dg, sd

    def rayleigh(self, psi: np.ndarray) -> float:
        psi = self._pack(psi)
        e = ctypes.c_double()
        check(lib().ed_rayleigh_f64(self.handle, self.vlen, dptr(psi), ctypes.byref(e)))
        return e.value


# --------------------------------------------------------------------------------
# One-global-sector surface of H_class.f90 (sector_h, :11)

_sector_h: Sector | None = None


def new_H(model: AIMParams, nup: int, ndn: int) -> Sector:
    """new_H (H_class.f90:28-37): builds the sector tables and makes it current."""
    global _sector_h
    delete_H()
    _sector_h = Sector(model, nup, ndn)
    return _sector_h


def delete_H() -> None:
    """delete_H (H_class.f90:48-58): no-op if H was never created."""
    global _sector_h
    if _sector_h is not None:
        _sector_h.close()
        _sector_h = None


def _current() -> Sector:
    if _sector_h is None:
        raise EDError(-5, "Hmultr, no array are associated!")   # H_class.f90:132
    return _sector_h


def Hmult__(n: int, vec_out: np.ndarray, vec_in: np.ndarray) -> None:
    """Hmultr (H_class.f90:109-136): vec_out(1:n) = H vec_in(1:n), in place."""
    s = _current()
    if n != s.dim or vec_out.shape != (n,) or vec_in.shape != (n,):
        raise EDError(-1, f"Hmult__: n={n} differs from dimen_H()={s.dim}")
    if vec_out.dtype != np.float64 or not vec_out.flags.c_contiguous:
        raise EDError(-1, "Hmult__: vec_out must be a contiguous float64 array")
    vin = np.ascontiguousarray(vec_in, np.float64)
    check(lib().ed_hmult_f64(s.handle, n, dptr(vec_out), dptr(vin)))


def dimen_H() -> int:
    """dimen_H (H_class.f90:149-152)."""
    return _current().dim


def Lanczos_fast_diagonalize(neigen: int = 1, nitermax: int = 128, tolerance: float = 1e-12, precision="f64",
                             start=None, sector: Sector | None = None, **kw) -> LanczosResult:
    """Lanczos_fast_diagonalize (Lanczos_plain.f90:33-134) on the current sector."""
    s = sector or _current()
    return s.lanczos(nitermax=nitermax, neigen=neigen, tolerance=tolerance, precision=precision, start=start, **kw)


def Lanczos_get_GS_sector(res: LanczosResult, k: int = 0, sector: Sector | None = None) -> np.ndarray:
    """Lanczos_get_GS_sector (Lanczos_plain.f90:156-229)."""
    s = sector or _current()
    return s.ground_state(res, k)


def start_vector(n: int) -> np.ndarray:
    out = np.zeros(n)
    check(lib().ed_start_vector(n, dptr(out)))
    return out


def tridiag_eig(diag, subdiag):
    d = np.ascontiguousarray(diag, np.float64); b = np.ascontiguousarray(subdiag, np.float64)
    n = len(d)
    w = np.zeros(n); z = np.zeros(n * n)
    check(lib().ed_tridiag_eig(n, dptr(d), dptr(b), dptr(w), dptr(z)))
    return w, z.reshape(n, n)


def tridiag_lowest(diag, subdiag, k: int, prev=None) -> np.ndarray:
    """The k lowest eigenvalues of T (the Lanczos convergence test's eigenvalues); prev:
    those of T with its last row and column removed (interlacing bounds), or None."""
    d = np.ascontiguousarray(diag, np.float64); b = np.ascontiguousarray(subdiag, np.float64)
    w = np.zeros(k)
    pp = None if prev is None else np.ascontiguousarray(prev, np.float64)
    if pp is not None and pp.shape != (k,):
        raise ValueError("prev must hold k values")
    check(lib().ed_tridiag_lowest(len(d), dptr(d), dptr(b), int(k), None if pp is None else dptr(pp), dptr(w)))
    return w


def band_eig(a: np.ndarray, b: int, nev: int, vectors: bool = True):
    """nev lowest eigenpairs of a symmetric band matrix (half-bandwidth b): the Ritz step
    of the block Lanczos driver.  Returns (w, z) with z of shape (nev, n)."""
    a = np.ascontiguousarray(a, np.float64)
    n = a.shape[0]
    w = np.zeros(nev)
    z = np.zeros((nev, n)) if vectors else None
    check(lib().ed_band_eig(n, int(b), dptr(a), int(nev), dptr(w), dptr(z) if vectors else None))
    return w, z


def apply_C(src: Sector, dst: Sector, site: int, spin: int, pm: int, psi: np.ndarray) -> np.ndarray:
    """apply_Cupdo (apply_C.f90:497-598): c+ (pm=+1) / c (pm=-1) on impurity orbital `site`.
    S_z sectors (apply_Csz, apply_C.f90:277-327): creates (pm=+1) or destroys the Nambu
    orbital (site, spin), npart -> npart + pm; the Nambu orbital of spin 1 is the
    particle-hole transformed down orbital, so pm=+1 there applies the physical c_dn."""
    psi = src._pack(psi, "apply_C: psi")
    phi = np.zeros(dst.vlen)
    check(lib().ed_apply_c(src.handle, dst.handle, site, spin, pm, dptr(psi), dptr(phi)))
    return dst._unpack(phi)


# flag or-ed into a Nambu orbital of dynamic_from_state's terms: the correlated
# excitation of apply_C_n for that term (ED_SITE_NCOND of the C-ABI)
SITE_NCOND = 0x10000


def apply_C_n(src: "SzSector", dst: "SzSector", site: int, spin: int, pm: int, psi: np.ndarray) -> np.ndarray:
    """apply_Csz_N_updo (apply_C.f90:404-483): apply_C on an S_z sector restricted by the
    other Nambu orbital of the site, the excitation of the correlated-hopping (FLAG_NCUP)
    Green's functions.  spin 0: the Nambu up orbital where the Nambu down orbital is empty
    (a physical down electron: n_dn c_up, c+_up n_dn); spin 1: the Nambu down orbital where
    the Nambu up orbital is occupied (n_up c+_dn, n_up c_dn).  pm as in apply_C."""
    psi = src._pack(psi, "apply_C_n: psi")
    phi = np.zeros(dst.vlen)
    check(lib().ed_apply_c_n(src.handle, dst.handle, site, spin, pm, dptr(psi), dptr(phi)))
    return dst._unpack(phi)


def matsubara(beta: float, nw: int) -> np.ndarray:
    """w_n = (2n-1) pi / beta, n = 1..nw (frequency_axis.f90:165-170)."""
    n = np.arange(1, nw + 1, dtype=np.float64)
    return (2.0 * np.pi / beta) * n - np.pi / beta


def invert_zmtridiag(a, b, wn) -> np.ndarray:
    """<0|(i wn - T)^-1|0> for every wn (matrices_tridiag.f90:163-198), on the device."""
    a = np.ascontiguousarray(a, np.float64); b = np.ascontiguousarray(b, np.float64)
    wn = np.ascontiguousarray(wn, np.float64)
    out = np.zeros(2 * len(wn))
    check(lib().ed_continued_fraction(len(a), dptr(a), dptr(b), len(wn), dptr(wn), dptr(out)))
    return out[0::2] + 1j * out[1::2]


class SzSector(Sector):
    """One S_z (Nambu) sector of an Anderson impurity model with a superconducting bath
    (new_HAIMsz, HAIMsz.f90:51-177): 2 Ns orbitals, the down spin particle-hole
    transformed, npart = Ns + 2 S_z.  pb[nb][nb] = bath pairing Pb, pvbc[2][nb][nc] =
    impurity-bath pairing (used when pairing_imp_to_bath).  With J, the Hund terms of
    HAIMsz_multc are included.  All H.v / Lanczos / dense-ED methods of Sector apply;
    physical energies are eigenvalues + energy_shift."""

    def __init__(self, model: AIMParams, npart: int, pb: np.ndarray | None = None, pvbc: np.ndarray | None = None,
                 pairing_imp_to_bath: bool = False):
        self.model = model
        self.npart = int(npart)
        self.nup, self.ndn = self.npart, -1
        self._cm = CModel.from_params(model)
        self._pb = None if pb is None else np.ascontiguousarray(pb, np.float64).reshape(model.nb, model.nb)
        self._pvbc = None if pvbc is None else np.ascontiguousarray(pvbc, np.float64).reshape(2, model.nb, model.nc)
        h = ctypes.c_void_p()
        check(lib().ed_sector_create_sz(ctypes.byref(self._cm), dptr(self._pb) if self._pb is not None else None,
                                        dptr(self._pvbc) if self._pvbc is not None else None,
                                        1 if pairing_imp_to_bath else 0, self.npart, ctypes.byref(h)))
        self._h = h
        self.dim = int(lib().ed_sector_dim(h))
        self.dup, self.ddn = self.dim, 1

    def quart(self) -> np.ndarray:
        """QUART_INT_SZ per state."""
        q = np.zeros(self.dim)
        check(lib().ed_sector_quart_sz(self.handle, dptr(q)))
        return q

    def set_fly(self, on: bool) -> None:
        """ON_FLY: generate the H.v entries on the fly (HAIMsz_mult_fly) instead of the table."""
        check(lib().ed_sector_set_fly(self.handle, 1 if on else 0))

    def hund(self):
        """Hund transitions of HAIMsz_multc as (rk[kh, dim], val[kh, dim]) (kh = 0 without J)."""
        kh = ctypes.c_int()
        check(lib().ed_sector_hund_sz(self.handle, ctypes.byref(kh), None, None))
        rk = np.zeros((kh.value, self.dim), np.int32); val = np.zeros((kh.value, self.dim))
        if kh.value:
            check(lib().ed_sector_hund_sz(self.handle, ctypes.byref(kh), iptr(rk), dptr(val)))
        return rk, val


class ComplexSector(Sector):
    """One (N_up, N_dn) sector of a complex AIM (tab_HAIM2 with COMPLEX(8) entries and
    HAIMupdo_multc, HAIMupdo.f90:150-656).  model holds the real parts of Ec, Eb, Vbc;
    ec_im[2, nc, nc], eb_im[2, nb, nb], vbc_im[2, nb, nc] the imaginary parts.  Vectors
    are complex numpy arrays of dim amplitudes; the device holds them planar (Re, Im),
    where H acts as [[A, -B], [B, A]] (H = A + iB) and the plain Lanczos has the complex
    Lanczos' real alpha and beta.  FP64 only; dense ED (Hermitian, from the real
    embedding) up to dimension 512; no block Lanczos or gradients."""

    def __init__(self, model: AIMParams, nup: int, ndn: int, ec_im, eb_im=None, vbc_im=None):
        self.model = model
        self.nup, self.ndn = int(nup), int(ndn)
        self._cm = CModel.from_params(model)
        nc, nb = model.nc, model.nb
        self._ec_im = np.ascontiguousarray(np.broadcast_to(ec_im, (2, nc, nc)), np.float64)
        self._eb_im = np.ascontiguousarray(np.zeros((2, nb, nb)) if eb_im is None
                                           else np.broadcast_to(eb_im, (2, nb, nb)), np.float64)
        self._vbc_im = np.ascontiguousarray(np.zeros((2, nb, nc)) if vbc_im is None
                                            else np.broadcast_to(vbc_im, (2, nb, nc)), np.float64)
        h = ctypes.c_void_p()
        check(lib().ed_sector_create_c(ctypes.byref(self._cm), dptr(self._ec_im), dptr(self._eb_im),
                                       dptr(self._vbc_im), self.nup, self.ndn, ctypes.byref(h)))
        self._h = h
        self.dim = int(lib().ed_sector_dim(h))
        du, dd = ctypes.c_int32(), ctypes.c_int32()
        check(lib().ed_sector_spin_dims(h, ctypes.byref(du), ctypes.byref(dd)))
        self.dup, self.ddn = du.value, dd.value

    @property
    def vlen(self) -> int:
        return 2 * self.dim

    def _pack(self, v: np.ndarray, what: str = "vector") -> np.ndarray:
        v = np.asarray(v)
        if v.shape != (self.dim,):
            raise EDError(-1, f"{what} of length {v.size}, sector dimension {self.dim}")
        return np.ascontiguousarray(np.concatenate([v.real, v.imag]), np.float64)

    def _unpack(self, v: np.ndarray) -> np.ndarray:
        return v[:self.dim] + 1j * v[self.dim:]

    def tables_im(self, spin: int) -> np.ndarray:
        """Imaginary parts of one spin's offdiag (in the order of tables(spin))."""
        nnz = self.nnz(spin)
        im = np.zeros(max(nnz, 1))
        check(lib().ed_sector_tables_im(self.handle, spin, dptr(im)))
        return im[:nnz]

    def hmult(self, v: np.ndarray) -> np.ndarray:
        """Hmultc: complex host vector in, complex host vector out."""
        v = np.ascontiguousarray(v, np.complex128)
        if v.shape != (self.dim,):
            raise EDError(-1, f"Hmult: vector of length {v.size}, sector dimension {self.dim}")
        out = np.empty_like(v)
        check(lib().ed_hmult_c128(self.handle, self.dim, dptr(out.view(np.float64)), dptr(v.view(np.float64))))
        return out

    def hmult_c128_dev(self, out_ptr: int, in_ptr: int, stream: int | None = None) -> None:
        """Hmultc on device COMPLEX(8) vectors (e.g. torch.complex128 data_ptr()); asynchronous."""
        check(lib().ed_hmult_c128_dev(self.handle, self.dim, ctypes.c_void_p(out_ptr), ctypes.c_void_p(in_ptr),
                                      ctypes.c_void_p(stream) if stream else None))


class ComplexSzSector(ComplexSector):
    """One S_z (Nambu) sector of a complex model with a superconducting bath (the COMPLEX(8)
    new_HAIMsz / HAIMsz_multc, HAIMsz.f90:188-452): model holds the real parts, im the
    imaginary parts as for ComplexSector ({"ec": [2, nc, nc], "eb": [2, nb, nb], "vbc":
    [2, nb, nc]}); pb[nb][nb] and pvbc[2][nb][nc] may be complex.  Vectors are complex numpy
    arrays (planar on the device), FP64 only; physical energies are eigenvalues +
    energy_shift."""

    def __init__(self, model: AIMParams, npart: int, im: dict, pb=None, pvbc=None, pairing_imp_to_bath: bool = False):
        self.model = model
        self.npart = int(npart)
        self.nup, self.ndn = self.npart, -1
        self._cm = CModel.from_params(model)
        nc, nb = model.nc, model.nb
        self._ec_im = np.ascontiguousarray(np.broadcast_to(im["ec"], (2, nc, nc)), np.float64)
        self._eb_im = np.ascontiguousarray(np.broadcast_to(im.get("eb", np.zeros((2, nb, nb))), (2, nb, nb)),
                                           np.float64)
        self._vbc_im = np.ascontiguousarray(np.broadcast_to(im.get("vbc", np.zeros((2, nb, nc))), (2, nb, nc)),
                                            np.float64)
        keep = {}
        for name, a, shape in (("pb", pb, (nb, nb)), ("pvbc", pvbc, (2, nb, nc))):
            if a is None:
                keep[name] = keep[name + "_im"] = None
            else:
                a = np.asarray(a).reshape(shape)
                keep[name] = np.ascontiguousarray(a.real, np.float64)
                keep[name + "_im"] = np.ascontiguousarray(a.imag, np.float64)
        self._pair = keep
        p = lambda a: dptr(a) if a is not None else None
        h = ctypes.c_void_p()
        check(lib().ed_sector_create_sz_c(ctypes.byref(self._cm), dptr(self._ec_im), dptr(self._eb_im),
                                          dptr(self._vbc_im), p(keep["pb"]), p(keep["pb_im"]), p(keep["pvbc"]),
                                          p(keep["pvbc_im"]), 1 if pairing_imp_to_bath else 0, self.npart,
                                          ctypes.byref(h)))
        self._h = h
        self.dim = int(lib().ed_sector_dim(h))
        self.dup, self.ddn = self.dim, 1

    quart = SzSector.quart
    set_fly = SzSector.set_fly
    hund = SzSector.hund


def compute_dynamic(sector: Sector, iph: int, energy: float, phi: np.ndarray, nitergreen: int, wn: np.ndarray,
                    cutoff_dynamic: float = 1e-8, precision="f64"):
    """compute_dynamic_cpu (green_class_compute_dynamic.f90:191-407): returns (dyn, a, b)."""
    if np.shape(phi) != (sector.dim,):
        raise EDError(-1, "compute_dynamic inconsistent dimensions, CRITICAL")
    phi = sector._pack(phi)
    wn = np.ascontiguousarray(wn, np.float64)
    n = max(1, min(nitergreen, sector.dim))
    dyn = np.zeros(2 * len(wn)); a = np.zeros(n); b = np.zeros(n)
    check(lib().ed_dynamic(sector.handle, iph, energy, dptr(phi), nitergreen, cutoff_dynamic, _prec(precision),
                           len(wn), dptr(wn), dptr(dyn), dptr(a), dptr(b)))
    return dyn[0::2] + 1j * dyn[1::2], a, b


def dynamic_full_ed(iph: int, energy: float, phi: np.ndarray, evals: np.ndarray, evecs: np.ndarray, wn: np.ndarray,
                    cutoff_dynamic: float = 1e-8) -> np.ndarray:
    """compute_dynamic_full_ed: Lehmann sum over the given eigenpairs of a small target
    sector (evecs one state per row); returns dyn(i w_n) complex."""
    cplx = np.iscomplexobj(phi) or np.iscomplexobj(evecs)
    phi = np.asarray(phi)
    dim = phi.size
    ev = np.ascontiguousarray(evals, np.float64).reshape(-1)
    vec = np.asarray(evecs).reshape(len(ev), -1) if len(ev) else np.zeros((0, dim))
    if len(ev) and vec.shape[1] != dim:
        raise EDError(-1, "dynamic_full_ed: eigenvector length differs from phi")
    if cplx:   # planar (Re, Im) rows for the complex entry point
        phi = np.concatenate([phi.real, phi.imag])
        vec = np.concatenate([vec.real, vec.imag], axis=1)
    phi = np.ascontiguousarray(phi, np.float64)
    vec = np.ascontiguousarray(vec, np.float64)
    wn = np.ascontiguousarray(wn, np.float64)
    out = np.zeros(2 * len(wn))
    f = lib().ed_dynamic_full_ed_c if cplx else lib().ed_dynamic_full_ed
    check(f(dim, int(iph), float(energy), dptr(phi), len(ev), dptr(ev) if len(ev) else None,
            dptr(vec) if len(ev) else None, float(cutoff_dynamic), len(wn), dptr(wn), dptr(out)))
    return out[0::2] + 1j * out[1::2]


def dynamic_from_state(gs: Sector, psi: np.ndarray, energy: float, jobs, spin: int, nitergreen: int,
                       wn: np.ndarray, cutoff_dynamic: float = 1e-8, precision="f64") -> np.ndarray:
    """The excitation runs of one eigenstate with every vector on the device: jobs is a list
    of (dst_sector, iph, [(site, coef), ...]); returns dyn (njobs, nw) complex.  For an
    S_z sector a site is the Nambu orbital spin * nc + site and `spin` is unused."""
    psi = gs._pack(psi, "dynamic_from_state: psi")
    nj = len(jobs)
    wn = np.ascontiguousarray(wn, np.float64)
    ds = (ctypes.c_void_p * nj)(*[j[0].handle.value for j in jobs])
    iph = np.array([int(j[1]) for j in jobs], np.int32)
    nterm = np.array([len(j[2]) for j in jobs], np.int32)
    sites = np.array([int(c) for j in jobs for c, _ in j[2]], np.int32)
    coefs = np.array([float(a) for j in jobs for _, a in j[2]], np.float64)
    dyn = np.zeros(2 * nj * len(wn))
    check(lib().ed_dynamic_from_state(gs.handle, dptr(psi), float(energy), nj, ds, iptr(iph), iptr(nterm), iptr(sites),
                                      dptr(coefs), int(spin), int(nitergreen), float(cutoff_dynamic), _prec(precision),
                                      len(wn), dptr(wn), dptr(dyn)))
    d = dyn.reshape(nj, len(wn), 2)
    return d[..., 0] + 1j * d[..., 1]


def dynamic_from_states(states, jobs, spin: int, nitergreen: int, wn: np.ndarray, cutoff_dynamic: float = 1e-8,
                        precision="f64") -> np.ndarray:
    """The excitation runs of several eigenstates in one call: states is a list of
    (sector, psi, energy), jobs a list of (state index, dst_sector, iph, [(site, coef), ...]);
    returns dyn (njobs, nw) complex, unweighted."""
    nst, nj = len(states), len(jobs)
    psis = [st[0]._pack(st[1], "dynamic_from_states: psi") for st in states]
    gs = (ctypes.c_void_p * nst)(*[st[0].handle.value for st in states])
    pp = (ctypes.POINTER(ctypes.c_double) * nst)(*[dptr(p) for p in psis])
    en = np.array([float(st[2]) for st in states])
    wn = np.ascontiguousarray(wn, np.float64)
    sof = np.array([int(j[0]) for j in jobs], np.int32)
    ds = (ctypes.c_void_p * nj)(*[j[1].handle.value for j in jobs])
    iph = np.array([int(j[2]) for j in jobs], np.int32)
    nterm = np.array([len(j[3]) for j in jobs], np.int32)
    sites = np.array([int(c) for j in jobs for c, _ in j[3]], np.int32)
    coefs = np.array([float(a) for j in jobs for _, a in j[3]], np.float64)
    dyn = np.zeros(2 * nj * len(wn))
    check(lib().ed_dynamic_from_states(nst, gs, pp, dptr(en), nj, iptr(sof), ds, iptr(iph), iptr(nterm), iptr(sites),
                                       dptr(coefs), int(spin), int(nitergreen), float(cutoff_dynamic),
                                       _prec(precision), len(wn), dptr(wn), dptr(dyn)))
    d = dyn.reshape(nj, len(wn), 2)
    return d[..., 0] + 1j * d[..., 1]


def lanczos_batched(sectors, niter: int, precision="f64"):
    """Lanczos_GPU_updo for many sectors at once (ed_lanczos_batched): niter fixed steps from
    each sector's default start vector, all advancing together.  Returns (alpha, beta,
    steps) with alpha, beta of shape (nsec, niter) and steps[k] the completed steps."""
    n = len(sectors)
    hs = (ctypes.c_void_p * n)(*[s.handle.value for s in sectors])
    a = np.zeros((n, niter)); b = np.zeros((n, niter)); st = np.zeros(n, np.int32)
    check(lib().ed_lanczos_batched(n, hs, int(niter), _prec(precision), dptr(a), dptr(b), iptr(st)))
    return a, b, st


def lanczos_converged(alpha, beta, neigen: int = 1, tolerance: float = 1e-12, nitermax: int | None = None):
    """Lanczos_fast_diagonalize's convergence test replayed on recorded coefficients:
    (iterations, the neigen lowest eigenvalues of T at that iteration)."""
    alpha = np.ascontiguousarray(alpha, np.float64); beta = np.ascontiguousarray(beta, np.float64)
    it = ctypes.c_int()
    ev = np.zeros(neigen)
    check(lib().ed_lanczos_converged(len(alpha), dptr(alpha), dptr(beta), int(neigen), float(tolerance),
                                     int(nitermax or len(alpha)), ctypes.byref(it), dptr(ev)))
    return it.value, ev


def batched_stats(last=False):
    """(launches of the multi-run persistent kernel, launches the runtime refused) of the
    batched excitation runs since the library was loaded; with last=True also a dict of the
    last launch: its duration in ms (HIP events), job-steps and algorithmic bytes."""
    n, r = ctypes.c_int(), ctypes.c_int()
    v = [ctypes.c_double() for _ in range(3)]
    check(lib().ed_dynamic_batched_stats(ctypes.byref(n), ctypes.byref(r), *[ctypes.byref(x) for x in v]))
    if not last:
        return n.value, r.value
    return n.value, r.value, {"ms": v[0].value, "job_steps": v[1].value, "bytes": v[2].value}


def compute_dynamic_batched(sectors, phis, iphs, energies, nitergreen: int, wn: np.ndarray,
                            cutoff_dynamic: float = 1e-8, precision="f64"):
    """All excitation runs in one call; returns dyn (njobs, nw) complex."""
    nj = len(sectors)
    phis = [s._pack(p, "compute_dynamic_batched: phi") for s, p in zip(sectors, phis)]
    wn = np.ascontiguousarray(wn, np.float64)
    hs = (ctypes.c_void_p * nj)(*[s.handle.value for s in sectors])
    ps = (ctypes.POINTER(ctypes.c_double) * nj)(*[dptr(p) for p in phis])
    ip = (ctypes.c_int * nj)(*[int(i) for i in iphs])
    en = np.ascontiguousarray(energies, np.float64)
    dyn = np.zeros(nj * 2 * len(wn)); a = np.zeros(nj * nitergreen); b = np.zeros(nj * nitergreen)
    check(lib().ed_dynamic_batched(nj, hs, ps, ip, dptr(en), nitergreen, cutoff_dynamic, _prec(precision), len(wn),
                                   dptr(wn), dptr(dyn), dptr(a), dptr(b)))
    d = dyn.reshape(nj, len(wn), 2)
    return d[..., 0] + 1j * d[..., 1], a.reshape(nj, nitergreen), b.reshape(nj, nitergreen)
