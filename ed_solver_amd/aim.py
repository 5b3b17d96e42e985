"""Anderson-impurity model parameters and the synthetic workloads.

The parameter layout follows the reference's AIM / AIM2 types
(ed_solver/quantum_algebra_Anderson_Impurity_class.f90:59-87 and
quantum_algebra_Anderson_Impurity_class2.f90:62-135):

* per spin sigma: impurity levels Ec(sigma) [nc x nc], bath levels Eb(sigma)
  [nb x nb], hybridisation Vbc(sigma) [nb x nc] (bath row, impurity column);
* per-spin bit layout: bath site k -> bit k, impurity orbital c -> bit nb + c;
* interaction matrix U [nc x nc] with U_ii = U and U_ij = U - 2J for i != j (the
  reference reduces the off-diagonal U by 2J for Slater-type Hund coupling,
  ed_solver/dmft_ed_solver.f90:731-741), Hund matrix JJ [nc x nc] and the scalar
  Jhund that switches the spin-flip / pair-hopping terms on when |Jhund| > 1e-5
  (ed_solver/HAIMupdo.f90:1017, 1087);
* cutoff_hamilt_param: off-diagonal one-body entries with |t| <= cutoff are not
  tabulated (ed_solver/HAIM_offdiag_arrays.f90:432, 458, 483).

The synthetic workloads draw the bath from ``numpy.random.default_rng(seed)``:
first eps_k, then V_k; bath site k couples only to orbital k mod nc.  The frozen
draws live in tests/golden/models.json and a test checks this generator still
reproduces them bit for bit.
"""
from __future__ import annotations

import ctypes
import dataclasses
from math import comb

import numpy as np

__all__ = ["AIMParams", "kanamori", "synthetic", "WORKLOADS", "sector_dims"]


@dataclasses.dataclass
class AIMParams:
    nc: int
    nb: int
    ec: np.ndarray   # (2, nc, nc)
    eb: np.ndarray   # (2, nb, nb)
    vbc: np.ndarray  # (2, nb, nc)
    u: np.ndarray    # (nc, nc)
    jj: np.ndarray   # (nc, nc)
    jhund: float
    cutoff: float = 1e-8

    def __post_init__(self):
        self.ec = np.ascontiguousarray(np.asarray(self.ec, dtype=np.float64).reshape(2, self.nc, self.nc))
        self.eb = np.ascontiguousarray(np.asarray(self.eb, dtype=np.float64).reshape(2, self.nb, self.nb))
        self.vbc = np.ascontiguousarray(np.asarray(self.vbc, dtype=np.float64).reshape(2, self.nb, self.nc))
        self.u = np.ascontiguousarray(np.asarray(self.u, dtype=np.float64).reshape(self.nc, self.nc))
        self.jj = np.ascontiguousarray(np.asarray(self.jj, dtype=np.float64).reshape(self.nc, self.nc))
        self.jhund = float(self.jhund)
        self.cutoff = float(self.cutoff)
        if self.nc < 1 or self.nb < 0 or self.nc + self.nb > 15:
            raise ValueError("need 1 <= nc and nc + nb <= 15 orbitals per spin")

    @property
    def ns(self) -> int:
        return self.nc + self.nb

    def to_json(self) -> dict:
        return {
            "nc": self.nc, "nb": self.nb, "ec": self.ec.tolist(), "eb": self.eb.tolist(),
            "vbc": self.vbc.tolist(), "u": self.u.tolist(), "jj": self.jj.tolist(),
            "jhund": self.jhund, "cutoff": self.cutoff,
        }

    @classmethod
    def from_json(cls, d: dict) -> "AIMParams":
        return cls(nc=d["nc"], nb=d["nb"], ec=d["ec"], eb=d["eb"], vbc=d["vbc"], u=d["u"],
                   jj=d["jj"], jhund=d["jhund"], cutoff=d.get("cutoff", 1e-8))

    def with_(self, **kw) -> "AIMParams":
        d = dataclasses.asdict(self)
        d.update(kw)
        return AIMParams(**d)


def kanamori(nc: int, nb: int, U: float, J: float, eps_k, v_k, eimp: float | None = None,
             cutoff: float = 1e-8) -> AIMParams:
    """Kanamori impurity (U, U' = U - 2J, J) with a diagonal bath, spin-independent.

    Bath site k couples to orbital k mod nc with amplitude v_k.  The default
    impurity level is the half-filling value -[U/2 + (nc-1)(2U'-J)/2].
    """
    eps_k = np.asarray(eps_k, dtype=np.float64)
    v_k = np.asarray(v_k, dtype=np.float64)
    assert eps_k.shape == (nb,) and v_k.shape == (nb,)
    up = U - 2.0 * J
    if eimp is None:
        eimp = -(U / 2.0 + (nc - 1) * (2.0 * up - J) / 2.0)
    ec1 = eimp * np.eye(nc)
    eb1 = np.diag(eps_k) if nb else np.zeros((0, 0))
    vbc1 = np.zeros((nb, nc))
    for k in range(nb):
        vbc1[k, k % nc] = v_k[k]
    u = np.full((nc, nc), up)
    np.fill_diagonal(u, U)
    jj = np.full((nc, nc), J)
    return AIMParams(nc=nc, nb=nb, ec=np.stack([ec1, ec1]), eb=np.stack([eb1, eb1]),
                     vbc=np.stack([vbc1, vbc1]), u=u, jj=jj, jhund=J, cutoff=cutoff)


# name -> (nc, nb, U, J, eps range, V range, seed, sector)
WORKLOADS = {
    "A": dict(nc=1, nb=5, U=2.0, J=0.0, eps=(-1.0, 1.0), v=(0.2, 0.6), seed=1001, sector=(3, 3)),
    "B": dict(nc=2, nb=8, U=3.0, J=0.5, eps=(-2.0, 2.0), v=(0.1, 0.5), seed=1002, sector=(5, 5)),
    "C": dict(nc=3, nb=9, U=2.5, J=0.4, eps=(-2.0, 2.0), v=(0.1, 0.5), seed=1003, sector=(6, 6)),
    "E": dict(nc=2, nb=12, U=3.0, J=0.5, eps=(-2.0, 2.0), v=(0.1, 0.5), seed=1005, sector=(7, 7)),
}


def slater_tensor(model: AIMParams) -> np.ndarray:
    """The model's density, Hund exchange and pair-hopping terms as a Slater tensor
    U(s1, s2, s3, s4) of flag_slater_int (HAIMupdo.f90:905-999): U_aa / 2 at (a, a, a, a),
    U_ab / 2 at (a, b, b, a), J_ab / 2 at (b, a, b, a) and (b, b, a, a), a != b.  A model
    with U = J = 0 plus this tensor has the spectrum of the model itself (spin flip and
    pair hopping come from the exchange and pair entries)."""
    nc = model.nc
    u4 = np.zeros((nc, nc, nc, nc))
    for a in range(nc):
        u4[a, a, a, a] = model.u[a, a] / 2
        for b in range(nc):
            if a != b:
                u4[a, b, b, a] = model.u[a, b] / 2
                u4[b, a, b, a] = model.jj[a, b] / 2
                u4[b, b, a, a] = model.jj[a, b] / 2
    return u4


def synthetic(name: str) -> tuple[AIMParams, tuple[int, int]]:
    """The seeded model of workload `name` and its headline (N_up, N_dn) sector."""
    w = WORKLOADS[name]
    rng = np.random.default_rng(w["seed"])
    eps = rng.uniform(w["eps"][0], w["eps"][1], size=w["nb"])
    v = rng.uniform(w["v"][0], w["v"][1], size=w["nb"])
    return kanamori(w["nc"], w["nb"], w["U"], w["J"], eps, v), w["sector"]


def sector_dims(ns: int, nup: int, ndn: int) -> tuple[int, int, int]:
    du, dd = comb(ns, nup), comb(ns, ndn)
    return du, dd, du * dd


class CModel(ctypes.Structure):
    """C mirror of ``ed_model`` (include/ed_hip.h); keeps the arrays alive."""
    _fields_ = [
        ("nc", ctypes.c_int32), ("nb", ctypes.c_int32),
        ("ec", ctypes.POINTER(ctypes.c_double)), ("eb", ctypes.POINTER(ctypes.c_double)),
        ("vbc", ctypes.POINTER(ctypes.c_double)), ("u", ctypes.POINTER(ctypes.c_double)),
        ("jj", ctypes.POINTER(ctypes.c_double)), ("jhund", ctypes.c_double),
        ("cutoff", ctypes.c_double),
    ]

    @classmethod
    def from_params(cls, p: AIMParams) -> "CModel":
        keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (p.ec, p.eb, p.vbc, p.u, p.jj)]
        ptrs = [a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for a in keep]
        m = cls(p.nc, p.nb, *ptrs, p.jhund, p.cutoff)
        m._keep = keep
        return m
