"""The reference's ground-state checkpoint GS.raw (OLDGSFILE), read and written.

write_raw_eigensectorlist / read_raw_eigensectorlist (eigen_sector_class.f90:514-557)
write the eigen-sector list as Fortran list-directed records, item by item:

    T                                   another sector follows
    updo                                write_raw_sector (quantum_sector_class.f90:174-184)
    norbs_up / npart_up / norbs_dn / npart_dn       (write_raw_fermion_sector2, one per line)
      -- or --
    sz / norbs (= 2 Ns) / npart                     (write_raw_fermion_sector)
    T                                   another eigenstate follows (write_raw_eigenlist,
    val / converged / rank              eigen_class.f90:479-490)
    n / v(1) ... v(n)                   write_raw_rcvector (rcvector_class.f90:217-229)
    ...
    F                                   end of this sector's eigenstates
    ...
    F                                   end of the list

so that a solver run on the GPU can hand its states to the reference's DMFT loop and
back (solver.f90:107, solver_init.h:168).  The reader takes any list-directed layout
(values split over lines, `r*value` repeats, D exponents, .TRUE./.FALSE.).
"""
from __future__ import annotations

import numpy as np

from .solver import EigenState


def _fmt(x: float) -> str:
    return repr(float(x))


def write_gs_raw(path: str, states: list[EigenState], ns: int, values_per_line: int = 4) -> None:
    """Write `states` grouped by sector (first appearance order).  A state's sector is
    (N_up, N_dn), or ("sz", npart) for an S_z sector."""
    order: list = []
    groups: dict = {}
    for st in states:
        key = tuple(st.sector)
        if key not in groups:
            groups[key] = []
            order.append(key)
        groups[key].append(st)
    with open(path, "w") as f:
        for key in order:
            f.write(" T\n")
            if key[0] == "sz":
                f.write(" sz\n")
                f.write(f" {2 * ns}\n {int(key[1])}\n")
            else:
                f.write(" updo\n")
                f.write(f" {ns}\n {int(key[0])}\n {ns}\n {int(key[1])}\n")
            for rank, st in enumerate(groups[key], start=1):
                v = np.asarray(st.vector, dtype=np.float64).ravel()
                f.write(" T\n")
                f.write(f" {_fmt(st.energy)}\n")
                f.write(" T\n")
                f.write(f" {rank}\n")
                f.write(f" {v.size}\n")
                for i in range(0, v.size, values_per_line):
                    f.write(" " + " ".join(_fmt(x) for x in v[i:i + values_per_line]) + "\n")
            f.write(" F\n")
        f.write(" F\n")


def _tokens(path: str):
    with open(path) as f:
        for line in f:
            for tok in line.replace(",", " ").split():
                if "*" in tok and not tok.startswith("'"):
                    r, val = tok.split("*", 1)
                    for _ in range(int(r)):
                        yield val
                else:
                    yield tok.strip("'\"")


def _logical(tok: str) -> bool:
    t = tok.strip(".").upper()
    if t.startswith("T"):
        return True
    if t.startswith("F"):
        return False
    raise ValueError(f"GS.raw: expected a logical, found {tok!r}")


def _real(tok: str) -> float:
    return float(tok.replace("D", "E").replace("d", "e"))


def read_gs_raw(path: str) -> list[EigenState]:
    """Read a GS.raw eigen-sector list.  Returns EigenStates with sector (N_up, N_dn) for
    'updo' sectors and ("sz", npart) for S_z sectors; vectors as float64 arrays."""
    it = _tokens(path)
    out: list[EigenState] = []
    while _logical(next(it)):
        kind = next(it).lower()
        if "updo" in kind:
            int(next(it)); nup = int(next(it)); int(next(it)); ndn = int(next(it))
            key = (nup, ndn)
        elif "sz" in kind:
            int(next(it)); npart = int(next(it))
            key = ("sz", npart)
        else:
            raise ValueError(f"GS.raw: unknown sector type {kind!r}")
        while _logical(next(it)):
            val = _real(next(it))
            _logical(next(it))   # converged
            int(next(it))        # rank
            n = int(next(it))
            vec = np.array([_real(next(it)) for _ in range(n)])
            out.append(EigenState(key, val, vec))
    return out
