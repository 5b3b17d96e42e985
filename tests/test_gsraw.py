"""GS.raw checkpoint: round trip, and list-directed interchange with a Fortran program
(the reference's own records are Fortran list-directed I/O; compiled with amdflang
when it is present)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from ed_solver_amd.gsraw import read_gs_raw, write_gs_raw
from ed_solver_amd.solver import EigenState

FLANG = os.environ.get("ED_FLANG", "/opt/rocm/llvm/bin/amdflang")
HERE = os.path.dirname(os.path.abspath(__file__))


def sample_states():
    r = np.random.default_rng(4)
    return [EigenState((3, 3), -5.978224258990123, r.normal(size=400)),
            EigenState((3, 3), -5.1, r.normal(size=400)),
            EigenState((4, 3), -4.25e-3, r.normal(size=300) * 1e-200),
            EigenState(("sz", 6), 1.0 / 3.0, r.normal(size=924))]


def test_round_trip_bit_exact(tmp_path):
    st = sample_states()
    p = tmp_path / "GS.raw"
    write_gs_raw(str(p), st, ns=6)
    back = read_gs_raw(str(p))
    assert [b.sector for b in back] == [s.sector for s in st]
    for a, b in zip(st, back):
        assert a.energy == b.energy
        assert np.array_equal(a.vector, b.vector)


def test_reader_takes_list_directed_variants(tmp_path):
    p = tmp_path / "v.raw"
    p.write_text(" .TRUE.\n 'updo'\n 4 2 4 1\n T\n -1.25D+00\n .true.\n 1\n 5\n 2*0.5d0, 1.0\n -2.5E0 3\n F\n F\n")
    st = read_gs_raw(str(p))
    assert len(st) == 1 and st[0].sector == (2, 1) and st[0].energy == -1.25
    assert np.array_equal(st[0].vector, [0.5, 0.5, 1.0, -2.5, 3.0])


@pytest.mark.skipif(not shutil.which(FLANG) and not os.path.exists(FLANG), reason="no Fortran compiler")
def test_fortran_list_directed_interchange(tmp_path):
    exe = tmp_path / "gsraw_io"
    r = subprocess.run([FLANG, os.path.join(HERE, "fortran", "gsraw_io.f90"), "-o", str(exe)], capture_output=True,
                       text=True)
    if r.returncode != 0:
        pytest.skip("Fortran compile failed: " + r.stderr[-300:])
    # Fortran writes, Python reads
    fw = tmp_path / "f.raw"
    subprocess.run([str(exe), "write", str(fw)], check=True)
    st = read_gs_raw(str(fw))
    assert [s.sector for s in st] == [(2, 2), (3, 2), (3, 2)]
    for s_, (sec, k) in zip(st, [(1, 1), (2, 1), (2, 2)]):
        v = np.array([i / (sec + k) - 0.125 * i * i for i in range(1, 8)])
        assert s_.energy == -1.5 * sec - 0.25 * k
        assert np.max(np.abs(s_.vector - v)) <= 1e-15 * np.max(np.abs(v))
    # Python writes, Fortran reads
    st = [s for s in sample_states() if s.sector[0] != "sz"]
    pw = tmp_path / "p.raw"
    write_gs_raw(str(pw), st, ns=6)
    out = subprocess.run([str(exe), "read", str(pw)], check=True, capture_output=True, text=True).stdout.split("\n")
    rows = [line.split() for line in out if line.strip()]
    assert len(rows) == len(st)
    for row, s_ in zip(rows, st):
        assert (int(row[0]), int(row[1])) == s_.sector
        assert float(row[2]) == s_.energy
        assert int(row[5]) == s_.vector.size
        assert abs(float(row[3]) - s_.vector.sum()) <= 1e-12 * np.abs(s_.vector).sum()
        i = np.arange(1, s_.vector.size + 1)
        assert abs(float(row[4]) - (i * s_.vector).sum()) <= 1e-12 * (i * np.abs(s_.vector)).sum()
