"""Entry points of module w90_library beside the wannierisation step, called
from a Fortran program (tests/f_demo/library_demo.F90) linked against the
host library: w90_input_reader, w90_set_constant_bohr_to_ang,
w90_get_fortran_file and w90_get_proj (src/library_interface.F90:154, 655,
902).  Host code only (no GPU call).

Expected projections follow the reference's expansion rules
(src/readwrite.F90:3390-4076): per line, every site (all atoms of a label,
or one f=/c= position), then l from -5 to 3, then mr; names map to (l, mr)
as s = (0,1), p = (1,1..3) with pz, px, py = 1, 2, 3, d = (2,1..5) with
dxy = 5, f-names 1..7 in the order fz3, fxz2, fyz2, fxyz, fz(x2-y2),
fx(x2-3y2), fy(3x2-y2), sp..sp3d2 = l -1..-5 with |l|+1 functions; sites
are fractional; c= and atoms in a bohr block scale by the Bohr length.
"""
import gzip
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT

LIB = os.path.join(ROOT, "wannier90_amd", "lib")
SRC = os.path.join(ROOT, "tests", "f_demo", "library_demo.F90")
FC = "/opt/rocm/bin/amdflang"
INPUTS = os.path.join(ROOT, "tests", "golden", "inputs")

needs_host = pytest.mark.skipif(not (os.path.exists(os.path.join(LIB, "libw90host.so")) and os.path.exists(FC)),
                                reason="Fortran host library or compiler missing")


@pytest.fixture(scope="module")
def demo(tmp_path_factory):
    d = tmp_path_factory.mktemp("fdemo")
    exe = str(d / "library_demo")
    subprocess.run([FC, "-O1", "-I", os.path.join(LIB, "fobj"), SRC, f"-L{LIB}", "-lw90host", "-lw90gpu",
                    f"-Wl,-rpath,{LIB}", "-o", exe], check=True, cwd=d)
    return exe


def run_demo(exe, cwd, seed, bohr=None):
    args = [exe, seed] + ([repr(bohr)] if bohr else [])
    r = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = open(os.path.join(cwd, seed + ".proj.txt")).read().split("\n")
    ierr, n = (int(x) for x in rows[0].split())
    lat = np.array([[float(x) for x in rows[1 + i].split()] for i in range(3)])
    projs = []
    for r_ in rows[4:4 + n]:
        v = r_.split()
        projs.append(dict(l=int(v[0]), m=int(v[1]), s=int(v[2]), rad=int(v[3]),
                          site=np.array([float(x) for x in v[4:7]]), z=np.array([float(x) for x in v[7:10]]),
                          x=np.array([float(x) for x in v[10:13]]), zona=float(v[13])))
    return ierr, lat, projs


def unpack(d, name, seed, dst):
    src = os.path.join(INPUTS, d, seed + ".win.gz")
    with gzip.open(src, "rb") as fi, open(os.path.join(dst, seed + ".win"), "wb") as fo:
        shutil.copyfileobj(fi, fo)


@needs_host
def test_get_proj_reference_inputs(demo, tmp_path):
    # Pb:sp3 on the atom at the origin: four sp3 hybrids
    unpack("testw90_example02", "lead", "lead", tmp_path)
    ierr, lat, pr = run_demo(demo, tmp_path, "lead")
    assert ierr == 0 and [(p["l"], p["m"]) for p in pr] == [(-3, 1), (-3, 2), (-3, 3), (-3, 4)]
    assert all(np.array_equal(p["site"], [0, 0, 0]) for p in pr)
    assert all(np.array_equal(p["z"], [0, 0, 1]) and np.array_equal(p["x"], [1, 0, 0]) for p in pr)
    assert all(p["rad"] == 1 and p["zona"] == 1.0 for p in pr)
    # the cell is given in Bohr: the default Bohr length of the build
    # (0.52917720859 Ang) and the caller's after w90_set_constant_bohr_to_ang
    _, lat2, _ = run_demo(demo, tmp_path, "lead", bohr=0.5)
    np.testing.assert_allclose(lat2, lat * (0.5 / 0.52917720859), rtol=1e-15, atol=0)
    # diamond: four s functions at the given fractional sites
    unpack("testw90_example05", "diamond", "diamond", tmp_path)
    ierr, _, pr = run_demo(demo, tmp_path, "diamond")
    assert ierr == 0 and [(p["l"], p["m"]) for p in pr] == [(0, 1)] * 4
    sites = [[0, 0, 0], [0, 0, 0.5], [0, 0.5, 0], [0.5, 0, 0]]
    assert all(np.array_equal(p["site"], s) for p, s in zip(pr, sites))


WIN = """num_wann = {nw}
begin unit_cell_cart
bohr
10.0 0.0 0.0
0.0 10.0 0.0
0.0 0.0 10.0
end unit_cell_cart
begin atoms_cart
bohr
Ga 0.0 0.0 0.0
As 2.5 2.5 2.5
Ga 5.0 5.0 0.0
end atoms_cart
{proj}
mp_grid = 1 1 1
begin kpoints
0 0 0
end kpoints
"""

PROJ = """begin projections
bohr
Ga:s;p
As : l=2,mr=1,4 : z=0,1,0 : x=0,0,1 : r=2 : zona=1.5
c=1.0,2.0,3.0:sp3-2;dxy
f=0.5,0.5,0.5:l=-1;fz(x2-y2)
end projections"""


@needs_host
def test_get_proj_expansion_rules(demo, tmp_path):
    (tmp_path / "pt.win").write_text(WIN.format(nw=2, proj=PROJ))
    ierr, lat, pr = run_demo(demo, tmp_path, "pt", bohr=0.5)
    assert ierr == 0
    np.testing.assert_array_equal(lat, 5.0 * np.eye(3))
    got = [(p["l"], p["m"], tuple(p["site"])) for p in pr]
    want = [(0, 1, (0, 0, 0)), (1, 1, (0, 0, 0)), (1, 2, (0, 0, 0)), (1, 3, (0, 0, 0)),
            (0, 1, (0.5, 0.5, 0)), (1, 1, (0.5, 0.5, 0)), (1, 2, (0.5, 0.5, 0)), (1, 3, (0.5, 0.5, 0)),
            (2, 1, (0.25, 0.25, 0.25)), (2, 4, (0.25, 0.25, 0.25)),
            (-3, 2, (0.1, 0.2, 0.3)), (2, 5, (0.1, 0.2, 0.3)),
            (-1, 1, (0.5, 0.5, 0.5)), (-1, 2, (0.5, 0.5, 0.5)), (3, 5, (0.5, 0.5, 0.5))]
    assert [(a, b) for a, b, _ in got] == [(a, b) for a, b, _ in want]
    for (_, _, s), (_, _, w) in zip(got, want):
        np.testing.assert_allclose(s, w, atol=1e-15)
    for p in pr[8:10]:
        assert np.array_equal(p["z"], [0, 1, 0]) and np.array_equal(p["x"], [0, 0, 1])
        assert p["rad"] == 2 and p["zona"] == 1.5
    assert all(p["rad"] == 1 and p["zona"] == 1.0 for p in pr[:8] + pr[10:])


@needs_host
@pytest.mark.parametrize("proj,nw", [("", 2), ("begin projections\nGa:s\nend projections", 4),
                                     ("begin projections\nIn:s\nend projections", 1),
                                     ("begin projections\nGa:q\nend projections", 1)])
def test_get_proj_errors(demo, tmp_path, proj, nw):
    """No projections block; fewer functions than num_wann; an unknown atom
    label; an unknown angular name: an error status, no functions."""
    (tmp_path / "pe.win").write_text(WIN.format(nw=nw, proj=proj))
    ierr, _, pr = run_demo(demo, tmp_path, "pe")
    assert ierr > 0
