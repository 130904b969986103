"""Runs the reference wannierisation (oracle/_ref/ref_driver.x, the
reference's w90_library built from its sources by oracle/Makefile.ref) on an
initial state, serially or over MPI ranks (oracle/_ref/mpi/ref_driver.x).

TEST INFRASTRUCTURE AND CPU BASELINE ONLY: used by tests/golden/make_golden.py
to write fixtures and by bench.py's cpu_baseline leg to time the reference.
"""
import os
import re
import shutil
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SERIAL = os.path.join(HERE, "_ref", "ref_driver.x")
PARALLEL = os.path.join(HERE, "_ref", "mpi", "ref_driver.x")
MPIEXEC = "/opt/conda/bin/mpiexec"


def available(ranks=1):
    exe = SERIAL if ranks <= 1 else PARALLEL
    return os.path.exists(exe) and (ranks <= 1 or os.path.exists(MPIEXEC))


def write_input(path, real_lattice, mp, kpts, kmesh, m0, u0, ctl):
    nw, nk, nn = u0.shape[0], u0.shape[2], kmesh.nntot
    head = np.array([nw, nk, nn, *mp, ctl["num_iter"], ctl.get("num_cg_steps", 5), ctl.get("conv_window", -1),
                     ctl.get("num_print_cycles", 1)], dtype="<i4")
    with open(path, "wb") as f:
        f.write(head.tobytes())
        f.write(np.array([ctl.get("conv_tol", 1e-10), ctl.get("trial_step", 2.0)], dtype="<f8").tobytes())
        # column i = lattice vector i
        f.write(np.asarray(real_lattice, dtype="<f8").T.tobytes(order="F"))
        f.write(np.asarray(kpts, dtype="<f8").T.tobytes(order="F"))
        f.write(np.asarray(kmesh.nnlist, dtype="<i4").tobytes(order="F"))
        f.write(np.asarray(m0, dtype="<c16").tobytes(order="F"))
        f.write(np.asarray(u0, dtype="<c16").tobytes(order="F"))


def read_conv(path):
    """(iteration, Omega, wall seconds) of the CONV lines of a .wout."""
    it, om, wall = [], [], []
    pat = re.compile(r"^\s*(\d+)\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)\s+<-- CONV")
    for line in open(path):
        m = pat.match(line)
        if m:
            it.append(int(m.group(1)))
            om.append(float(m.group(4)))
            wall.append(float(m.group(5)))
    return np.array(it), np.array(om), np.array(wall)


def run(real_lattice, mp, kpts, kmesh, m0, u0, ctl, ranks=1, timeout=3600, env=None):
    """The reference's wann_main on (M0, U0).  Returns seconds (wall time of
    w90_wannierise), omega (invariant, tilde, total), centres, spreads, U,
    and the CONV trajectory (iterations, Omega, wall column)."""
    nw, nk = u0.shape[0], u0.shape[2]
    tmp = tempfile.mkdtemp(prefix="w90ref_")
    try:
        fin, fout = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
        write_input(fin, real_lattice, mp, kpts, kmesh, m0, u0, ctl)
        e = dict(os.environ, OMP_NUM_THREADS="1", MKL_NUM_THREADS="1", MKL_THREADING_LAYER="SEQUENTIAL")
        e.update(env or {})
        cmd = [SERIAL, fin, fout, "w"] if ranks <= 1 else [MPIEXEC, "-n", str(ranks), PARALLEL, fin, fout, "w"]
        r = subprocess.run(cmd, cwd=tmp, env=e, capture_output=True, text=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError(f"ref_driver failed ({r.returncode}): {r.stdout[-2000:]} {r.stderr[-2000:]}")
        with open(fout, "rb") as f:
            sec, oi, ot, otot = np.frombuffer(f.read(32), dtype="<f8")
            cen = np.frombuffer(f.read(24 * nw), dtype="<f8").reshape((3, nw), order="F")
            spr = np.frombuffer(f.read(8 * nw), dtype="<f8").copy()
            u = np.frombuffer(f.read(16 * nw * nw * nk), dtype="<c16").reshape((nw, nw, nk), order="F")
        it, om, wall = read_conv(os.path.join(tmp, "w.wout"))
        return dict(seconds=float(sec), omega=(oi, ot, otot), centres=np.array(cen, order="F"), spreads=spr,
                    u=np.array(u, order="F"), conv_iter=it, conv_omega=om, conv_wall=wall,
                    iterations=int(it[-1]) if len(it) else 0)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
