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


BIG = 1.7976931348623157e308
DIS_DEFAULTS = dict(win_min=-BIG, win_max=BIG, froz_min=-BIG, froz_max=BIG, frozen_states=0, num_iter=200,
                    mix_ratio=0.5, conv_tol=1e-10, conv_window=3)


def write_input(path, real_lattice, mp, kpts, kmesh, m0, u0, ctl, dis=None):
    """dis: None (m0 (nw,nw,nn,nk), u0 (nw,nw,nk)), or a dict with
    num_bands, eig (nb,nk), a (nb,nw,nk) and the w90gpu_dis_control fields
    (m0 then (nb,nb,nn,nk), u0 unused)."""
    nk, nn = kpts.shape[0], kmesh.nntot
    if dis is None:
        nw = u0.shape[0]
        nb = nw
        d = dict(DIS_DEFAULTS)
    else:
        nb = int(dis["num_bands"])
        nw = dis["a"].shape[1]
        d = dict(DIS_DEFAULTS)
        d.update({k: v for k, v in dis.items() if k in DIS_DEFAULTS})
    head = np.array([nw, nk, nn, *mp, ctl["num_iter"], ctl.get("num_cg_steps", 5), ctl.get("conv_window", -1),
                     ctl.get("num_print_cycles", 1), nb, d["num_iter"], d["conv_window"], d["frozen_states"]],
                    dtype="<i4")
    with open(path, "wb") as f:
        f.write(head.tobytes())
        f.write(np.array([ctl.get("conv_tol", 1e-10), ctl.get("trial_step", 2.0), d["win_min"], d["win_max"],
                          d["froz_min"], d["froz_max"], d["mix_ratio"], d["conv_tol"]], dtype="<f8").tobytes())
        # column i = lattice vector i
        f.write(np.asarray(real_lattice, dtype="<f8").T.tobytes(order="F"))
        f.write(np.asarray(kpts, dtype="<f8").T.tobytes(order="F"))
        f.write(np.asarray(kmesh.nnlist, dtype="<i4").tobytes(order="F"))
        f.write(np.asarray(m0, dtype="<c16").tobytes(order="F"))
        if dis is None:
            f.write(np.asarray(u0, dtype="<c16").tobytes(order="F"))
        else:
            f.write(np.asarray(dis["a"], dtype="<c16").tobytes(order="F"))
            f.write(np.asarray(dis["eig"], dtype="<f8").tobytes(order="F"))


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


def read_dis(path):
    """(iteration, Omega_I(i-1), Omega_I(i), delta, wall) of the DIS lines."""
    rows = []
    pat = re.compile(r"^\s*(\d+)\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)\s+<-- DIS")
    for line in open(path):
        m = pat.match(line)
        if m:
            rows.append([float(m.group(i)) for i in range(1, 6)])
    return np.array(rows).reshape(-1, 5)


def run(real_lattice, mp, kpts, kmesh, m0, u0, ctl, ranks=1, timeout=3600, env=None, dis=None):
    """The reference's wann_main on (M0, U0), or with dis (see write_input)
    its w90_disentangle first.  Returns seconds (wall time of
    w90_wannierise), omega (invariant, tilde, total), centres, spreads, U,
    the CONV trajectory (iterations, Omega, wall column), and with dis the
    seconds of w90_disentangle, U_opt and the DIS lines."""
    nk = kpts.shape[0]
    nw = u0.shape[0] if dis is None else dis["a"].shape[1]
    tmp = tempfile.mkdtemp(prefix="w90ref_")
    try:
        fin, fout = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
        write_input(fin, real_lattice, mp, kpts, kmesh, m0, u0, ctl, dis)
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
            out = {}
            if dis is not None:
                nb = int(dis["num_bands"])
                out["dis_seconds"] = float(np.frombuffer(f.read(8), dtype="<f8")[0])
                out["u_opt"] = np.array(np.frombuffer(f.read(16 * nb * nw * nk), dtype="<c16").reshape(
                    (nb, nw, nk), order="F"), order="F")
        it, om, wall = read_conv(os.path.join(tmp, "w.wout"))
        if dis is not None:
            out["dis_history"] = read_dis(os.path.join(tmp, "w.wout"))
        out.update(seconds=float(sec), omega=(oi, ot, otot), centres=np.array(cen, order="F"), spreads=spr,
                   u=np.array(u, order="F"), conv_iter=it, conv_omega=om, conv_wall=wall,
                   iterations=int(it[-1]) if len(it) else 0)
        return out
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
