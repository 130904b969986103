"""The drop-in boundary: the wannier90.x-style driver (wannier90_gpu.x, the
Fortran host over libw90gpu.so) on the reference's own regression inputs
(test-suite/tests/testw90_basic1, _example02, _example05; copied as data
under tests/golden/inputs) and on seeded synthetic inputs.

-pp needs no GPU (the k-mesh is host code) and runs everywhere; the full
wannierisation runs on the GPU and is compared with the reference's
checkpoint values stored in tests/golden.
"""
import gzip
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT, Case

EXE = os.path.join(ROOT, "wannier90_amd", "lib", "wannier90_gpu.x")
INPUTS = os.path.join(ROOT, "tests", "golden", "inputs")
REF_CASES = {
    "ref_basic1": ("testw90_basic1", "wannier"),
    "ref_example02": ("testw90_example02", "lead"),
    "ref_example05": ("testw90_example05", "diamond"),
}

needs_exe = pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran host not built")


def unpack(case, dst):
    d, seed = REF_CASES[case]
    for f in os.listdir(os.path.join(INPUTS, d)):
        with gzip.open(os.path.join(INPUTS, d, f), "rb") as fi, open(os.path.join(dst, f[:-3]), "wb") as fo:
            shutil.copyfileobj(fi, fo)
    return seed


def run(seed, cwd, *args, timeout=120):
    print(f"[dropin] {cwd}: wannier90_gpu.x {' '.join(args)} {seed}", flush=True)
    return subprocess.run([EXE, *args, seed], cwd=cwd, capture_output=True, text=True, timeout=timeout)


@needs_exe
@pytest.mark.parametrize("case", sorted(REF_CASES))
def test_pp_writes_reference_nnkp(case, tmp_path):
    """-pp (wannier_prog.F90 post-processing setup) writes the neighbour
    list the reference computes for the same .win (cell in Bohr for basic1)."""
    from wannier90_amd import w90io

    seed = unpack(case, tmp_path)
    r = run(seed, tmp_path, "-pp")
    assert r.returncode == 0, r.stdout + r.stderr
    nn, nnlist, nncell = w90io.read_nnkp(os.path.join(tmp_path, seed + ".nnkp"))
    c = Case(case)
    assert nn == c.nn
    assert np.array_equal(nnlist, c.kmesh.nnlist) and np.array_equal(nncell, c.kmesh.nncell)


@needs_exe
def test_pp_rejects_paths_outside_this_build(tmp_path):
    seed = unpack("ref_basic1", tmp_path)
    win = os.path.join(tmp_path, seed + ".win")
    with open(win) as f:
        txt = f.read()
    with open(win, "w") as f:
        f.write("precond = true\n" + txt)
    r = run(seed, tmp_path, "-pp")
    assert r.returncode != 0
    assert "precond" in open(os.path.join(tmp_path, seed + ".werr")).read()


@pytest.mark.gpu
@needs_exe
@pytest.mark.parametrize("case", sorted(REF_CASES))
def test_driver_matches_reference_checkpoint(case, tmp_path):
    """wannier90_gpu.x seedname on the reference's regression inputs writes
    the .chk (U, M, centres, spreads) and the Omega of the reference's run."""
    from wannier90_amd import w90io

    seed = unpack(case, tmp_path)
    r = run(seed, tmp_path)
    assert r.returncode == 0, open(os.path.join(tmp_path, seed + ".wout")).read()[-2000:]
    c = Case(case)
    chk = w90io.read_chk(os.path.join(tmp_path, seed + ".chk"))
    assert chk["checkpoint"] == "postwann" and not chk["have_disentangled"]
    tol_u = max(1e-7, 20 * c.sens_u)
    assert np.abs(chk["u_matrix"] - c.z["u_final"]).max() <= tol_u
    assert np.abs(chk["m_matrix"] - c.z["m_final"]).max() <= tol_u
    assert np.abs(chk["centres"] - c.z["centres"]).max() <= 1e-6
    assert np.abs(chk["spreads"] - c.z["spreads"]).max() <= 1e-8
    cen, spr, om = w90io.read_wout_final(os.path.join(tmp_path, seed + ".wout"))
    # .wout prints Omega with 9 decimals, as the reference
    assert abs(om["omega_total"] - c.z["omega"][3]) <= 2e-9
    assert abs(om["omega_i"] - c.z["omega"][0]) <= 2e-9
    assert not os.path.exists(os.path.join(tmp_path, seed + ".werr"))


@pytest.mark.gpu
@needs_exe
def test_driver_outputs_and_restart(tmp_path):
    """_centres.xyz and _u.mat when asked for; restart = wannierise resumes
    from the U and M of the checkpoint (wannier_prog.F90:204-235): the
    restarted run equals a fresh minimisation started from them."""
    from wannier90_amd import capi, w90io

    seed = unpack("ref_example05", tmp_path)
    win = os.path.join(tmp_path, seed + ".win")
    txt = open(win).read()
    txt = "\n".join(l for l in txt.splitlines() if not l.strip().lower().startswith("num_iter"))
    open(win, "w").write("num_iter = 10\nwrite_xyz = true\nwrite_u_matrices = true\n" + txt + "\n")
    assert run(seed, tmp_path).returncode == 0
    ua, kp = w90io.read_u_mat(os.path.join(tmp_path, seed + "_u.mat"))
    chk_a = w90io.read_chk(os.path.join(tmp_path, seed + ".chk"))
    assert np.abs(ua - chk_a["u_matrix"]).max() < 1e-9  # f15.10 in _u.mat
    lines = open(os.path.join(tmp_path, seed + "_centres.xyz")).read().splitlines()
    assert int(lines[0]) >= chk_a["num_wann"] and lines[2].startswith("X")
    open(win, "a").write("restart = wannierise\n")
    assert run(seed, tmp_path).returncode == 0
    chk_r = w90io.read_chk(os.path.join(tmp_path, seed + ".chk"))
    c = Case("ref_example05")
    ctx = capi.Context(0)
    m = np.array(chk_a["m_matrix"], order="F")
    u = np.array(chk_a["u_matrix"], order="F")
    cen, spr, res = ctx.wann_main(c.kmesh, m, u, c.gpu_control(num_iter=10))
    ctx.close()
    assert np.abs(chk_r["u_matrix"] - u).max() < 1e-12
    assert np.abs(chk_r["centres"] - cen).max() < 1e-12


HOST_LIB = os.path.join(ROOT, "wannier90_amd", "lib", "libw90host.so")


def test_host_library_exports_the_reference_c_interface():
    """libw90host.so exports every routine include/w90_library_c.h declares,
    under the names of src/c_interface.F90."""
    import ctypes
    import re

    if not os.path.exists(HOST_LIB):
        pytest.skip("Fortran host not built")
    txt = open(os.path.join(ROOT, "include", "w90_library_c.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = set(re.findall(r"\b(c\w+|w90_\w+)\s*\(", txt)) - {"cplusplus"}
    assert {"w90_create", "w90_delete", "cwannierise", "cset_m_local", "cget_centres"} <= names
    lib = ctypes.CDLL(HOST_LIB)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.gpu
@needs_exe
def test_c_caller_of_reference_interface(tmp_path):
    """A C++ program using the reference's C interface (options, k-mesh,
    caller-owned arrays, project, wannierise, getters) reproduces the
    reference run of the same problem (cfg1, 60 iterations, no window)."""
    import struct

    from wannier90_amd import synth

    exe = os.path.join(tmp_path, "libc_demo")
    b = subprocess.run(["g++", "-O1", "-std=c++17", f"-I{ROOT}/include", "-I/opt/rocm/llvm/include/flang",
                        os.path.join(ROOT, "tests", "c_demo", "libc_demo.cc"),
                        f"-L{ROOT}/wannier90_amd/lib", "-lw90host", "-lw90gpu",
                        f"-Wl,-rpath,{ROOT}/wannier90_amd/lib", "-o", exe], capture_output=True, text=True)
    assert b.returncode == 0, b.stderr
    c = Case("cfg1_nowindow")
    prob = synth.config(1)
    nw, nk, nn = prob.num_wann, prob.num_kpts, prob.kmesh.nntot
    ctl = c.ctl
    with open(os.path.join(tmp_path, "in.bin"), "wb") as f:
        f.write(struct.pack("<8i", nw, nk, *prob.mp, ctl["num_iter"], ctl["num_cg_steps"], ctl["conv_window"]))
        f.write(struct.pack("<2d", ctl["conv_tol"], ctl["trial_step"]))
        f.write(np.ascontiguousarray(prob.real_lattice, dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(prob.kpts, dtype="<f8").tobytes())
        f.write(struct.pack("<i", nn))
        f.write(np.asarray(prob.m_orig, dtype="<c16").tobytes(order="F"))
        f.write(np.asarray(np.transpose(prob.amat, (1, 2, 0)), dtype="<c16").tobytes(order="F"))
    r = subprocess.run([exe, "in.bin", "out.bin"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = open(os.path.join(tmp_path, "out.bin"), "rb").read()
    ierr, nn_out = struct.unpack_from("<2i", raw, 0)
    pos = 8
    nnlist = np.frombuffer(raw, "<i4", nk * nn_out, pos).reshape((nk, nn_out), order="F")
    pos += 4 * nk * nn_out
    u = np.frombuffer(raw, "<c16", nw * nw * nk, pos).reshape((nw, nw, nk), order="F")
    pos += 16 * nw * nw * nk
    cen = np.frombuffer(raw, "<f8", 3 * nw, pos).reshape((3, nw), order="F")
    pos += 24 * nw
    spr = np.frombuffer(raw, "<f8", nw, pos)
    pos += 8 * nw
    om = np.frombuffer(raw, "<f8", 4, pos)
    assert ierr == 0 and nn_out == nn and np.array_equal(nnlist, c.kmesh.nnlist)
    assert np.abs(u - c.z["u_final"]).max() <= 1e-7
    assert np.abs(cen - c.z["centres"]).max() <= 1e-6
    assert np.abs(spr - c.z["spreads"]).max() <= 1e-8
    assert abs(om[3] - c.z["omega"][3]) <= 2e-9


@pytest.mark.gpu
@needs_exe
def test_driver_reports_missing_inputs(tmp_path):
    seed = unpack("ref_basic1", tmp_path)
    os.remove(os.path.join(tmp_path, seed + ".mmn"))
    r = run(seed, tmp_path)
    assert r.returncode != 0
    assert "mmn" in open(os.path.join(tmp_path, seed + ".werr")).read()
