"""The drop-in boundary: the wannier90.x-style driver (wannier90_gpu.x, the
Fortran host over libw90gpu.so) on the reference's own regression inputs
(test-suite/tests/testw90_basic1, _example02, _example05,
_write_u_matrices; copied as data under tests/golden/inputs) and on seeded
synthetic inputs.

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

from conftest import ROOT, Case, tol

EXE = os.path.join(ROOT, "wannier90_amd", "lib", "wannier90_gpu.x")
INPUTS = os.path.join(ROOT, "tests", "golden", "inputs")
REF_CASES = {
    "ref_basic1": ("testw90_basic1", "wannier"),
    "ref_example02": ("testw90_example02", "lead"),
    "ref_example05": ("testw90_example05", "diamond"),
    "ref_write_u_matrices": ("testw90_write_u_matrices", "gaas"),
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


@pytest.mark.gpu
@needs_exe
def test_driver_writes_u_matrices_like_the_reference(tmp_path):
    """The reference's write_u_matrices case (GaAs, 4 WFs on a 2x2x2 mesh,
    write_u_matrices in its .win): _u.mat holds the k-points and the final
    U of the run, the checkpoint's to the file's 10 decimals, and the
    reference's final U to the checkpoint test's bound."""
    from wannier90_amd import w90io

    seed = unpack("ref_write_u_matrices", tmp_path)
    r = run(seed, tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    u, kp = w90io.read_u_mat(os.path.join(tmp_path, seed + "_u.mat"))
    chk = w90io.read_chk(os.path.join(tmp_path, seed + ".chk"))
    c = Case("ref_write_u_matrices")
    assert np.abs(kp - c.z["kpt_latt"]).max() <= 1e-10
    assert np.abs(u - chk["u_matrix"]).max() <= 1e-9
    assert np.abs(u - c.z["u_final"]).max() <= max(1e-7, 20 * c.sens_u)


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


def run_c_demo(tmp_path, nw, nb, mp, ctl, real_lattice, kpts, nn, m, a, eig=None, dctl=None):
    """Build tests/c_demo/libc_demo.cc against the host library, run it on
    one problem (m (nb,nb,nn,nk), a (nb,nw,nk), Fortran order) and return
    ierr, nnlist, U, U_opt, centres, spreads, omega."""
    import struct

    exe = os.path.join(tmp_path, "libc_demo")
    b = subprocess.run(["g++", "-O1", "-std=c++17", f"-I{ROOT}/include", "-I/opt/rocm/llvm/include/flang",
                        os.path.join(ROOT, "tests", "c_demo", "libc_demo.cc"),
                        f"-L{ROOT}/wannier90_amd/lib", "-lw90host", "-lw90gpu",
                        f"-Wl,-rpath,{ROOT}/wannier90_amd/lib", "-o", exe], capture_output=True, text=True)
    assert b.returncode == 0, b.stderr
    nk = kpts.shape[0]
    d = dctl or {"num_iter": 0, "conv_window": 0, "win_min": -2e300, "win_max": 2e300, "froz_min": -2e300,
                 "froz_max": 2e300, "mix_ratio": 0.5, "conv_tol": 1e-10, "frozen_states": 0}
    froz = (d["froz_min"], d["froz_max"]) if d["frozen_states"] else (-2e300, 2e300)
    with open(os.path.join(tmp_path, "in.bin"), "wb") as f:
        f.write(struct.pack("<11i", nw, nb, nk, *mp, ctl["num_iter"], ctl["num_cg_steps"], ctl["conv_window"],
                            d["num_iter"], d["conv_window"]))
        f.write(struct.pack("<8d", ctl["conv_tol"], ctl["trial_step"], d["win_min"], d["win_max"], *froz,
                            d["mix_ratio"], d["conv_tol"]))
        f.write(np.ascontiguousarray(real_lattice, dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(kpts, dtype="<f8").tobytes())
        f.write(struct.pack("<i", nn))
        f.write(np.asarray(m, dtype="<c16").tobytes(order="F"))
        f.write(np.asarray(a, dtype="<c16").tobytes(order="F"))
        if nb > nw:
            f.write(np.asarray(eig, dtype="<f8").tobytes(order="F"))
    r = subprocess.run([exe, "in.bin", "out.bin"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = open(os.path.join(tmp_path, "out.bin"), "rb").read()
    ierr, nn_out = struct.unpack_from("<2i", raw, 0)
    pos = 8
    nnlist = np.frombuffer(raw, "<i4", nk * nn_out, pos).reshape((nk, nn_out), order="F")
    pos += 4 * nk * nn_out
    u = np.frombuffer(raw, "<c16", nw * nw * nk, pos).reshape((nw, nw, nk), order="F")
    pos += 16 * nw * nw * nk
    uopt = np.frombuffer(raw, "<c16", nb * nw * nk, pos).reshape((nb, nw, nk), order="F")
    pos += 16 * nb * nw * nk
    cen = np.frombuffer(raw, "<f8", 3 * nw, pos).reshape((3, nw), order="F")
    pos += 24 * nw
    spr = np.frombuffer(raw, "<f8", nw, pos)
    pos += 8 * nw
    om = np.frombuffer(raw, "<f8", 4, pos)
    assert nn_out == nn
    return ierr, nnlist, u, uopt, cen, spr, om


@pytest.mark.gpu
@needs_exe
def test_c_caller_of_reference_interface(tmp_path):
    """A C++ program using the reference's C interface (options, k-mesh,
    caller-owned arrays, project, wannierise, getters) reproduces the
    reference run of the same problem (cfg1, 60 iterations, no window)."""
    from wannier90_amd import synth

    c = Case("cfg1_nowindow")
    prob = synth.config(1)
    nw, nn = prob.num_wann, prob.kmesh.nntot
    ierr, nnlist, u, _, cen, spr, om = run_c_demo(tmp_path, nw, nw, prob.mp, c.ctl, prob.real_lattice, prob.kpts,
                                                  nn, prob.m_orig, np.transpose(prob.amat, (1, 2, 0)))
    assert ierr == 0 and np.array_equal(nnlist, c.kmesh.nnlist)
    assert np.abs(u - c.z["u_final"]).max() <= 1e-7
    assert np.abs(cen - c.z["centres"]).max() <= 1e-6
    assert np.abs(spr - c.z["spreads"]).max() <= 1e-8
    assert abs(om[3] - c.z["omega"][3]) <= 2e-9


@pytest.mark.gpu
@needs_exe
def test_c_caller_disentangles_like_the_reference(tmp_path):
    """The same C++ caller with num_bands > num_wann (the reference's silicon
    example, 12 bands to 8 WFs, outer and frozen windows): band energies
    through cset_eigval, the projections in U_opt, cdisentangle, then
    cwannierise, against the reference's run: the optimal subspace (the
    projector U_opt U_opt^+), U_opt U, the spreads and Omega (tolerances of
    the driver's test of the same run)."""
    import json

    z = np.load(os.path.join(ROOT, "tests", "golden", "dis", "dis_silicon.npz"))
    nb, nw = int(z["num_bands"]), int(z["num_wann"])
    ctl = json.loads(str(z["control"]))
    dctl = json.loads(str(z["dis_control"]))
    ierr, nnlist, u, uopt, cen, spr, om = run_c_demo(tmp_path, nw, nb, tuple(int(x) for x in z["mp_grid"]), ctl,
                                                     z["real_lattice"], z["kpt_latt"], int(z["nnlist"].shape[1]),
                                                     z["m"], z["a"], z["eig"], dctl)
    assert ierr == 0 and np.array_equal(nnlist, z["nnlist"])
    p = np.einsum("iwk,jwk->ijk", uopt, uopt.conj())
    pr = np.einsum("iwk,jwk->ijk", z["u_opt"], z["u_opt"].conj())
    assert np.abs(p - pr).max() <= 1e-8
    w = np.einsum("bwk,wvk->bvk", uopt, u)
    wr = np.einsum("bwk,wvk->bvk", z["u_opt"], z["u_final"])
    assert np.abs(w - wr).max() <= 1e-5
    assert np.abs(np.sort(spr) - np.sort(z["spreads"])).max() <= 5e-6
    assert abs(om[3] - float(z["omega"][3])) <= 2e-8


@pytest.mark.gpu
@needs_exe
def test_c_caller_disentangles_above_64_bands(tmp_path):
    """The same C++ caller on a synthetic problem of 72 bands to 12 WFs (fcc,
    2x2x2 mesh): the host library's size checks and the disentanglement
    kernels whose matrices live in global memory, behind cdisentangle, then
    cwannierise.  The optimal subspace against the oracle's dis_main (1e-8);
    Omega against the same two steps through the GPU's C ABI (1e-10
    relative: the host adds no arithmetic) and against the oracle's
    wann_main from that disentanglement's M and U (1e-8 Ang^2, or 20x the
    trajectory's own sensitivity to rounding where that is larger: here a
    1e-15 perturbation of M moves the final Omega by 2e-6 Ang^2, and the
    GPU ends 1.5e-5 from the oracle)."""
    po = pytest.importorskip("oracle.pyoracle")
    from wannier90_amd import capi, synth

    nb, nw = 72, 12
    p, dc = synth.make_dis_problem(7 + nb, kind="fcc", num_bands=nb, num_wann=nw, mp=(2, 2, 2))
    dc = dict(dc, num_iter=30)
    ctl = dict(num_iter=20, num_cg_steps=5, conv_window=-1, conv_tol=1e-10, trial_step=2.0)
    dctl = {"num_iter": dc["num_iter"], "conv_window": dc["conv_window"], "win_min": -2e300, "win_max": 2e300,
            "froz_min": -2e300, "froz_max": 2e300, "mix_ratio": dc["mix_ratio"], "conv_tol": dc["conv_tol"],
            "frozen_states": 0}
    print(f"[c demo] dis nb={nb} nw={nw} nk={p.num_kpts}", flush=True)
    ierr, _, _, uopt, _, _, om = run_c_demo(tmp_path, nw, nb, p.mp, ctl, p.real_lattice, p.kpts, p.kmesh.nntot,
                                            p.m_orig, p.amat_dis, p.eig, dctl)
    assert ierr == 0
    o = po.dis_main(nb, nw, p.kmesh, dc, p.eig, p.amat_dis, p.m_orig)

    def proj(x):
        return np.einsum("iwk,jwk->ijk", x, x.conj())

    assert np.abs(proj(uopt) - proj(o["u_opt"])).max() <= 1e-8
    ctx = capi.Context(0)
    g = ctx.disentangle(p.kmesh, nb, dc, p.eig, p.amat_dis, p.m_orig)
    sess = ctx.session(p.kmesh, nw, p.num_kpts, capi.control(**ctl))
    sess.load(np.asfortranarray(g["m"]), np.asfortranarray(g["u"]))
    res = sess.solve()
    sess.close()
    ctx.close()
    assert abs(om[3] - res.omega_total) <= 1e-10 * max(1.0, abs(res.omega_total))
    op = po.Problem(nw, p.num_kpts, p.kmesh)
    mg, ug = np.asfortranarray(g["m"]), np.asfortranarray(g["u"])
    ow = po.wann_main(op, ctl, mg, ug)
    sens = 0.0
    for seed in (1, 2):
        rng = np.random.default_rng(seed)
        o2 = po.wann_main(op, ctl, mg * (1.0 + 1e-15 * rng.standard_normal(mg.shape)), ug)
        sens = max(sens, abs(o2["omega"][3] - ow["omega"][3]))
    print(f"[c demo] Omega {om[3]:.10f}, oracle {ow['omega'][3]:.10f}, sensitivity {sens:.1e}", flush=True)
    assert abs(om[3] - ow["omega"][3]) <= tol(1e-8 * max(1.0, abs(ow["omega"][3])), sens)


@pytest.mark.gpu
@needs_exe
def test_driver_reports_missing_inputs(tmp_path):
    seed = unpack("ref_basic1", tmp_path)
    os.remove(os.path.join(tmp_path, seed + ".mmn"))
    r = run(seed, tmp_path)
    assert r.returncode != 0
    assert "mmn" in open(os.path.join(tmp_path, seed + ".werr")).read()


def write_dis_inputs(name, dst, params=None):
    """The .win/.mmn/.amn/.eig of a disentanglement fixture (the reference's
    test-suite inputs, stored as data in tests/golden/dis)."""
    import json

    from wannier90_amd import capi, w90io

    z = np.load(os.path.join(ROOT, "tests", "golden", "dis", name + ".npz"))
    km = capi.KMesh(int(z["nnlist"].shape[1]), np.asfortranarray(z["nnlist"].astype(np.int32)),
                    np.asfortranarray(z["nncell"]), np.asfortranarray(z["bk"]), np.ascontiguousarray(z["wb"]),
                    float(z["wbtot"]))
    dctl = json.loads(str(z["dis_control"]))
    ctl = json.loads(str(z["control"]))
    p = {"dis_num_iter": dctl["num_iter"], "dis_mix_ratio": dctl["mix_ratio"], "dis_conv_tol": dctl["conv_tol"],
         "dis_conv_window": dctl["conv_window"], "num_iter": ctl["num_iter"], "num_cg_steps": ctl["num_cg_steps"],
         "conv_window": ctl["conv_window"], "conv_tol": ctl["conv_tol"], "trial_step": ctl["trial_step"]}
    for key in ("win_min", "win_max", "froz_max", "froz_min"):
        if abs(dctl[key]) < 1e300 and (key != "froz_min" or dctl["frozen_states"]):
            p["dis_" + key] = repr(float(dctl[key]))
    p.update(params or {})
    seed = "w"
    w90io.write_win(os.path.join(dst, seed + ".win"), z["real_lattice"], tuple(z["mp_grid"]), z["kpt_latt"],
                    int(z["num_wann"]), int(z["num_bands"]), p)
    w90io.write_mmn(os.path.join(dst, seed + ".mmn"), z["m"], km)
    w90io.write_amn(os.path.join(dst, seed + ".amn"), np.transpose(z["a"], (2, 0, 1)))
    w90io.write_eig(os.path.join(dst, seed + ".eig"), z["eig"].T)
    return seed, z


@pytest.mark.gpu
@needs_exe
def test_driver_disentangles_like_the_reference(tmp_path):
    """wannier90_gpu.x on the reference's silicon example (12 bands, 8 WFs,
    outer and frozen windows): the DIS lines, the disentangled checkpoint
    (Omega_I, ndimwin, the subspaces of U_opt, U_opt U) and the final Omega
    of the reference's run; restart = wannierise reads the disentangled
    checkpoint back."""
    from wannier90_amd import w90io

    seed, z = write_dis_inputs("dis_silicon", tmp_path)
    r = run(seed, tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr + open(os.path.join(tmp_path, seed + ".werr")).read()
    hist, om_i = w90io.read_wout_dis(os.path.join(tmp_path, seed + ".wout"))
    ref = z["dis_history"]
    assert hist.shape == ref.shape
    assert np.abs(hist[:, :2] - ref[:, :2]).max() <= 2e-8  # both printed with 8 decimals
    assert abs(om_i - float(z["omega_invariant"])) <= 2e-8
    chk = w90io.read_chk(os.path.join(tmp_path, seed + ".chk"))
    assert chk["have_disentangled"] and chk["checkpoint"] == "postwann"
    assert np.array_equal(chk["ndimwin"], z["ndimwin"]) and np.array_equal(chk["lwindow"], z["lwindow"])
    assert abs(chk["omega_invariant"] - float(z["omega_invariant"])) <= 1e-10 * float(z["omega_invariant"])
    p = np.einsum("iwk,jwk->ijk", chk["u_matrix_opt"], chk["u_matrix_opt"].conj())
    pr = np.einsum("iwk,jwk->ijk", z["u_opt"], z["u_opt"].conj())
    assert np.abs(p - pr).max() <= 1e-8
    w = np.einsum("bwk,wvk->bvk", chk["u_matrix_opt"], chk["u_matrix"])
    wr = np.einsum("bwk,wvk->bvk", z["u_opt"], z["u_final"])
    assert np.abs(w - wr).max() <= 1e-5
    cen, spr, om = w90io.read_wout_final(os.path.join(tmp_path, seed + ".wout"))
    assert abs(om["omega_total"] - float(z["omega"][3])) <= 2e-8
    assert np.abs(np.sort(spr) - np.sort(z["spreads"])).max() <= 5e-6
    # restart from the disentangled checkpoint: 5 more iterations
    with open(os.path.join(tmp_path, seed + ".win"), "a") as f:
        f.write("restart = wannierise\n")
    r = run(seed, tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    chk2 = w90io.read_chk(os.path.join(tmp_path, seed + ".chk"))
    assert chk2["have_disentangled"] and np.array_equal(chk2["u_matrix_opt"], chk["u_matrix_opt"])
    _, _, om2 = w90io.read_wout_final(os.path.join(tmp_path, seed + ".wout"))
    assert om2["omega_total"] <= om["omega_total"] + 1e-8


@pytest.mark.gpu
@needs_exe
def test_driver_writes_u_matrices_of_the_disentanglement(tmp_path):
    """write_u_matrices after a disentanglement, on the reference's
    write_u_matrices_disent inputs (8 bands to 4 WFs, outer window only):
    _u_dis.mat holds U_opt with its window rows at their band positions and
    zeros elsewhere (plot_u_matrices, src/plot.F90:2266-2285), _u.mat the
    unitary U.  Compared with the reference's U_opt through the projector
    onto the optimal subspace, to the files' 10 decimals: U_opt's columns
    are the subspace Hamiltonian's eigenvectors, whose phases LAPACK and
    this build fix differently, and U_opt^+ A is rank-deficient here, so U
    itself is not unique."""
    from wannier90_amd import w90io

    seed, z = write_dis_inputs("dis_li_window", tmp_path, {"write_u_matrices": "true"})
    r = run(seed, tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr + open(os.path.join(tmp_path, seed + ".werr")).read()
    uopt, kpts = w90io.read_u_dis_mat(os.path.join(tmp_path, seed + "_u_dis.mat"))
    u, kpts2 = w90io.read_u_mat(os.path.join(tmp_path, seed + "_u.mat"))
    nb, nw = int(z["num_bands"]), int(z["num_wann"])
    nk = z["u_opt"].shape[2]
    assert uopt.shape == (nb, nw, nk) and u.shape == (nw, nw, nk)
    assert np.abs(kpts - z["kpt_latt"]).max() <= 1e-10 and np.abs(kpts2 - z["kpt_latt"]).max() <= 1e-10
    first = np.argmax(z["lwindow"], axis=0)
    nd = z["ndimwin"]
    ref = np.zeros_like(uopt)
    for k in range(nk):
        ref[first[k]: first[k] + nd[k], :, k] = z["u_opt"][: nd[k], :, k]
        outside = np.ones(nb, bool)
        outside[first[k]: first[k] + nd[k]] = False
        assert np.abs(uopt[outside, :, k]).max(initial=0.0) == 0.0
        assert np.abs(u[:, :, k].conj().T @ u[:, :, k] - np.eye(nw)).max() <= 1e-9
    p = np.einsum("iwk,jwk->ijk", uopt, uopt.conj())
    pr = np.einsum("iwk,jwk->ijk", ref, ref.conj())
    assert np.abs(p - pr).max() <= 1e-8


def test_pp_accepts_disentanglement_inputs(tmp_path):
    """-pp with num_bands > num_wann and the dis_* keywords: the k-mesh and
    the keyword checks of the disentanglement run before any GPU work."""
    if not os.path.exists(EXE):
        pytest.skip("Fortran host not built")
    seed, z = write_dis_inputs("dis_silicon", tmp_path)
    r = run(seed, tmp_path, "-pp")
    assert r.returncode == 0, r.stdout + r.stderr
    win = os.path.join(tmp_path, seed + ".win")
    txt = open(win).read()
    for bad, msg in (("dis_mix_ratio = 0.0", "dis_mix_ratio must be greater than 0.0"),
                     ("dis_froz_min = 1.0\ndis_froz_max = 0.5", "check disentanglement frozen windows")):
        lines = [l for l in txt.splitlines() if not l.startswith(bad.split()[0])]
        open(win, "w").write(bad + "\n" + "\n".join(lines) + "\n")
        r = run(seed, tmp_path, "-pp")
        assert r.returncode != 0
        assert msg in open(os.path.join(tmp_path, seed + ".werr")).read()


def _library_with_options(z, seedname):
    """wannier90_amd.library with the k-point keywords as options and the
    rest of seedname.win through the reader (the reference's order:
    input_setopt, then input_reader)."""
    from wannier90_amd import library

    lib = library.Library()
    lib.set_option("num_wann", int(z["num_wann"]))
    lib.set_option("num_bands", int(z["num_bands"]))
    lib.set_option("unit_cell_cart", np.asarray(z["real_lattice"], dtype=float))
    lib.set_option("mp_grid", np.asarray(z["mp_grid"], dtype=np.int32))
    lib.set_option("kpoints", np.asarray(z["kpt_latt"], dtype=float))
    assert lib.input_setopt(seedname) == 0
    assert lib.input_reader() == 0
    return lib


def test_python_library_interface_reads_options_and_win(tmp_path):
    """The Python mirror of the library interface (wannier90_amd.library,
    over libw90host.so's C interface): options, input_setopt, the .win
    through input_reader, and the k-mesh getters, which need no GPU; the
    reader with unspent options and no setopt is refused as the reference
    refuses it (src/library_interface.F90:312-318)."""
    if not os.path.exists(HOST_LIB):
        pytest.skip("Fortran host not built")
    from wannier90_amd import library

    seed, z = write_dis_inputs("dis_silicon", tmp_path)
    lib = _library_with_options(z, os.path.join(str(tmp_path), seed))
    nk = z["kpt_latt"].shape[0]
    assert lib.get_nn() == z["nnlist"].shape[1]
    assert np.array_equal(lib.get_nnkp(nk), z["nnlist"])
    lib.close()
    lib2 = library.Library()
    lib2.set_option("num_wann", int(z["num_wann"]))
    assert lib2.input_reader() > 0
    lib2.close()


@pytest.mark.gpu
@needs_exe
def test_python_library_interface_disentangles_like_the_reference(tmp_path):
    """The reference's silicon run (12 bands to 8 WFs, outer and frozen
    windows) from Python through the library interface: overlaps and band
    energies as caller-owned arrays, disentangle, wannierise, the getters;
    against the reference's run with the tolerances of the driver's test."""
    from wannier90_amd import capi

    seed, z = write_dis_inputs("dis_silicon", tmp_path)
    lib = _library_with_options(z, os.path.join(str(tmp_path), seed))
    nb, nw = int(z["num_bands"]), int(z["num_wann"])
    nk = z["kpt_latt"].shape[0]
    km = capi.kmesh_get(np.asarray(z["real_lattice"]), np.asarray(z["kpt_latt"]))
    m = capi.read_mmn(os.path.join(tmp_path, seed + ".mmn"), nb, km)
    a = capi.read_amn(os.path.join(tmp_path, seed + ".amn"), nb, nk, nw)
    eig = np.loadtxt(os.path.join(tmp_path, seed + ".eig"))[:, 2].reshape(nk, nb).T.copy(order="F")
    u = np.zeros((nw, nw, nk), dtype=np.complex128, order="F")
    lib.set_m_local(m)
    lib.set_u_matrix(u)
    lib.set_u_opt(a)
    lib.set_eigval(eig)
    assert lib.disentangle() == 0
    assert lib.wannierise() <= 0
    spr, om = lib.get_spreads(nw), lib.get_omega()
    lib.close()
    p = np.einsum("iwk,jwk->ijk", a, a.conj())
    pr = np.einsum("iwk,jwk->ijk", z["u_opt"], z["u_opt"].conj())
    assert np.abs(p - pr).max() <= 1e-8
    w = np.einsum("bwk,wvk->bvk", a, u)
    wr = np.einsum("bwk,wvk->bvk", z["u_opt"], z["u_final"])
    assert np.abs(w - wr).max() <= 1e-5
    assert np.abs(np.sort(spr) - np.sort(z["spreads"])).max() <= 5e-6
    assert abs(om[3] - float(z["omega"][3])) <= 2e-8
