"""CPU checks of the oracle, the k-mesh and the library boundary.

The oracle (oracle/wannierise_oracle.c) is pinned here against
  * the reference wannier90.x outputs stored in tests/golden (its own
    regression inputs and seeded synthetic problems), per iteration and at
    the end of the run,
  * mathematics: the closed-form minimum of the tight-binding problems
    (Omega = 0, centres = orbital positions), Omega_I = 0 for unitary M,
    sum of spreads = Omega, the gradient against finite differences of
    Omega, unitarity and the symmetry M(k,b)^+ = M(k+b,-b) under rotations.
No GPU is needed.
"""
import os
import re

import numpy as np
import pytest

from conftest import ROOT, Case, golden_names, stored_golden_names, tol

po = pytest.importorskip("oracle.pyoracle")


def _oracle_ok():
    try:
        po.load()
        return True
    except OSError:
        return False


needs_oracle = pytest.mark.skipif(not _oracle_ok(), reason="oracle not built (make -C oracle)")
FULL = stored_golden_names()


def oracle_problem(c):
    return po.Problem(c.nw, c.nk, c.kmesh)


# ------------------------------------------------------------------ oracle

@needs_oracle
@pytest.mark.parametrize("name", FULL)
def test_oracle_trajectory_matches_reference(name):
    """Omega after every iteration equals the reference's CONV lines (printed
    with 10 decimals) for all iterations the trajectory is pinned."""
    c = Case(name)
    r = po.wann_main(oracle_problem(c), c.ctl, c.m0, c.u0)
    pinned = int(c.z["pinned_iters"])
    it = c.z["conv_iter"]
    om = c.z["conv_omega"]
    sel = it <= min(pinned, r["iterations"])
    got = r["history"][it[sel], 0]
    assert np.abs(got - om[sel]).max() <= 2e-10, name
    if c.well_conditioned():
        assert r["iterations"] == int(c.z["iterations"])


@needs_oracle
@pytest.mark.parametrize("name", FULL)
def test_oracle_final_state_matches_reference(name):
    """The final state equals the reference's to the plain-path bounds, or
    to 20x the trajectory's measured sensitivity to rounding where that is
    larger.  On a path that passes saddles (rounding moves the final U by
    more than 1e-3), the reference converged to the closed-form global
    minimum and the oracle, given the iterations to converge, reaches the
    same minimum."""
    c = Case(name)
    r = po.wann_main(oracle_problem(c), c.ctl, c.m0, c.u0)
    if c.sens_u <= 1e-3:
        assert np.abs(r["u"] - c.z["u_final"]).max() <= tol(1e-7, c.sens_u)
        assert np.abs(r["m"] - c.z["m_final"]).max() <= tol(1e-7, c.sens_u)
        assert np.abs(r["centres"] - c.z["centres"]).max() <= tol(1e-6, c.sens_centres)
        # .wout prints Omega with 9 decimals and the .chk keeps spreads in full
        assert abs(r["omega"][3] - c.z["omega"][3]) <= 1e-9 + tol(1e-8, c.sens_omega)
        assert np.abs(r["spreads"] - c.z["spreads"]).max() <= tol(1e-8, c.sens_omega)
        return
    if int(c.z["iterations"]) >= c.ctl["num_iter"]:
        pytest.skip("the reference run ends unconverged on a path that rounding changes "
                    "(sensitivity of Omega %.1e); no final state to compare" % c.sens_omega)
    assert abs(c.z["omega"][3]) <= 1e-8
    r = po.wann_main(oracle_problem(c), dict(c.ctl, num_iter=6000), c.m0, c.u0)
    assert r["converged"] and abs(r["omega"][3] - c.z["omega"][3]) <= 1e-8
    tau, lat = c.z["tau_cart"], c.z["real_lattice"]
    inv = np.linalg.inv(lat)
    for cen in (r["centres"], c.z["centres"]):
        hit = []
        for x in cen.T:
            d = (x - tau) @ inv
            d -= np.round(d)
            dist = np.linalg.norm(d @ lat, axis=1)
            assert dist.min() < 1e-6
            hit.append(int(np.argmin(dist)))
        assert sorted(hit) == list(range(c.nw))


@needs_oracle
def test_closed_form_minimum():
    """cfg1 reaches the global minimum in the reference: Omega = 0 and the
    centres are the orbital positions (up to lattice vectors and order)."""
    c = Case("cfg1")
    assert abs(c.z["omega"][3]) < 1e-8
    tau = c.z["tau_cart"]
    lat = c.z["real_lattice"]
    cen = c.z["centres"].T
    inv = np.linalg.inv(lat)
    for x in cen:
        d = (x - tau) @ inv
        d -= np.round(d)
        assert np.min(np.linalg.norm(d @ lat, axis=1)) < 1e-6
    # the oracle's Omega at the reference's final M is zero too
    rave, r2ave, om, _ = po.omega(oracle_problem(c), c.z["m_final"])
    assert abs(om[3]) < 1e-9 and np.allclose(rave, c.z["centres"], atol=1e-8)


@needs_oracle
@pytest.mark.parametrize("name", ["cfg1", "bcc_nw6", "hex_nw5", "fcc_nw8", "sc_nw1"])
def test_omega_invariants(name):
    c = Case(name)
    pr = oracle_problem(c)
    rave, r2ave, om, _ = po.omega(pr, c.m0)
    # unitary overlaps of a complete tight-binding basis: Omega_I = 0
    assert abs(om[0]) < 1e-10
    # sum of spreads = Omega (the .wout "Sum of centres and spreads" line)
    spreads = r2ave - (rave ** 2).sum(axis=0)
    assert abs(spreads.sum() - om[3]) < 1e-9 * max(1.0, abs(om[3]))
    # Omega_I is unchanged by a rotation
    rng = np.random.default_rng(3)
    w = rng.normal(size=c.u0.shape) + 1j * rng.normal(size=c.u0.shape)
    w = 0.05 * (w - np.conj(np.transpose(w, (1, 0, 2))))
    _, m1 = po.new_u_and_m(pr, np.asfortranarray(w), c.u0, c.m0)
    assert abs(po.omega(pr, m1)[2][0] - om[0]) < 1e-10


@needs_oracle
@pytest.mark.parametrize("name", ["ref_basic1", "ref_example05", "hex_nw5", "fcc_nw8"])
def test_gradient_finite_difference(name):
    """d Omega(U exp(eps W)) / d eps at 0 equals -Re<G, W> (the slope used by
    internal_search_direction), checked by a central difference."""
    c = Case(name)
    pr = oracle_problem(c)
    g, _ = po.domega(pr, c.m0)
    rng = np.random.default_rng(7)
    w = rng.normal(size=c.u0.shape) + 1j * rng.normal(size=c.u0.shape)
    w = np.asfortranarray(w - np.conj(np.transpose(w, (1, 0, 2))))
    h = 1e-5
    _, mp = po.new_u_and_m(pr, np.asfortranarray(h * w), c.u0, c.m0)
    _, mm = po.new_u_and_m(pr, np.asfortranarray(-h * w), c.u0, c.m0)
    om0 = po.omega(pr, c.m0)[2][0]
    fd = (po.omega(pr, mp, False, om0)[2][3] - po.omega(pr, mm, False, om0)[2][3]) / (2 * h)
    slope = -np.real(np.vdot(g, w))
    assert abs(fd - slope) <= 1e-6 * max(1.0, abs(slope)), (fd, slope)
    # the gradient is anti-Hermitian
    assert np.abs(g + np.conj(np.transpose(g, (1, 0, 2)))).max() < 1e-12


@needs_oracle
@pytest.mark.parametrize("name", ["ref_example02", "bcc_nw6"])
def test_rotation_keeps_unitarity_and_symmetry(name):
    c = Case(name)
    pr = oracle_problem(c)
    rng = np.random.default_rng(11)
    w = rng.normal(size=c.u0.shape) + 1j * rng.normal(size=c.u0.shape)
    w = np.asfortranarray(0.3 * (w - np.conj(np.transpose(w, (1, 0, 2)))))
    u1, m1 = po.new_u_and_m(pr, w, c.u0, c.m0)
    assert po.unitarity(pr, u1) < 1e-12
    # M(k,b)^+ = M(k+b,-b) holds before and after the rotation
    km = c.kmesh
    bk = km.bk
    for k in range(0, c.nk, 7):
        for b in range(c.nn):
            k2 = km.nnlist[k, b] - 1
            bneg = [bb for bb in range(c.nn) if km.nnlist[k2, bb] - 1 == k
                    and np.allclose(bk[:, bb, k2], -bk[:, b, k], atol=1e-8)][0]
            for m in (c.m0, m1):
                assert np.abs(np.conj(m[:, :, b, k].T) - m[:, :, bneg, k2]).max() < 1e-10


@needs_oracle
def test_check_unitarity_closed_forms():
    """wann_check_unitarity's deviation max |UU^+ - 1|, |U^+U - 1| on
    non-unitary U with known deviations: a block scaled by (1 + d) deviates
    by 2d + d^2 on the diagonal, and the identity with one off-diagonal entry
    e deviates by e (off-diagonal) when e < 1; the largest block decides."""
    c = Case("ref_example05")
    pr = oracle_problem(c)
    u = np.array(c.u0, order="F")
    assert po.unitarity(pr, u) < 1e-13
    for d in (1e-6, 3e-4, 0.25):
        v = np.array(u, order="F")
        v[:, :, 5] *= 1.0 + d
        assert abs(po.unitarity(pr, v) - (2 * d + d * d)) <= 1e-12 * max(1.0, d)
    for e in (2e-5, 0.1, 0.6):
        v = np.array(u, order="F")
        v[:, :, 7] = np.eye(c.nw)
        v[2, 0, 7] = e
        assert abs(po.unitarity(pr, v) - e) <= 1e-15
        # a second, smaller deviation elsewhere does not change the maximum
        v[:, :, 9] *= 1.0 + 0.1 * e
        assert abs(po.unitarity(pr, v) - e) <= 1e-15
    # a phase change of one column keeps U unitary
    v = np.array(u, order="F")
    v[:, 1, 3] *= np.exp(0.7j)
    assert po.unitarity(pr, v) < 1e-13


@needs_oracle
def test_exp_of_antihermitian_matches_series():
    """exp(cdq) built from zheev agrees with a long Taylor series in extended
    precision (np.longdouble) on a small random step."""
    c = Case("ref_basic1")
    pr = oracle_problem(c)
    rng = np.random.default_rng(5)
    w = rng.normal(size=c.u0.shape) + 1j * rng.normal(size=c.u0.shape)
    w = np.asfortranarray(0.7 * (w - np.conj(np.transpose(w, (1, 0, 2)))))
    eye = np.asfortranarray(np.repeat(np.eye(c.nw)[:, :, None], c.nk, axis=2).astype(np.complex128))
    u1, _ = po.new_u_and_m(pr, w, eye, c.m0)
    for k in range(c.nk):
        x = w[:, :, k].astype(np.clongdouble)
        term = np.eye(c.nw, dtype=np.clongdouble)
        acc = term.copy()
        for j in range(1, 60):
            term = term @ x / j
            acc = acc + term
        assert np.abs(u1[:, :, k] - acc.astype(np.complex128)).max() < 1e-13


# ------------------------------------------------------------------ k-mesh

@pytest.mark.parametrize("name", FULL)
def test_kmesh_reproduces_reference_tables(name):
    """The C++ k-mesh reproduces nnlist/nncell of the reference (checked
    against its .nnkp when the fixture was made) and its b-vectors/weights."""
    from wannier90_amd import capi

    c = Case(name)
    km = capi.kmesh_get(c.z["real_lattice"], c.z["kpt_latt"])
    assert km.nntot == c.nn
    assert np.array_equal(km.nnlist, c.kmesh.nnlist)
    assert np.array_equal(km.nncell, c.kmesh.nncell)
    np.testing.assert_allclose(km.bk, c.kmesh.bk, atol=1e-12)
    np.testing.assert_allclose(km.wb, c.kmesh.wb, rtol=1e-12)
    # B1 condition: sum_b wb b_i b_j = delta_ij at every k
    for k in range(c.nk):
        b1 = np.einsum("b,ib,jb->ij", km.wb, km.bk[:, :, k], km.bk[:, :, k])
        assert np.abs(b1 - np.eye(3)).max() < 1e-10


def test_kmesh_reference_wout_tables():
    """testw90_basic1's .wout lists 8 b-vectors (+-0.553079, +-0.553079,
    +-0.553079) 1/Ang with w_b = 0.408635 Ang^2
    (benchmark.out.default.inp=wannier.win:171-182)."""
    c = Case("ref_basic1")
    assert c.nn == 8
    np.testing.assert_allclose(np.abs(c.kmesh.bk[:, :, 0]), 0.553079, atol=1e-6)
    np.testing.assert_allclose(c.kmesh.wb, 0.408635, atol=1e-6)


def test_kmesh_closed_forms():
    from wannier90_amd import capi, synth

    a = 5.0
    km = capi.kmesh_get(synth.lattice("sc", a), synth.mp_kpoints((4, 4, 4)))
    b = 2 * np.pi / (4 * a)
    assert km.nntot == 6
    np.testing.assert_allclose(km.wb, 1 / (2 * b * b), rtol=1e-12)
    np.testing.assert_allclose(km.wbtot, 3 / (b * b), rtol=1e-12)
    assert capi.kmesh_get(synth.lattice("bcc", 4.0), synth.mp_kpoints((4, 4, 4))).nntot == 12
    assert capi.kmesh_get(synth.lattice("fcc", 5.43), synth.mp_kpoints((4, 4, 4))).nntot == 8
    assert capi.kmesh_get(synth.lattice("hex", 4.0, 1.5), synth.mp_kpoints((4, 4, 4))).nntot == 8


def test_kmesh_rejects_bad_input():
    from wannier90_amd import capi

    with pytest.raises(capi.W90Error):
        capi.kmesh_get(np.zeros((3, 3)), np.zeros((1, 3)))


# --------------------------------------------------------------- boundary

def _header_functions():
    txt = open(os.path.join(ROOT, "include", "w90gpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(w90gpu_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from wannier90_amd import capi

    lib = capi.load()
    names = _header_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    declared = {n for n, _, _ in capi.SIGNATURES}
    assert set(names) == declared


def test_control_defaults_follow_reference():
    from wannier90_amd import capi

    c = capi.control()
    assert (c.num_iter, c.num_cg_steps, c.conv_window) == (100, 5, -1)
    assert (c.conv_tol, c.trial_step, c.fixed_step) == (1e-10, 2.0, -999.0)


def test_session_bytes_matches_footprint():
    """Admission by HBM bytes: the session holds 3 M buffers and 8 nw^2 nk
    blocks plus small tables."""
    from wannier90_amd import capi

    nw, nk, nn = 16, 512, 8
    b = capi.session_bytes(nw, nk, nn)
    core = 16 * nk * nw * nw * (3 * nn + 8)
    assert core <= b <= core * 1.2 + (1 << 20)


def test_no_gpu_raises_loudly():
    """Without a HIP device the library refuses to create a context; there is
    no CPU fallback."""
    from wannier90_amd import capi

    try:
        ctx = capi.Context(0)
    except capi.W90Error as e:
        assert e.code == capi.ERR_FATAL
        return
    ctx.close()
    pytest.skip("GPU present")


# ------------------------------------------------------------ projection

LOWDIN_CASES = {
    "cfg1": lambda s: s.config(1),
    "bcc_nw6": lambda s: s.make_problem(77, kind="bcc", a=4.2, nw=6, mp=(4, 4, 4), decay=1.5),
    "fcc_nw8": lambda s: s.make_problem(79, kind="fcc", a=5.43, nw=8, mp=(4, 4, 4), decay=2.0),
}


def lowdin_inputs(name):
    """The original A and M of a synthetic fixture, regenerated from its seed."""
    from wannier90_amd import synth

    prob = LOWDIN_CASES[name](synth)
    a = np.asfortranarray(np.transpose(prob.amat, (1, 2, 0)))
    return prob, a, np.asfortranarray(prob.m_orig)


@needs_oracle
@pytest.mark.parametrize("name", sorted(LOWDIN_CASES))
def test_oracle_lowdin_matches_reference(name):
    """overlap_project restated with zgesvd reproduces the reference's initial
    U and M (its num_iter = 0 checkpoint) on the same .amn/.mmn data."""
    c = Case(name)
    prob, a, m = lowdin_inputs(name)
    u, m1 = po.lowdin(oracle_problem(c), a, m)
    assert np.abs(u - c.u0).max() < 1e-12
    assert np.abs(m1 - c.m0).max() < 1e-12


def test_generator_does_not_load_the_gpu_runtime():
    """Problem generation (k-mesh search included) loads only the host
    library: no HIP runtime in the process; threads give the results of a
    sequential run."""
    import subprocess
    import sys

    code = ("import sys; sys.path.insert(0, %r)\n"
            "from wannier90_amd import workloads\n"
            "p = workloads._generate(workloads.batch_params(3))\n"
            "maps = open('/proc/self/maps').read()\n"
            "print(p['nw'], p['nn'], 'amdhip64' in maps, 'libw90gpu' in maps, 'libw90io' in maps)\n" % ROOT)
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    nw, nn, hip, gpu, host = out.stdout.split()
    assert (hip, gpu, host) == ("False", "False", "True")
    import numpy as np

    from wannier90_amd import workloads

    plist = [workloads.batch_params(i) for i in (15, 19, 28, 29)]
    a = workloads.generate(plist, workers=1)
    b = workloads.generate(plist, workers=4)
    for x, y in zip(a, b):
        assert np.array_equal(x["m0"], y["m0"]) and np.array_equal(x["u0"], y["u0"])
