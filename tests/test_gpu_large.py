"""num_wann above 16 (17-256): the MFMA GEMM rotation, checked
against the oracle on seeded synthetic problems (sc, bcc, fcc, triclinic
lattices; nntot 6-12) small enough in k for the oracle to finish in seconds.

Tolerances: single phases 1e-12 relative (FP64 MFMA and the oracle's zgemm
sum in different orders); 20-iteration trajectories 1e-8 Ang^2 in Omega at
every iteration, 1e-6 Ang in the centres and 1e-7 in U, or 20x the
trajectory's own sensitivity to rounding where that is larger (measured in
the test by rerunning the oracle from M0 perturbed by 1e-15 relative).
"""
import numpy as np
import pytest

from conftest import tol

pytestmark = pytest.mark.gpu

po = pytest.importorskip("oracle.pyoracle")

# (lattice, num_wann, MP grid): 32x32 MFMA tiles (17-32), 64x64 tiles with
# the LDS retraction (33-64) and with the global-memory retraction (65-256)
LARGE = [("sc", 40, (3, 3, 3)), ("bcc", 48, (3, 3, 3)), ("fcc", 33, (2, 2, 2)),
         ("triclinic", 64, (2, 2, 3)), ("fcc", 24, (3, 3, 3)), ("hex", 17, (3, 3, 3)),
         ("bcc", 32, (3, 3, 3)), ("sc", 96, (2, 2, 2)), ("hex", 70, (2, 2, 2))]


@pytest.fixture(scope="module")
def ctx():
    from wannier90_amd import capi

    c = capi.Context(0)
    yield c
    c.close()


_cache = {}


def problem(kind, nw, mp):
    from wannier90_amd import synth

    key = (kind, nw, mp)
    if key not in _cache:
        p = synth.make_problem(4242 + nw, kind=kind, nw=nw, mp=mp, decay=2.0)
        u, m = synth.initial_state(p)
        _cache[key] = (p, u, m)
    return _cache[key]


def oprob(p):
    return po.Problem(p.num_wann, p.num_kpts, p.kmesh)


def antihermitian(shape, scale, seed):
    rng = np.random.default_rng(seed)
    w = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    return np.asfortranarray(scale * (w - np.conj(np.transpose(w, (1, 0, 2)))))


def ids(v):
    return f"{v[0]}-nw{v[1]}"


def check_phases(ctx, case):
    p, u0, m0 = problem(*case)
    op = oprob(p)
    w = antihermitian(u0.shape, 0.05, 7)
    u_ref, m_ref = po.new_u_and_m(op, w, u0, m0)
    u = np.array(u0, order="F")
    m = np.array(m0, order="F")
    ctx.new_u_and_m(p.kmesh, w, u, m)
    assert np.abs(u - u_ref).max() <= 1e-12
    assert np.abs(m - m_ref).max() <= 1e-12 * max(1.0, np.abs(m_ref).max())
    for first in (True, False):
        ref_r, ref_r2, ref_om, _ = po.omega(op, m_ref, first, 0.5)
        r, r2, om = ctx.omega(p.kmesh, np.asfortranarray(m_ref), first, 0.5)
        assert np.abs(om - ref_om).max() <= 1e-12 * max(1.0, abs(ref_om[3]))
        assert np.abs(r - ref_r).max() <= 1e-12 * max(1.0, np.abs(ref_r).max())
    g_ref, _ = po.domega(op, m_ref)
    g = ctx.domega(p.kmesh, np.asfortranarray(m_ref))
    assert np.abs(g - g_ref).max() <= 1e-12 * max(1.0, np.abs(g_ref).max())


def check_trajectory(ctx, case, num_iter):
    from wannier90_amd import capi

    p, u0, m0 = problem(*case)
    ctl = dict(num_iter=num_iter, num_cg_steps=5, conv_window=-1, conv_tol=1e-10, trial_step=2.0)
    o = po.wann_main(oprob(p), ctl, m0, u0)
    s = ctx.session(p.kmesh, p.num_wann, p.num_kpts, capi.control(**ctl))
    s.load(np.array(m0, order="F"), np.array(u0, order="F"))
    res = s.solve()
    hist = s.history()
    _, u, cen, spr = s.fetch(want_m=False)
    s.close()
    # the trajectory's own sensitivity to rounding: the oracle rerun from M0
    # perturbed by 1e-15 relative (as the golden fixtures measure it)
    s_h = s_u = s_c = 0.0
    for seed in (1, 2):
        rng = np.random.default_rng(seed)
        o2 = po.wann_main(oprob(p), ctl, m0 * (1.0 + 1e-15 * rng.standard_normal(m0.shape)), u0)
        s_h = max(s_h, np.abs(o2["history"][:, 0] - o["history"][:, 0]).max())
        s_u = max(s_u, np.abs(o2["u"] - o["u"]).max())
        s_c = max(s_c, np.abs(o2["centres"] - o["centres"]).max())
    print(f"[large trajectory] sensitivity: Omega {s_h:.1e} U {s_u:.1e} centres {s_c:.1e}", flush=True)
    assert res.iterations == o["iterations"] == num_iter
    assert np.abs(hist[:, 0] - o["history"][:, 0]).max() <= tol(1e-8, s_h)
    assert abs(res.omega_total - o["omega"][3]) <= tol(1e-8, s_h)
    assert np.abs(cen - o["centres"]).max() <= tol(1e-6, s_c)
    assert np.abs(u - o["u"]).max() <= tol(1e-7, s_u)


@pytest.mark.parametrize("case", LARGE, ids=ids)
def test_large_phases_match_oracle(ctx, case):
    print(f"[large phases] {case}", flush=True)
    check_phases(ctx, case)


@pytest.mark.parametrize("case", LARGE, ids=ids)
def test_large_trajectory_matches_oracle(ctx, case):
    print(f"[large trajectory] {case}", flush=True)
    check_trajectory(ctx, case, 20)


def test_largest_num_wann_matches_oracle(ctx):
    """num_wann = 256, the largest the library takes (the per-k partial
    sums give each Wannier function a thread of a 256-thread workgroup):
    the single phases and a 5-iteration trajectory on a 2x2x2 simple-cubic
    mesh (64 x 64 panels, four per dimension)."""
    case = ("sc", 256, (2, 2, 2))
    print(f"[largest num_wann] {case}", flush=True)
    check_phases(ctx, case)
    check_trajectory(ctx, case, 5)


def test_triclinic_800k_matches_reference(ctx):
    """The triclinic 64-band problem on its own 10x10x8 mesh (800 k, nntot
    12), regenerated from its seed: Omega after each of the reference's 40
    iterations (restarted from the same initial state), and its final
    centres, spreads and U blocks; the first 3 iterations against the
    oracle; a full 500-iteration solve keeps the invariants (Omega_I = 0,
    U unitary, spreads summing to Omega)."""
    from conftest import Case
    from wannier90_amd import capi, synth

    print("[cfg3] triclinic nw 64, 800 k", flush=True)
    c = Case("cfg3")
    prob = synth.config(3)
    assert np.array_equal(prob.kmesh.nnlist, c.kmesh.nnlist) and prob.kmesh.nntot == 12
    u0, m0 = synth.initial_state(prob)
    s = ctx.session(prob.kmesh, prob.num_wann, prob.num_kpts, c.gpu_control())
    s.load(m0, u0)
    res = s.solve()
    hist = s.history()
    _, u, cen, spr = s.fetch(want_m=False)
    s.close()
    print(f"[cfg3] {res.iterations} iterations, Omega {res.omega_total:.10f} "
          f"(reference {c.z['omega'][3]:.10f})", flush=True)
    assert res.status == capi.OK and res.iterations == int(c.z["iterations"])
    it = c.z["conv_iter"]
    sel = it <= int(c.z["pinned_iters"])
    # the plain-path bound (1e-8 Ang^2) at every iteration: Omega ~ 349 is a
    # sum of 800 x 12 x 64 terms, so summation order alone moves it by ~1e-9
    dev = np.abs(hist[it[sel], 0] - c.z["conv_omega"][sel])
    print(f"[cfg3] Omega vs reference per iteration: max {dev.max():.2e} (first {dev[1]:.2e}, "
          f"last {dev[-1]:.2e})", flush=True)
    assert dev.max() <= 1e-8
    assert abs(res.omega_total - c.z["omega"][3]) <= 1e-9 + tol(1e-8, c.sens_omega)
    assert np.abs(cen - c.z["centres"]).max() <= tol(1e-6, c.sens_centres)
    assert np.abs(spr - c.z["spreads"]).max() <= tol(1e-8, c.sens_omega)
    assert np.abs(u[:, :, :8] - c.z["u_final"]).max() <= tol(1e-7, c.sens_u)

    # the first 3 iterations against the oracle, U included
    ctl3 = dict(num_iter=3, num_cg_steps=5, conv_window=3, conv_tol=1e-10, trial_step=2.0)
    o = po.wann_main(oprob(prob), ctl3, m0, u0)
    s = ctx.session(prob.kmesh, prob.num_wann, prob.num_kpts, capi.control(**ctl3))
    s.load(m0, u0)
    r3 = s.solve()
    h3 = s.history()
    _, u3, cen3, _ = s.fetch(want_m=False)
    s.close()
    assert r3.iterations == o["iterations"] == 3
    assert np.abs(h3[:, 0] - o["history"][:, 0]).max() <= 1e-8
    assert np.abs(u3 - o["u"]).max() <= 1e-7
    assert np.abs(cen3 - o["centres"]).max() <= 1e-6

    # the full solve of the benchmark configuration
    s = ctx.session(prob.kmesh, prob.num_wann, prob.num_kpts,
                    capi.control(num_iter=500, num_cg_steps=5, conv_window=3, conv_tol=1e-10, trial_step=2.0))
    s.load(m0, u0)
    rf = s.solve()
    _, _, _, sprf = s.fetch(want_m=False)
    s.close()
    print(f"[cfg3] 500-iteration solve: {rf.iterations} iterations, Omega {rf.omega_total:.6f} "
          f"in {rf.solve_seconds:.3f} s", flush=True)
    assert rf.status == capi.OK and rf.unitarity_ok == 1
    assert abs(rf.omega_invariant) <= 1e-10
    assert abs(sprf.sum() - rf.omega_total) <= 1e-9 * rf.omega_total
    assert rf.omega_total < hist[0, 0]


def test_num_wann_above_the_limit_is_rejected(ctx):
    """num_wann above W90GPU_MAX_WANN (256) is an input error, as every other
    size the kernels do not take."""
    from wannier90_amd import capi

    p, _, _ = problem("sc", 40, (2, 2, 2))
    with pytest.raises(capi.W90Error, match=r"num_wann=257 outside \[1, 256\]"):
        ctx.session(p.kmesh, capi.MAX_WANN + 1, p.num_kpts, capi.control(num_iter=1))
