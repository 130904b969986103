"""num_wann above 16 (17-128): the MFMA GEMM rotation, checked
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
# the LDS retraction (33-64) and with the global-memory retraction (65-128)
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


@pytest.mark.parametrize("case", LARGE, ids=ids)
def test_large_phases_match_oracle(ctx, case):
    print(f"[large phases] {case}", flush=True)
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


@pytest.mark.parametrize("case", LARGE[:7], ids=ids)
def test_large_trajectory_matches_oracle(ctx, case):
    from wannier90_amd import capi

    print(f"[large trajectory] {case}", flush=True)
    p, u0, m0 = problem(*case)
    ctl = dict(num_iter=20, num_cg_steps=5, conv_window=-1, conv_tol=1e-10, trial_step=2.0)
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
    assert res.iterations == o["iterations"] == 20
    assert np.abs(hist[:, 0] - o["history"][:, 0]).max() <= tol(1e-8, s_h)
    assert abs(res.omega_total - o["omega"][3]) <= tol(1e-8, s_h)
    assert np.abs(cen - o["centres"]).max() <= tol(1e-6, s_c)
    assert np.abs(u - o["u"]).max() <= tol(1e-7, s_u)
