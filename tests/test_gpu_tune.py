"""Rotation tile shapes and their autotuner on the GPU.

Every rotation kernel of every num_wann class above 16, fixed through
w90gpu_tune_set, is checked against the oracle on the large synthetic
problems: the single phases (rotation, wann_omega with the kernel's tiles,
gradient) to 1e-12 relative, and a 20-iteration trajectory to 1e-8 Ang^2 in
Omega at every iteration, 1e-6 Ang in the centres and 1e-7 in U (or 20x the
trajectory's own sensitivity to rounding, measured on the oracle, where that
is larger).  The FP32 instances of the 33-48-band kernels are checked in
the mixed bulk phase (fused48 against fused64 and against FP64).  The
online tuner (W90GPU_TUNE=1) is checked to time both
candidates of a key, write them to its file, keep the faster one only by its
3% rule, and give the same result bits on a second solve of the key.
"""
import os

import numpy as np
import pytest

from conftest import tol

pytestmark = pytest.mark.gpu

po = pytest.importorskip("oracle.pyoracle")

# (lattice, num_wann, MP grid) and the kernels of its class
CASES = [(("fcc", 24, (3, 3, 3)), ("fused32", "fused64")),
         (("bcc", 48, (3, 3, 3)), ("fused64", "direct64", "fused48")),
         (("sc", 40, (3, 3, 3)), ("fused48",)),
         (("hex", 70, (2, 2, 2)), ("gemm64", "gemm32")),
         (("sc", 96, (2, 2, 2)), ("gemm64", "gemm32")),
         (("sc", 160, (2, 2, 2)), ("gemm64", "gemm32"))]
PARAMS = [(case, k) for case, ks in CASES for k in ks]


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


class fixed_kernel:
    """w90gpu_tune_set for the duration of a block; the choice is dropped
    afterwards (back to the table)."""

    def __init__(self, nw, nn, name):
        self.key = (nw, nn)
        self.name = name

    def __enter__(self):
        from wannier90_amd import capi

        capi.tune_set(*self.key, self.name)
        assert capi.tune_get(*self.key) == self.name

    def __exit__(self, *exc):
        from wannier90_amd import capi

        capi.tune_set(*self.key, None)


def solve(ctx, p, u0, m0, ctl):
    from wannier90_amd import capi

    s = ctx.session(p.kmesh, p.num_wann, p.num_kpts, capi.control(**ctl))
    s.load(np.array(m0, order="F"), np.array(u0, order="F"))
    res = s.solve()
    hist = s.history()
    _, u, cen, spr = s.fetch(want_m=False)
    s.close()
    return res, hist, u, cen


def pid(v):
    (kind, nw, _), k = v
    return f"{kind}-nw{nw}-{k}"


@pytest.mark.parametrize("param", PARAMS, ids=pid)
def test_rotation_kernel_phases_match_oracle(ctx, param):
    case, kernel = param
    print(f"[tune phases] {case} {kernel}", flush=True)
    p, u0, m0 = problem(*case)
    op = oprob(p)
    with fixed_kernel(p.num_wann, p.kmesh.nntot, kernel):
        w = antihermitian(u0.shape, 0.05, 7)
        u_ref, m_ref = po.new_u_and_m(op, w, u0, m0)
        u = np.array(u0, order="F")
        m = np.array(m0, order="F")
        ctx.new_u_and_m(p.kmesh, w, u, m)
        assert np.abs(u - u_ref).max() <= 1e-12
        assert np.abs(m - m_ref).max() <= 1e-12 * max(1.0, np.abs(m_ref).max())
        for first in (True, False):
            ref_r, _, ref_om, _ = po.omega(op, m_ref, first, 0.5)
            r, _, om = ctx.omega(p.kmesh, np.asfortranarray(m_ref), first, 0.5)
            assert np.abs(om - ref_om).max() <= 1e-12 * max(1.0, abs(ref_om[3]))
            assert np.abs(r - ref_r).max() <= 1e-12 * max(1.0, np.abs(ref_r).max())
            r, _, om = ctx.omega_noalloc(p.kmesh, np.asfortranarray(m_ref), first, 0.5)
            assert np.abs(om - ref_om).max() <= 1e-12 * max(1.0, abs(ref_om[3]))


@pytest.mark.parametrize("param", PARAMS, ids=pid)
def test_rotation_kernel_trajectory_matches_oracle(ctx, param):
    case, kernel = param
    print(f"[tune trajectory] {case} {kernel}", flush=True)
    p, u0, m0 = problem(*case)
    ctl = dict(num_iter=20, num_cg_steps=5, conv_window=-1, conv_tol=1e-10, trial_step=2.0)
    o = po.wann_main(oprob(p), ctl, m0, u0)
    with fixed_kernel(p.num_wann, p.kmesh.nntot, kernel):
        res, hist, u, cen = solve(ctx, p, u0, m0, ctl)
    s_h = s_u = s_c = 0.0
    for seed in (1, 2):
        rng = np.random.default_rng(seed)
        o2 = po.wann_main(oprob(p), ctl, m0 * (1.0 + 1e-15 * rng.standard_normal(m0.shape)), u0)
        s_h = max(s_h, np.abs(o2["history"][:, 0] - o["history"][:, 0]).max())
        s_u = max(s_u, np.abs(o2["u"] - o["u"]).max())
        s_c = max(s_c, np.abs(o2["centres"] - o["centres"]).max())
    print(f"[tune trajectory] sensitivity: Omega {s_h:.1e} U {s_u:.1e} centres {s_c:.1e}; deviation "
          f"Omega {np.abs(hist[:, 0] - o['history'][:, 0]).max():.1e}", flush=True)
    assert res.iterations == o["iterations"] == 20
    assert np.abs(hist[:, 0] - o["history"][:, 0]).max() <= tol(1e-8, s_h)
    assert np.abs(cen - o["centres"]).max() <= tol(1e-6, s_c)
    assert np.abs(u - o["u"]).max() <= tol(1e-7, s_u)


def test_mixed_bulk_phase_with_the_48_tile(ctx):
    """The FP32 instances of the 33-48-band rotation kernels in the mixed
    bulk phase (num_wann 40): the Omega history of a 10-iteration mixed
    solve with fused48 equals the one with fused64 to 1e-8 relative (both
    form every fragment by the same FP32 MFMA sequence; only the order of
    the per-tile spread sums differs), and follows the FP64 solve's over
    the bulk iterations to 1e-4 relative (FP32 rounding of M)."""
    from wannier90_amd import capi

    case = ("sc", 40, (3, 3, 3))
    print(f"[tune mixed] {case}", flush=True)
    p, u0, m0 = problem(*case)
    nw, nn = p.num_wann, p.kmesh.nntot
    ctl = dict(num_iter=10, num_cg_steps=5, conv_window=-1, conv_tol=1e-10, trial_step=2.0)
    r64, h64, _, _ = solve(ctx, p, u0, m0, ctl)
    runs = {}
    for kernel in ("fused48", "fused64"):
        with fixed_kernel(nw, nn, kernel):
            runs[kernel] = solve(ctx, p, u0, m0, dict(ctl, precision=capi.PREC_MIXED))
    r48, h48, _, _ = runs["fused48"]
    rf64, hf64, _, _ = runs["fused64"]
    assert r48.unitarity_ok and rf64.unitarity_ok
    assert r48.iterations == rf64.iterations and r48.bulk_iterations == rf64.bulk_iterations
    nb = r48.bulk_iterations
    assert nb > 0
    print(f"[tune mixed] bulk {nb} of {r48.iterations}; Omega {h48[nb - 1, 0]:.10f} (fp64 {h64[nb - 1, 0]:.10f})",
          flush=True)
    scale = np.abs(h64[:nb, 0]).max()
    assert np.abs(h48[:, 0] - hf64[:, 0]).max() <= 1e-8 * scale
    assert np.abs(h48[:nb, 0] - h64[:nb, 0]).max() <= 1e-4 * scale


def test_online_tuner_times_candidates_and_keeps_its_choice(tmp_path):
    """W90GPU_TUNE=1 on a key without a choice: the first solve times both
    class-3 kernels and appends them to W90GPU_TUNE_FILE; the kept kernel is
    the default unless the other is at least 3% faster; a second solve of
    the key takes the same kernel, bit for bit, and both match the oracle."""
    from wannier90_amd import capi

    case = ("sc", 100, (2, 2, 2))
    p, u0, m0 = problem(*case)
    nw, nn = p.num_wann, p.kmesh.nntot
    path = str(tmp_path / "tune.txt")
    capi.tune_set(nw, nn, None)
    ctl = dict(num_iter=5, num_cg_steps=5, conv_window=-1, conv_tol=1e-10, trial_step=2.0)
    os.environ["W90GPU_TUNE"] = "1"
    os.environ["W90GPU_TUNE_FILE"] = path
    try:
        c = capi.Context(0)  # the environment is read when a context or engine is created
        print(f"[tune online] {case}", flush=True)
        r1, h1, u1, cen1 = solve(c, p, u0, m0, ctl)
        r2, h2, u2, cen2 = solve(c, p, u0, m0, ctl)
        c.close()
    finally:
        del os.environ["W90GPU_TUNE"]
        del os.environ["W90GPU_TUNE_FILE"]
        # the library reads the environment at context creation: one more
        # context turns the online tuner off for the rest of the process
        capi.Context(0).close()
    lines = [ln.split() for ln in open(path) if ln.strip()]
    mine = [ln for ln in lines if ln[:2] == [str(nw), str(nn)]]
    assert len(mine) == 1, lines
    chosen = mine[0][2]
    times = dict(f.split("=") for f in mine[0][3:])
    t64, t32 = float(times["gemm64"]), float(times["gemm32"])
    print(f"[tune online] gemm64 {t64:.1f} us, gemm32 {t32:.1f} us -> {chosen}", flush=True)
    assert t64 > 0 and t32 > 0
    assert chosen == ("gemm32" if t32 <= 0.97 * t64 else "gemm64")
    assert capi.tune_get(nw, nn) == chosen
    np.testing.assert_array_equal(h1, h2)
    np.testing.assert_array_equal(u1, u2)
    o = po.wann_main(oprob(p), ctl, m0, u0)
    assert np.abs(h1[:, 0] - o["history"][:, 0]).max() <= 1e-8
    assert np.abs(cen1 - o["centres"]).max() <= 1e-6
    capi.tune_set(nw, nn, None)
