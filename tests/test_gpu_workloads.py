"""The ragged batch and the serving loop on subsets of their own problem
distributions (wannier90_amd.workloads).

Ragged batch: 24 problems of the 256 (sc/fcc/bcc/hex, num_wann U[4,48],
meshes 4^3/6^3/8^3), two or three from each pair of num_wann class (<= 16,
17-32, 33-48) and mesh, solved as one batch and one per launch: every
result bit-identical (the batch sums each problem in its own fixed order);
the small ones equal the oracle over 20 iterations (Omega 1e-8 Ang^2 at
every iteration, centres 1e-6 Ang, U 1e-7, or 20x the trajectory's own
measured sensitivity to rounding where that is larger).

Serving: requests of the serving distribution (num_wann up to 96, so the
65-256 kernel class runs), arriving while others are on the device, under
the continuous-batching loop of the serving workload: each result equals
the same request solved alone.
"""
import numpy as np
import pytest

from conftest import tol

pytestmark = pytest.mark.gpu

po = pytest.importorskip("oracle.pyoracle")

NUM_ITER = 100


def _wclass(nw):
    return 0 if nw <= 16 else 1 if nw <= 32 else 2


def batch_subset():
    """Indices into the 256-problem batch: 3, 3 and 2 from each num_wann
    class on the 4^3, 6^3 and 8^3 meshes."""
    from wannier90_amd import workloads

    want = {4: 3, 6: 3, 8: 2}
    got, out = {}, []
    for i in range(256):
        p = workloads.batch_params(i)
        key = (_wclass(p["nw"]), p["mp"][0])
        if got.get(key, 0) < want[p["mp"][0]]:
            got[key] = got.get(key, 0) + 1
            out.append(i)
    assert len(out) == 24
    return out


@pytest.fixture(scope="module")
def ctx():
    from wannier90_amd import capi

    c = capi.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def batch_problems():
    from wannier90_amd import workloads

    ids = batch_subset()
    # generated in this process: no worker processes next to a GPU context
    return ids, workloads.generate([workloads.batch_params(i) for i in ids], workers=1)


def solve_alone(ctx, p, ctl):
    from wannier90_amd import workloads

    s = ctx.session(workloads.kmesh_of(p), p["nw"], p["nk"], ctl)
    s.load(p["m0"], p["u0"])
    r = s.solve()
    hist = s.history()
    _, u, cen, spr = s.fetch(want_m=False)
    s.close()
    return r, u, cen, spr, hist


def test_ragged_batch_subset_equals_alone(ctx, batch_problems):
    from wannier90_amd import capi, workloads

    ids, probs = batch_problems
    print(f"[ragged] {len(probs)} problems, num_wann {sorted(p['nw'] for p in probs)}", flush=True)
    ctl = workloads.control(NUM_ITER, False)
    alone = [solve_alone(ctx, p, ctl) for p in probs]
    pool = sum(capi.problem_bytes(p["nw"], p["nk"], p["nn"], NUM_ITER) for p in probs)
    b = ctx.batch(pool)
    hs = [b.add(workloads.kmesh_of(p), p["m0"], p["u0"], ctl) for p in probs]
    assert None not in hs
    b.run()
    for i, h, p, a in zip(ids, hs, probs, alone):
        r = b.result(h)
        _, u, cen, spr = b.fetch(h)
        what = f"batch item {i} ({p['kind']}, num_wann {p['nw']}, {p['mp'][0]}^3)"
        assert r.status == capi.OK and r.unitarity_ok == 1, what
        assert r.iterations == a[0].iterations and r.omega_total == a[0].omega_total, what
        assert np.array_equal(u, a[1]) and np.array_equal(cen, a[2]) and np.array_equal(spr, a[3]), what
        assert abs(r.omega_invariant) <= 1e-10, what
    b.close()


def test_ragged_batch_small_problems_match_oracle(ctx, batch_problems):
    from wannier90_amd import capi, workloads

    ids, probs = batch_problems
    ctl = dict(num_iter=20, num_cg_steps=5, conv_window=3, conv_tol=1e-10, trial_step=2.0)
    n = 0
    for i, p in zip(ids, probs):
        if p["nw"] > 16 or p["mp"][0] > 6:
            continue
        n += 1
        km = workloads.kmesh_of(p)
        op = po.Problem(p["nw"], p["nk"], km)
        o = po.wann_main(op, ctl, p["m0"], p["u0"])
        r, u, cen, _, hist = solve_alone(ctx, p, capi.control(**ctl))
        s_h = s_u = s_c = 0.0
        for seed in (1, 2):
            rng = np.random.default_rng(seed)
            o2 = po.wann_main(op, ctl, p["m0"] * (1.0 + 1e-15 * rng.standard_normal(p["m0"].shape)), p["u0"])
            n2 = min(len(o2["history"]), len(o["history"]))
            s_h = max(s_h, np.abs(o2["history"][:n2, 0] - o["history"][:n2, 0]).max())
            s_u = max(s_u, np.abs(o2["u"] - o["u"]).max())
            s_c = max(s_c, np.abs(o2["centres"] - o["centres"]).max())
        what = f"batch item {i} ({p['kind']}, num_wann {p['nw']}, {p['mp'][0]}^3)"
        print(f"[ragged oracle] {what}: sensitivity Omega {s_h:.1e} U {s_u:.1e}", flush=True)
        assert r.iterations == o["iterations"], what
        assert np.abs(hist[:, 0] - o["history"][:, 0]).max() <= tol(1e-8, s_h), what
        assert np.abs(u - o["u"]).max() <= tol(1e-7, s_u), what
        assert np.abs(cen - o["centres"]).max() <= tol(1e-6, s_c), what
    assert n >= 4


# serving requests: num_wann 74 and 71 (the 65-256 class, same lane), 24,
# 26, 18 (4^3 meshes) and 14 (6^3)
SERVE_IDS = [6, 13, 1, 10, 31, 54]


def test_serving_with_large_requests_equals_alone(ctx):
    from wannier90_amd import capi, workloads

    plist = [workloads.serve_params(i) for i in SERVE_IDS]
    probs = workloads.generate(plist, workers=1)
    print(f"[serve] requests num_wann {[p['nw'] for p in probs]}", flush=True)
    assert sum(p["nw"] > 64 for p in probs) == 2
    ctl = workloads.control(NUM_ITER, False)
    alone = [solve_alone(ctx, p, ctl) for p in probs]
    # all requests arrive within 50 ms, so they overlap on the device
    arrivals = np.cumsum(np.random.default_rng(3).exponential(0.01, size=len(probs)))
    arrivals -= arrivals[0]
    lat, res, peak, _, makespan = workloads.serve(probs, arrivals, ctl, NUM_ITER, cap_running=16, cap_large=2,
                                                  nlanes=2, pool_bytes=8 << 30)
    print(f"[serve] makespan {makespan:.3f} s, latency p50 {workloads.pct(lat, 0.5):.3f} s "
          f"p99 {workloads.pct(lat, 0.99):.3f} s, peak in flight {peak}", flush=True)
    assert all(x is not None and x > 0 for x in lat)
    assert sum(peak) >= 3
    for i, (r, cen, spr), a in zip(SERVE_IDS, res, alone):
        what = f"request {i}"
        assert r.status == capi.OK, what
        assert r.iterations == a[0].iterations and r.omega_total == a[0].omega_total, what
        assert np.array_equal(cen, a[2]) and np.array_equal(spr, a[3]), what
