"""Ragged-batch and serving workloads over the batch C ABI (w90gpu_batch_*).

batch: a ragged batch of independent problems (lattice uniform over sc, fcc,
bcc, hexagonal with c/a in [1.2, 1.8]; num_wann uniform in [4, 48]; k-mesh
uniform over 4^3, 6^3, 8^3; each from its own sub-seed), solved
  - one problem per launch: a device-resident session per problem, FP64 (the
    plain drop-in path), problems one after another;
  - as one batch in FP64 (the results must equal the plain ones bit for bit);
  - as one batch in mixed precision, whose Omega and centres are compared
    with the plain FP64 result of each problem.

serve: open-loop Poisson arrivals of problems drawn from the same
distribution extended to num_wann <= 96 and k-meshes up to 12^3; requests are
admitted into a running batch as they arrive and as device memory allows
(continuous batching), their M/U/centres stay on the device until fetched;
p50/p99 time-to-converged from arrival and the throughput are reported next
to a serial plain FP64 server (one request at a time, service times measured
on the same device).
"""
import json
import multiprocessing as mp
import os
import sys
import threading
import time

import numpy as np

from . import synth

METRIC = "MLWF solves/s at ΔΩ≤1e-8 Å² vs FP64; p50/p99 latency; mixed vs plain GPU"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ------------------------------------------------------------ problem sets

def batch_params(i):
    """Item i of the ragged batch (256 items)."""
    return synth.batch_params(i)


def serve_params(i):
    """Request i of the serving workload: the batch distribution extended to
    num_wann <= 96 and k-meshes up to 12^3."""
    return synth.batch_params(i, seed=synth.BASE_SEED + 5, nw_range=(4, 96), meshes=(4, 6, 8, 10, 12))


def _generate(params):
    p = dict(params)
    seed = p.pop("seed")
    prob = synth.make_problem(seed, **p)
    u0, m0 = synth.initial_state(prob)
    km = prob.kmesh
    return dict(nw=prob.num_wann, nk=prob.num_kpts, nn=km.nntot, kind=p["kind"], mp=p["mp"],
                real_lattice=prob.real_lattice, nnlist=km.nnlist, nncell=km.nncell, bk=km.bk, wb=km.wb,
                wbtot=km.wbtot, m0=m0, u0=u0, tau_cart=prob.tau_cart)


# Worker processes of the problem generator.  They load only the host-side
# library (libw90io.so), never the HIP runtime; the cap keeps the workers
# plus the parent at 15 processes, so that even a tool that starts the GPU
# runtime in every process (a profiler's preload) stays within the 16
# processes one GPU admits.
MAX_WORKERS = 14


def generate(plist, workers=None):
    """Problems for a list of parameter sets, generated in worker processes
    (spawned, so no GPU state is shared; at most MAX_WORKERS)."""
    workers = min(workers or (os.cpu_count() or 1), MAX_WORKERS)
    t = time.perf_counter()
    if workers <= 1 or len(plist) <= 1:
        out = [_generate(p) for p in plist]
    else:
        # one BLAS thread per worker; largest problems first
        saved = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "MKL_NUM_THREADS", "OPENBLAS_NUM_THREADS")}
        for k in saved:
            os.environ[k] = "1"
        order = sorted(range(len(plist)), key=lambda i: -plist[i]["nw"] ** 2 * plist[i]["mp"][0] ** 3)
        try:
            with mp.get_context("spawn").Pool(workers) as pool:
                res = pool.map(_generate, [plist[i] for i in order], chunksize=1)
        finally:
            for k, v in saved.items():
                if v is None:
                    os.environ.pop(k, None)
                else:
                    os.environ[k] = v
        out = [None] * len(plist)
        for j, i in enumerate(order):
            out[i] = res[j]
    log(f"[workload] generated {len(plist)} problems in {time.perf_counter() - t:.1f}s ({workers} workers)")
    return out


def kmesh_of(p):
    from . import capi

    return capi.KMesh(p["nn"], np.asfortranarray(p["nnlist"]), np.asfortranarray(p["nncell"]),
                      np.asfortranarray(p["bk"]), np.ascontiguousarray(p["wb"]), p["wbtot"])


def control(num_iter, mixed):
    from . import capi

    return capi.control(num_iter=num_iter, conv_window=3, conv_tol=1e-10, trial_step=2.0, num_cg_steps=5,
                        precision=capi.PREC_MIXED if mixed else capi.PREC_FP64)


def solve_alone(ctx, p, ctl):
    """One problem per launch (device-resident session): (result, centres)."""
    s = ctx.session(kmesh_of(p), p["nw"], p["nk"], ctl)
    s.load(p["m0"], p["u0"])
    r = s.solve()
    _, _, cen, spr = s.fetch(want_m=False)
    s.close()
    return r, cen, spr


def same_set(cen_a, spr_a, cen_b, spr_b, lattice, tol_c=1e-6, tol_s=1e-6):
    """Centres/spreads equal up to a relabelling of the Wannier functions and
    lattice translations of their centres."""
    lat = np.asarray(lattice)
    inv = np.linalg.inv(lat.T)
    used = set()
    for n in range(cen_a.shape[1]):
        hit = None
        for m in range(cen_b.shape[1]):
            if m in used or abs(spr_a[n] - spr_b[m]) > tol_s:
                continue
            f = inv @ (cen_a[:, n] - cen_b[:, m])
            if np.abs(lat.T @ (f - np.round(f))).max() <= tol_c:
                hit = m
                break
        if hit is None:
            return False
        used.add(hit)
    return True


def pct(v, q):
    v = sorted(v)
    if not v:
        return None
    return v[min(len(v) - 1, int(round(q * (len(v) - 1))))]


# ------------------------------------------------------------ batch

def run_batch(args):
    from . import capi

    probs = generate([batch_params(i) for i in range(args.problems)])
    n = len(probs)
    ctx = capi.Context(0)
    warm = probs[0]
    solve_alone(ctx, warm, control(20, False))

    # one problem per launch, FP64
    plain = []
    t0 = time.perf_counter()
    for i, p in enumerate(probs):
        r, cen, spr = solve_alone(ctx, p, control(args.num_iter, False))
        plain.append((r, cen, spr))
    t_plain_wall = time.perf_counter() - t0
    t_plain = sum(r.solve_seconds for r, _, _ in plain)
    log(f"[batch] plain: {t_plain:.2f}s of solves ({t_plain_wall:.2f}s wall with set-up)")

    def batched(mixed):
        ctl = control(args.num_iter, mixed)
        pool = sum(capi.problem_bytes(p["nw"], p["nk"], p["nn"], args.num_iter) for p in probs)
        b = ctx.batch(pool)
        ids = [b.add(kmesh_of(p), p["m0"], p["u0"], ctl) for p in probs]
        assert None not in ids
        t = time.perf_counter()
        b.run()
        dt = time.perf_counter() - t
        out = []
        for i in ids:
            r = b.result(i)
            _, _, cen, spr = b.fetch(i)
            out.append((r, cen, spr))
        b.close()
        return dt, out

    t_b64, res64 = batched(False)
    exact = sum(1 for (a, ca, _), (b, cb, _) in zip(plain, res64)
                if a.omega_total == b.omega_total and a.iterations == b.iterations and np.array_equal(ca, cb))
    ok64 = sum(1 for (a, ca, _), (b, cb, _) in zip(plain, res64)
               if abs(a.omega_total - b.omega_total) <= 1e-8 and float(np.abs(ca - cb).max()) <= 1e-6)
    log(f"[batch] batched FP64: {t_b64:.2f}s, {exact}/{n} bit-identical to one-per-launch")
    t_mix, resm = batched(True)
    d_om = [abs(a.omega_total - b.omega_total) for (a, _, _), (b, _, _) in zip(plain, resm)]
    d_c = [float(np.abs(ca - cb).max()) for (_, ca, _), (_, cb, _) in zip(plain, resm)]
    ok = [do <= 1e-8 and dc <= 1e-6 for do, dc in zip(d_om, d_c)]
    ok_set = [do <= 1e-8 and same_set(cb, sb, ca, sa, p["real_lattice"])
              for do, (_, ca, sa), (_, cb, sb), p in zip(d_om, plain, resm, probs)]
    conv = sum(1 for r, _, _ in plain if r.converged)
    log(f"[batch] batched mixed: {t_mix:.2f}s; within 1e-8/1e-6: {sum(ok)}/{n}, "
        f"same WF set: {sum(ok_set)}/{n}; plain converged {conv}/{n}")
    its = [r.iterations for r, _, _ in plain]
    # the value is the faster batched form counting only solves within
    # 1e-8 Ang^2 (and centres 1e-6 Ang) of the one-per-launch FP64 solve
    best = max((ok64 / t_b64, "batched fp64"), (sum(ok) / t_mix, "batched mixed"))
    return {
        "metric": METRIC,
        "value": best[0],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": 1e3 * t_mix,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if best[1] == "batched fp64" else "f32+f64",
        "data": "synthetic",
        "config": {
            "value_from": best[1],
            "workload": f"ragged batch of {n} problems: sc/fcc/bcc/hex, num_wann U[4,48], k-mesh "
                        f"4^3/6^3/8^3, num_iter={args.num_iter}, conv_window=3, conv_tol=1e-10",
            "plain_fp64_one_per_launch_s": t_plain,
            "batched_fp64_s": t_b64,
            "batched_mixed_s": t_mix,
            "speedup_batched_mixed_vs_plain": t_plain / t_mix,
            "speedup_batched_fp64_vs_plain": t_plain / t_b64,
            "batched_fp64_bit_identical": exact,
            "batched_fp64_within_1e8": ok64,
            "mixed_within_1e8_omega_1e6_centres": sum(ok),
            "mixed_same_wannier_set": sum(ok_set),
            "max_delta_omega": max(d_om),
            "median_delta_omega": pct(d_om, 0.5),
            "plain_converged": conv,
            "plain_iterations_median": pct(its, 0.5),
            "mixed_bulk_iterations_median": pct([r.bulk_iterations for r, _, _ in resm], 0.5),
            "problems": n,
        },
    }


# ------------------------------------------------------------ convergence

def at_orbitals(cen, tau, lattice, tol_c=1e-6):
    """Every centre on an orbital position (Cartesian) up to lattice
    vectors, one centre per orbital: the closed-form global minimum of the
    tight-binding problems (Omega = 0)."""
    lat = np.asarray(lattice)
    inv = np.linalg.inv(lat)
    hit = []
    for x in cen.T:
        d = (x - tau) @ inv
        d -= np.round(d)
        dist = np.linalg.norm(d @ lat, axis=1)
        if dist.min() >= tol_c:
            return False
        hit.append(int(np.argmin(dist)))
    return sorted(hit) == list(range(tau.shape[0]))


def run_converge(args):
    """The ragged batch run to convergence: FP64 and mixed precision, both
    batched, with num_iter high enough for the conv_window test to end the
    solves, and FP64 one problem per launch.  Per problem: iterations and
    time to converge, final Omega, and whether the result is the
    closed-form global minimum (Omega <= 1e-8 Ang^2, centres on the orbital
    positions) -- the same minimum for both precisions is what makes Omega
    agree within 1e-8 and the centres within 1e-6 (up to a relabelling of
    the functions)."""
    from . import capi

    ids = list(range(args.first, args.first + args.problems))
    probs = generate([batch_params(i) for i in ids])
    n = len(probs)
    ctx = capi.Context(0)
    solve_alone(ctx, probs[0], control(20, False))

    def batched(mixed):
        ctl = control(args.num_iter, mixed)
        pool = sum(capi.problem_bytes(p["nw"], p["nk"], p["nn"], args.num_iter) for p in probs)
        b = ctx.batch(pool)
        hs = [b.add(kmesh_of(p), p["m0"], p["u0"], ctl) for p in probs]
        t = time.perf_counter()
        b.run()
        dt = time.perf_counter() - t
        out = []
        for h in hs:
            r = b.result(h)
            _, _, cen, spr = b.fetch(h)
            out.append((r, cen, spr))
        b.close()
        return dt, out

    t64, r64 = batched(False)
    log(f"[converge] batched FP64: {t64:.2f}s")
    tmx, rmx = batched(True)
    log(f"[converge] batched mixed: {tmx:.2f}s")
    plain_t = 0.0
    if not args.no_plain:
        for p in probs:
            r, _, _ = solve_alone(ctx, p, control(args.num_iter, False))
            plain_t += r.solve_seconds
        log(f"[converge] one per launch FP64: {plain_t:.2f}s of solves")
    rows = []
    for i, p, (a, ca, sa), (m, cm, sm) in zip(ids, probs, r64, rmx):
        rows.append(dict(
            item=i, kind=p["kind"], nw=p["nw"], mesh=p["mp"][0],
            fp64_iterations=a.iterations, fp64_converged=a.converged, fp64_omega=a.omega_total,
            fp64_seconds=a.solve_seconds, fp64_global_min=bool(a.omega_total <= 1e-8 and
                                                                 at_orbitals(ca, p["tau_cart"], p["real_lattice"])),
            mixed_iterations=m.iterations, mixed_bulk=m.bulk_iterations, mixed_converged=m.converged,
            mixed_omega=m.omega_total, mixed_seconds=m.solve_seconds,
            mixed_global_min=bool(m.omega_total <= 1e-8 and at_orbitals(cm, p["tau_cart"], p["real_lattice"])),
            delta_omega=abs(a.omega_total - m.omega_total),
            same_set=bool(abs(a.omega_total - m.omega_total) <= 1e-8 and same_set(cm, sm, ca, sa, p["real_lattice"]))))
    its64 = [r["fp64_iterations"] for r in rows]
    itsmx = [r["mixed_iterations"] for r in rows]
    summary = dict(
        problems=n, num_iter=args.num_iter,
        fp64_converged=sum(r["fp64_converged"] for r in rows),
        mixed_converged=sum(r["mixed_converged"] for r in rows),
        fp64_global_min=sum(r["fp64_global_min"] for r in rows),
        mixed_global_min=sum(r["mixed_global_min"] for r in rows),
        within_1e8_same_set=sum(r["same_set"] for r in rows),
        fp64_iterations_p50=pct(its64, 0.5), fp64_iterations_max=max(its64),
        mixed_iterations_p50=pct(itsmx, 0.5), mixed_iterations_max=max(itsmx),
        batched_fp64_s=t64, batched_mixed_s=tmx, plain_fp64_one_per_launch_s=plain_t or None,
    )
    log("[converge] " + json.dumps(summary))
    return {"metric": METRIC, "value": summary["within_1e8_same_set"] / tmx, "unit": "solves/s", "n_gpus": 1,
            "steps": 1, "warmup": 1, "ms_per_step": 1e3 * tmx, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
            "config": {"workload": f"ragged batch items {ids[0]}..{ids[-1]} run to convergence "
                                   f"(num_iter={args.num_iter}, conv_window=3, conv_tol=1e-10)",
                       "summary": summary, "per_problem": rows}}


# ------------------------------------------------------------ serving

def serve(probs, arrivals, ctl, num_iter, cap_running=16, cap_large=2, nlanes=2, pool_bytes=0):
    """Continuous batching of requests arriving at `arrivals` (seconds from
    the start): each lane is one batch (own context, stream and share of the
    device pool) stepped from its own host thread; a request is admitted
    when it has arrived, its lane has fewer than its cap of problems in
    flight and its bytes fit the lane's pool; finished problems are fetched
    and removed at once.  With two lanes the requests of num_wann <= 32 and
    the larger ones run apart, so a small request does not advance in
    lockstep with the large problems (which fill the device on their own,
    so their lane admits only cap_large).  The library calls release the
    interpreter lock.

    Returns (latency from arrival to result per request, (result, centres,
    spreads) per request, peak problems in flight per lane, pool bytes per
    lane, makespan in seconds)."""
    from . import capi

    n = len(probs)
    need = [capi.problem_bytes(p["nw"], p["nk"], p["nn"], num_iter) for p in probs]
    pool_total = max(max(need), int(pool_bytes))
    lane_of = [0 if (nlanes == 1 or p["nw"] <= 32) else 1 for p in probs]
    lanes = [[i for i in range(n) if lane_of[i] == ln] for ln in range(nlanes)]
    # device pool per lane in proportion to the bytes its requests need
    lane_bytes = [sum(need[i] for i in ids) for ids in lanes]
    pools = [max([need[i] for i in ids] + [1 << 20], default=1 << 20) for ids in lanes]
    pools = [max(pl, int(pool_total * lb / max(1, sum(lane_bytes)))) for pl, lb in zip(pools, lane_bytes)]
    lat, res = [None] * n, [None] * n
    peak = [0] * nlanes
    errors = []

    def lane(ln, start):
        ids = lanes[ln]
        if not ids:
            return
        lctx = capi.Context(0)
        b = lctx.batch(pools[ln])
        queue, nxt, done = [], 0, 0
        owner = {}
        try:
            while done < len(ids):
                t = time.perf_counter() - start
                while nxt < len(ids) and arrivals[ids[nxt]] <= t:
                    queue.append(ids[nxt])
                    nxt += 1
                while queue and len(owner) < (cap_running if ln == 0 else cap_large):
                    i = queue[0]
                    p = probs[i]
                    h = b.add(kmesh_of(p), p["m0"], p["u0"], ctl)
                    if h is None:
                        break
                    owner[h] = i
                    queue.pop(0)
                info = b.info()
                peak[ln] = max(peak[ln], info["running"] + info["queued"])
                if info["running"] + info["queued"] == 0:
                    if nxt < len(ids):
                        time.sleep(max(0.0, arrivals[ids[nxt]] - (time.perf_counter() - start)))
                    continue
                # a few blocks per step while nobody is waiting to be admitted
                due = nxt < len(ids) and arrivals[ids[nxt]] <= time.perf_counter() - start
                finished = b.step(1 if queue or due else 2)
                t = time.perf_counter() - start
                for h in finished:
                    i = owner.pop(h)
                    lat[i] = t - arrivals[i]
                    r = b.result(h)
                    _, _, cen, spr = b.fetch(h)
                    res[i] = (r, cen, spr)
                    b.remove(h)
                    done += 1
        except Exception as ex:  # reported after the lanes have joined
            errors.append(ex)
        finally:
            b.close()
            lctx.close()

    start = time.perf_counter()
    threads = [threading.Thread(target=lane, args=(ln, start)) for ln in range(nlanes)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    if errors:
        raise errors[0]
    makespan = time.perf_counter() - start
    return lat, res, peak, pools, makespan


def run_serve(args):
    from . import capi

    probs = generate([serve_params(i) for i in range(args.requests)])
    n = len(probs)
    ctx = capi.Context(0)
    solve_alone(ctx, probs[0], control(20, False))
    # serial plain FP64 server: service time of each request alone
    service, plain = [], []
    for p in probs:
        r, cen, spr = solve_alone(ctx, p, control(args.num_iter, False))
        service.append(r.solve_seconds)
        plain.append((r, cen, spr))
    cap = n / sum(service)
    rate = args.rate if args.rate > 0 else 0.7 * cap
    rng = np.random.default_rng(synth.BASE_SEED + 6)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n))
    arrivals -= arrivals[0]
    # serial FCFS queue with the measured service times
    fin, lat_serial = 0.0, []
    for a, s in zip(arrivals, service):
        fin = max(fin, a) + s
        lat_serial.append(fin - a)
    log(f"[serve] {n} requests, serial capacity {cap:.2f}/s, arrival rate {rate:.2f}/s")

    # continuous batching, admission by device bytes and a cap on the
    # problems in flight; FP64 unless --precision mixed (FP64 results are
    # bit-identical to the same request solved alone)
    mixed = getattr(args, "precision", "fp64") == "mixed"
    ctl = control(args.num_iter, mixed)
    cap_running = max(1, int(getattr(args, "max_running", 16)))
    cap_large = max(1, int(getattr(args, "large_running", 2)))
    nlanes = 2 if int(getattr(args, "lanes", 2)) >= 2 else 1
    lat, res, peak, pools, makespan = serve(probs, arrivals, ctl, args.num_iter, cap_running=cap_running,
                                            cap_large=cap_large, nlanes=nlanes, pool_bytes=int(args.pool_gb * (1 << 30)))
    lane_of = [0 if (nlanes == 1 or p["nw"] <= 32) else 1 for p in probs]
    lanes = [[i for i in range(n) if lane_of[i] == ln] for ln in range(nlanes)]
    max_running = sum(peak)
    pool = sum(pools)
    d_om = [abs(a[0].omega_total - r[0].omega_total) for a, r in zip(plain, res)]
    ok = [do <= 1e-8 and float(np.abs(a[1] - r[1]).max()) <= 1e-6 for do, a, r in zip(d_om, plain, res)]
    ok_set = [do <= 1e-8 and same_set(r[1], r[2], a[1], a[2], p["real_lattice"])
              for do, a, r, p in zip(d_om, plain, res, probs)]
    thr = n / makespan
    log(f"[serve] continuous batching: p50 {pct(lat, 0.5):.3f}s p99 {pct(lat, 0.99):.3f}s, {thr:.2f} solves/s; "
        f"serial plain p50 {pct(lat_serial, 0.5):.3f}s p99 {pct(lat_serial, 0.99):.3f}s; "
        f"within 1e-8: {sum(ok)}/{n}, same set {sum(ok_set)}/{n}")
    return {
        "metric": METRIC,
        "value": sum(ok) / makespan,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": n,
        "warmup": 1,
        "ms_per_step": 1e3 * makespan / n,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if mixed else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"open-loop Poisson serving, {n} requests at {rate:.3f}/s: sc/fcc/bcc/hex, num_wann "
                        f"U[4,96], k-mesh 4^3..12^3, num_iter={args.num_iter}, conv_window=3, conv_tol=1e-10, "
                        f"{'mixed precision' if mixed else 'FP64'}, continuous batching of up to {cap_running} "
                        f"problems ({cap_large} in the lane of num_wann > 32), {nlanes} lane(s), admission by device "
                        f"bytes (pool {pool / (1 << 30):.0f} GiB)",
            "p50_s": pct(lat, 0.5),
            "p99_s": pct(lat, 0.99),
            "lanes": nlanes,
            "lane_p50_s": [pct([lat[i] for i in ids], 0.5) if ids else None for ids in lanes],
            "lane_p99_s": [pct([lat[i] for i in ids], 0.99) if ids else None for ids in lanes],
            "lane_requests": [len(ids) for ids in lanes],
            "throughput_solves_s": thr,
            "serial_plain_p50_s": pct(lat_serial, 0.5),
            "serial_plain_p99_s": pct(lat_serial, 0.99),
            "serial_plain_capacity_solves_s": cap,
            "arrival_rate": rate,
            "max_in_flight": max_running,
            "within_1e8_omega_1e6_centres_vs_alone": sum(ok),
            "same_wannier_set_vs_alone": sum(ok_set),
            "max_delta_omega": max(d_om),
            "requests": n,
        },
    }


def write_mmn_f18(path, nb, kmesh, seed=0):
    """A .mmn with uniform random entries in pw2wannier90's record format
    (2f18.12), blocks in the k-mesh's (k, b) order."""
    nk, nn = kmesh.nnlist.shape
    rng = np.random.default_rng(seed)
    with open(path, "w") as f:
        f.write(" Created by wannier90_amd.workloads\n")
        f.write(f"{nb:12d}{nk:12d}{nn:12d}\n")
        for k in range(nk):
            for b in range(nn):
                g = kmesh.nncell[:, k, b]
                f.write(f"{k + 1:5d}{kmesh.nnlist[k, b]:5d}{g[0]:5d}{g[1]:5d}{g[2]:5d}\n")
                v = rng.uniform(-1.0, 1.0, (nb * nb, 2))
                f.write("".join("%18.12f%18.12f\n" % (x, y) for x, y in v))


def run_mmn(args):
    """Parse of the benchmark problem's .mmn (w90gpu_read_mmn) with all
    host threads and with one; the two results must agree bit for bit, and
    sampled entries must equal float() of their text."""
    import tempfile

    from . import capi

    prob = synth.config(args.config)
    mesh = prob.kmesh
    nb = prob.num_wann
    nk, nn = mesh.nnlist.shape
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "bench.mmn")
        t = time.perf_counter()
        write_mmn_f18(path, nb, mesh)
        log(f"[mmn] wrote {os.path.getsize(path) / 1e6:.1f} MB in {time.perf_counter() - t:.1f}s")
        size = os.path.getsize(path)

        def timed(nt, reps):
            best, m = 1e30, None
            for _ in range(reps):
                t = time.perf_counter()
                m = capi.read_mmn(path, nb, mesh, nthreads=nt)
                best = min(best, time.perf_counter() - t)
            return best, m

        capi.read_mmn(path, nb, mesh)  # page cache warm
        t_all, m_all = timed(0, max(args.steps, 1))
        t_one, m_one = timed(1, max(min(args.steps, 3), 1))
        same = np.array_equal(m_all.ravel(order="F").view(np.uint64), m_one.ravel(order="F").view(np.uint64))
        # sampled records against float() of their text
        rng = np.random.default_rng(1)
        lines = open(path).read().split("\n")
        ok = 0
        picks = rng.integers(0, nk * nn * nb * nb, 2000)
        for q in picks:
            blk, e = divmod(int(q), nb * nb)
            k, b = divmod(blk, nn)
            re, im = (float(x) for x in lines[2 + blk * (nb * nb + 1) + 1 + e].split())
            z = m_all[e % nb, e // nb, b, k]
            ok += int(z.real == re and z.imag == im)
    nt = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    return {
        "metric": ".mmn parse throughput (w90gpu_read_mmn)",
        "value": size / t_all / 1e6,
        "unit": "MB/s",
        "n_gpus": 1,
        "steps": max(args.steps, 1),
        "warmup": 1,
        "ms_per_step": t_all * 1e3,
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f".mmn of num_bands={nb}, num_kpts={nk}, nntot={nn} (2f18.12 records, "
                        f"{nk * nn * nb * nb} value pairs, {size / 1e6:.1f} MB), host threads {nt}",
            "seconds_all_threads": t_all,
            "seconds_one_thread": t_one,
            "one_thread_mb_s": size / t_one / 1e6,
            "bit_identical_1_vs_all_threads": bool(same),
            "sampled_equal_float_of_text": f"{ok}/{len(picks)}",
        },
    }


def run(args):
    if args.workload == "batch":
        return run_batch(args)
    if args.workload == "converge":
        return run_converge(args)
    if args.workload == "mmn":
        return run_mmn(args)
    return run_serve(args)
