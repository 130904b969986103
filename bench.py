"""Benchmark of the MI355X wannierisation path.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp64|mixed]
                    [--config 1|2|3] [--extras batch,serve,dis|none] [--no-cpu-baseline]
                    [--workload single|batch|serve|converge|mmn|dis]

single (default), --config 3 (default): the largest single-GPU
configuration, a seeded triclinic tight-binding problem (a, b, c ~ U[4,7]
Ang, angles ~ U[75,105] deg, num_wann = num_bands = 64, decay 2.5 Ang,
10x10x8 MP grid = 800 k, nntot 12), solved with num_iter = 500,
conv_tol = 1e-10, conv_window = 3, trial_step = 2, num_cg_steps = 5 by the
plain FP64 drop-in path, one problem per launch.  One step = one complete
minimisation (wann_main) from the Lowdin-projected initial state, with M
and U resident in HBM before the timed region.  --config 2 is the fcc
16-band 512-k problem, --config 1 the 4-band 64-k one.

Prints one JSON line: value = FP64 solves/s on one GPU, with
  roofline      the dominant kernel (the M rotation with the fused omega
                partial sums), its duration from HIP events on its own
                stream, against the FP64 MFMA or HBM peak;
  mixed         the same problem in mixed precision (FP32 MFMA bulk, FP64
                phases and polish): solves/s and Omega / centres against the
                FP64 solve;
  ragged_batch  (extras) the ragged batch of 256 problems: one problem per
                launch in FP64, batched FP64, batched mixed;
  serving       (extras) open-loop Poisson serving with continuous
                batching: p50/p99 time to converged;
  disentangle   (extras) the disentanglement of a 48-band fcc problem to
                24 WFs on 512 k, beside the reference's w90_disentangle
                (MPI over k-points) on a 20-iteration sample;
  cpu_baseline  the reference (its w90_library built from its sources,
                driven by oracle/ref_driver.F90) with MPI over k-points on
                the host's cores, on a bounded sample of the same workload.

batch, serve, converge (wannier90_amd/workloads.py): the ragged batch of 256
problems, the serving loop, and the batch run to convergence.  mmn: the
parse of a benchmark problem's .mmn file by the library's parallel reader
(host only).
"""
import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MLWF solves/s at ΔΩ≤1e-8 Å² vs FP64; p50/p99 latency; mixed vs plain GPU"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="single: skip the mixed-precision comparison")
    ap.add_argument("--cpu-sample-iters", type=int, default=0,
                    help="iterations of the CPU baseline sample (0: 100 for configs 1-2, 20 for config 3)")
    ap.add_argument("--cpu-ranks", type=int, default=0,
                    help="MPI ranks of the CPU baseline (0: the host's cores, at most 16)")
    ap.add_argument("--extras", default="batch,serve,dis",
                    help="single: comma list of extra workloads reported beside the line (batch, serve, dis) "
                         "or none")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device once (a kernel of known traffic that "
                         "calibrates FETCH_SIZE/WRITE_SIZE counters)")
    ap.add_argument("--config", type=int, choices=[1, 2, 3], default=3,
                    help="single: configuration (3: num_wann 64 triclinic 800 k, the largest; 2: fcc 16 "
                         "bands 512 k; 1: sc 4 bands 64 k)")
    ap.add_argument("--workload", choices=["single", "batch", "serve", "mmn", "converge", "dis"], default="single")
    ap.add_argument("--problems", type=int, default=256, help="batch: number of problems")
    ap.add_argument("--first", type=int, default=0, help="converge: first batch item")
    ap.add_argument("--no-plain", action="store_true", help="converge: skip the one-per-launch FP64 solves")
    ap.add_argument("--requests", type=int, default=48, help="serve: number of requests")
    ap.add_argument("--rate", type=float, default=0.0, help="serve: arrivals per second (0: from a probe)")
    ap.add_argument("--num-iter", type=int, default=500)
    ap.add_argument("--pool-gb", type=float, default=160.0, help="serve: device pool for admission")
    ap.add_argument("--max-running", type=int, default=16, help="serve: problems in flight at most")
    ap.add_argument("--large-running", type=int, default=2,
                    help="serve: problems in flight at most in the lane of num_wann > 32")
    ap.add_argument("--lanes", type=int, choices=[1, 2], default=2,
                    help="serve: 2 runs num_wann <= 32 and larger requests in separate batches")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_ranks(requested):
    """MPI ranks of the CPU baseline: the requested count, or the host's
    cores (at most 16, the CPU share of one GPU on the benchmark host)."""
    if requested > 0:
        return requested
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(prob, m0, u0, sample_iters, ranks):
    """The reference's wann_main (its w90_library compiled from its sources
    by oracle/Makefile.ref, driven by oracle/ref_driver.F90) on the same
    initial state with num_iter cut to `sample_iters`, with MPI over
    k-points on `ranks` ranks (MKL sequential, one core per rank).  The loop
    time comes from the CONV wall-clock column (iteration 0 is printed when
    wann_main starts).  Falls back to the serial reference, then to the C
    restatement (kind "port")."""
    from oracle import refrun

    ctl = dict(num_iter=sample_iters, num_cg_steps=5, conv_window=3, conv_tol=1e-10, trial_step=2.0,
               num_print_cycles=1)
    for p in ([ranks, 1] if ranks > 1 else [1]):
        if not refrun.available(p):
            continue
        try:
            r = refrun.run(prob.real_lattice, prob.mp, prob.kpts, prob.kmesh, m0, u0, ctl, ranks=p)
        except (RuntimeError, OSError, subprocess.SubprocessError) as e:
            log(f"[bench] cpu baseline with {p} ranks failed: {e}")
            continue
        loop = float(r["conv_wall"][-1] - r["conv_wall"][0])
        return dict(seconds=loop, iters=r["iterations"], kind="reference", cores=p,
                    sample=f"reference w90_library (its sources, amdflang, MKL sequential) driven by "
                           f"oracle/ref_driver.F90, {'MPI over k-points on %d ranks' % p if p > 1 else 'serial'}, "
                           f"num_iter={sample_iters} from the same initial state; loop wall time {loop:.2f} s "
                           f"from its CONV lines ({r['seconds']:.1f} s for w90_wannierise)")
    from oracle import pyoracle as po

    pr = po.Problem(prob.num_wann, prob.num_kpts, prob.kmesh)
    t0 = time.perf_counter()
    r = po.wann_main(pr, ctl, m0, u0)
    dt = time.perf_counter() - t0
    return dict(seconds=dt, iters=r["iterations"], kind="port", cores=1,
                sample=f"C restatement of wann_main (1 thread, MKL sequential), num_iter={sample_iters}")


def dis_summary(args):
    """The disentanglement workload (workloads.run_dis) as an extra key, with
    the reference timed on a 20-iteration sample on the CPU baseline's
    ranks (or none with --no-cpu-baseline)."""
    from types import SimpleNamespace

    from wannier90_amd import workloads

    a = SimpleNamespace(reps=3, dis_cpu_iters=0 if args.no_cpu_baseline else 20,
                        cpu_ranks_resolved=cpu_ranks(args.cpu_ranks))
    return workloads.run_dis(a)


def batch_summary(args):
    """The ragged batch of 256 problems (workloads.run_batch) as an extra key."""
    from types import SimpleNamespace

    from wannier90_amd import workloads

    a = SimpleNamespace(problems=256, num_iter=500)
    r = workloads.run_batch(a)
    c = dict(r["config"])
    c["value_solves_s"] = r["value"]
    return c


def serve_summary(args):
    """The serving loop (workloads.run_serve) as an extra key."""
    from types import SimpleNamespace

    from wannier90_amd import workloads

    a = SimpleNamespace(requests=args.requests, num_iter=500, rate=args.rate, precision="fp64",
                        max_running=args.max_running, large_running=args.large_running, lanes=args.lanes,
                        pool_gb=args.pool_gb)
    r = workloads.run_serve(a)
    c = dict(r["config"])
    c["value_solves_s"] = r["value"]
    return c


WORKLOADS = {
    1: "sc a=5.0 A, num_wann=4, 4x4x4 MP grid (64 k)",
    2: "fcc a=5.43 A, num_wann=16, 8x8x8 MP grid (512 k)",
    3: "triclinic a,b,c~U[4,7] A, num_wann=64, 10x10x8 MP grid (800 k)",
}


# file stem of each workload's kernel profiles under profiles/
PROFILE_STEM = {1: "sc4_64k", 2: "fcc16_512k", 3: "triclinic64_800k"}


def measured_traffic(config, precision):
    """HBM bytes per launch of the rotation kernel from the committed PMC
    summary of the same workload (profiles/*_pmc.json: FETCH_SIZE and
    WRITE_SIZE per dispatch, scaled by the same counters on a copy kernel
    of known size), or None."""
    f = os.path.join(ROOT, "profiles", f"rotate_{PROFILE_STEM[config]}_{precision}_pmc.json")
    if not os.path.exists(f):
        return None
    with open(f) as fh:
        return json.load(fh).get("traffic_bytes_per_launch")


def mixed_comparison(ctx, prob, m0, u0, sess_fp64, r64, args):
    """The same problem in mixed precision: solves/s, and Omega / centres
    against the FP64 solve of the same inputs (sess_fp64 holds its state)."""
    from wannier90_amd import capi

    _, _, cen64, _ = sess_fp64.fetch(want_m=False)
    ctl = capi.control(num_iter=500, conv_tol=1e-10, conv_window=3, trial_step=2.0, num_cg_steps=5,
                       precision=capi.PREC_MIXED)
    s = ctx.session(prob.kmesh, prob.num_wann, prob.num_kpts, ctl)
    s.load(m0, u0)
    for _ in range(args.warmup):
        s.solve()
    t0 = time.perf_counter()
    res = [s.solve() for _ in range(args.steps)]
    dt = time.perf_counter() - t0
    _, _, cen, _ = s.fetch(want_m=False)
    s.close()
    r = res[-1]
    d_om = abs(r.omega_total - r64.omega_total)
    d_c = float(np.abs(cen - cen64).max())
    log(f"[bench] mixed: {args.steps / dt:.2f} solves/s, bulk {r.bulk_iterations}/{r.iterations} iterations, "
        f"dOmega {d_om:.2e}, dcentres {d_c:.2e}")
    return dict(value=args.steps / dt, unit="solves/s", ms_per_step=1e3 * dt / args.steps,
                iterations=r.iterations, bulk_iterations=r.bulk_iterations, omega_total=r.omega_total,
                delta_omega_vs_fp64=d_om, delta_centres_vs_fp64=d_c,
                within_1e8=bool(d_om <= 1e-8 and d_c <= 1e-6))


def copy_probe(mib):
    """One device-to-device copy of `mib` MiB through the HIP runtime's copy
    kernel: FETCH_SIZE and WRITE_SIZE of that dispatch should each read
    mib * 1024 KiB, which gives the scale of the counters."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    n = C.c_size_t(mib << 20)
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), n) or hip.hipMalloc(C.byref(b), n):
        raise RuntimeError("hipMalloc failed in the copy probe")
    hip.hipMemset(a, 1, n)
    hip.hipMemcpy(b, a, n, 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"[bench] copy probe: {mib} MiB device to device")


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(batched and serving workloads shard as whole problems per GPU)", file=sys.stderr)
        sys.exit(2)
    if args.workload != "single":
        from wannier90_amd import workloads

        args.cpu_ranks_resolved = cpu_ranks(args.cpu_ranks)
        args.dis_cpu_iters = 0 if args.no_cpu_baseline else 20
        args.reps = 3
        out = workloads.run(args)
        print(json.dumps(out), flush=True)
        return
    from wannier90_amd import capi, synth

    if args.copy_probe:
        copy_probe(args.copy_probe)
    t = time.perf_counter()
    prob = synth.config(args.config)
    u0, m0 = synth.initial_state(prob)
    log(f"[bench] generated cfg{args.config} problem in {time.perf_counter() - t:.1f}s: nk={prob.num_kpts} "
        f"nn={prob.kmesh.nntot} nw={prob.num_wann}")
    if args.cpu_sample_iters <= 0:
        args.cpu_sample_iters = 20 if args.config == 3 else 100
    prec = capi.PREC_FP64 if args.precision == "fp64" else capi.PREC_MIXED
    ctl = capi.control(num_iter=500, conv_tol=1e-10, conv_window=3, trial_step=2.0,
                       num_cg_steps=5, precision=prec)
    ctx = capi.Context(0)
    sess = ctx.session(prob.kmesh, prob.num_wann, prob.num_kpts, ctl)
    sess.load(m0, u0)
    for i in range(args.warmup):
        r = sess.solve()
        log(f"[bench] warmup {i}: iterations={r.iterations} Omega={r.omega_total:.10f} "
            f"t={r.solve_seconds * 1e3:.2f} ms")
    # timed region: K complete solves, inputs resident in HBM
    t0 = time.perf_counter()
    results = []
    for i in range(args.steps):
        results.append(sess.solve())
    t1 = time.perf_counter()
    total = t1 - t0
    lat = sorted(r.solve_seconds for r in results)
    its = results[-1].iterations
    omega = results[-1].omega_total
    log(f"[bench] {args.steps} solves in {total:.3f}s; iterations={its} Omega={omega:.10f}")

    # the dominant kernel (M rotation with the fused omega partial sums):
    # HIP events around back-to-back launches on the final state
    launches = 200 if prob.num_wann <= 32 else 20
    avg_ms = sess.rotation_time(launches)
    nk, nn, nw = prob.num_kpts, prob.kmesh.nntot, prob.num_wann
    rot_bytes = (16.0 if prec == capi.PREC_FP64 else 8.0) * nk * nw * nw * (2 * nn + 1)  # M read + write, E once
    rot_flops = 16.0 * nk * nn * nw ** 3               # two complex GEMMs per (k,b)
    ai = rot_flops / rot_bytes
    peak_flops = FP64_PEAK_TFLOPS if prec == capi.PREC_FP64 else FP32_PEAK_TFLOPS
    ridge = peak_flops * 1e12 / (HBM_PEAK_GBS * 1e9)
    if ai < ridge:
        achieved = rot_bytes / (avg_ms * 1e-3) / 1e9
        roof = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=achieved / HBM_PEAK_GBS, traffic=None)
    else:
        achieved = rot_flops / (avg_ms * 1e-3) / 1e12
        roof = dict(bound="mfma", achieved=achieved, peak=peak_flops, unit="TFLOP/s",
                    frac=achieved / peak_flops, traffic=None)
    kname = ("k_rotate_w16" if nw <= 16 else "k_rot_fused + k_kpart" if nw <= 64
             else "k_rot_gemm stage 1 + stage 2 + k_kpart")
    roof.update(kernel=kname, avg_launch_ms=avg_ms, launches=launches,
                algorithmic_bytes=rot_bytes, algorithmic_flops=rot_flops)
    roof["traffic"] = measured_traffic(args.config, args.precision)
    log(f"[bench] k_rotate: {launches} launches, avg {avg_ms * 1e3:.1f} us, AI={ai:.2f} flop/B")
    # the other phases of an iteration, timed the same way (the omega
    # reductions and control are the remainder of the per-iteration time)
    phase_us = {"rotation": avg_ms * 1e3,
                "gradient": sess.phase_time("gradient", launches) * 1e3,
                "retraction": sess.phase_time("retraction", launches) * 1e3,
                "reduction": sess.phase_time("reduction", launches) * 1e3,
                "iteration": 1e6 * total / args.steps / max(1, its)}
    log("[bench] phases (us): " + ", ".join(f"{k} {v:.1f}" for k, v in phase_us.items()))

    out = {
        "metric": METRIC,
        "value": args.steps / total,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == capi.PREC_FP64 else "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": WORKLOADS[args.config] + f", nntot={nn}, num_iter=500, conv_tol=1e-10, "
                        "conv_window=3, one problem per launch",
            "precision": args.precision,
            "iterations": its,
            "phase_us": phase_us,
            "omega_total": omega,
            "p50_ms": 1e3 * lat[len(lat) // 2],
            "p99_ms": 1e3 * lat[min(len(lat) - 1, int(round(0.99 * (len(lat) - 1))))],
        },
        "roofline": roof,
    }
    if prec == capi.PREC_FP64 and not args.no_mixed:
        out["mixed"] = mixed_comparison(ctx, prob, m0, u0, sess, results[-1], args)
    sess.close()
    ctx.close()
    extras = [] if args.extras in ("", "none") else args.extras.split(",")
    if "batch" in extras:
        out["ragged_batch"] = batch_summary(args)
    if "serve" in extras:
        out["serving"] = serve_summary(args)
    if "dis" in extras:
        out["disentangle"] = dis_summary(args)
    if not args.no_cpu_baseline:
        cb = cpu_baseline(prob, m0, u0, args.cpu_sample_iters, cpu_ranks(args.cpu_ranks))
        per_solve = cb["seconds"] * its / max(cb["iters"], 1)
        out["cpu_baseline"] = dict(value=1.0 / per_solve, unit="solves/s", cores=cb["cores"], kind=cb["kind"],
                                   sample=cb["sample"] + f"; scaled to the GPU solve's {its} iterations")
        log(f"[bench] cpu baseline: {cb['seconds']:.1f}s for {cb['iters']} iterations on {cb['cores']} cores")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
