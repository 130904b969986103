"""Sweep of the rotation autotuner over (num_wann, nntot).

The autotuner lives in the engine (w90gpu.hip: rot_choice, tune_member): with
W90GPU_TUNE=1 the first solve of each (num_wann, nntot) times every rotation
kernel of its num_wann class on the problem's own resident data and keeps a
candidate only if it is at least 3% faster than the class default; with
W90GPU_TUNE_FILE set, each choice is appended to that file as

    num_wann nntot kernel candidate=microseconds ... items=N

This module drives that tuner over a grid of keys (one short solve per key on
a Gamma-centred 4x4x4 mesh of a lattice with the wanted nntot, M(k,b) = I and
U = I: the rotation's cost does not depend on the values) and turns the
resulting file into the compiled-in table w90gpu_tune_table.h, whose rows
hold the keys where the tuner kept a non-default kernel, merged into
num_wann ranges; each measured nntot stands for the range of nntot nearest
to it.

    python -m wannier90_amd.tune --out FILE [--nw 17:128] [--nn 6,8,12]
    python -m wannier90_amd.tune --emit FILE [--header PATH]
"""
import argparse
import os
import sys

import numpy as np

# lattices giving each nntot on a cubic mesh: sc 6, fcc 8, bcc 12
LATTICE_OF_NN = {6: "sc", 8: "fcc", 12: "bcc"}
DEFAULTS = {1: "fused32", 2: "fused64", 3: "gemm64"}
KERNELS = ["w16", "fused32", "fused64", "direct64", "gemm64", "gemm32", "fused48"]
HEADER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "w90gpu_tune_table.h")


def wann_class(nw):
    return 0 if nw <= 16 else 1 if nw <= 32 else 2 if nw <= 64 else 3


def parse_range(text):
    lo, _, hi = text.partition(":")
    return range(int(lo), int(hi or lo) + 1)


def sweep(out, nws, nns, mp=(4, 4, 4)):
    """One solve per key with the online tuner on; choices go to `out`."""
    if os.environ.get("W90GPU_TUNE") != "1" or os.environ.get("W90GPU_TUNE_FILE") != out:
        raise RuntimeError("run through main(): it sets the environment the tuner reads")
    from . import capi, synth

    ctx = capi.Context()
    kpts = synth.mp_kpoints(mp)
    nk = kpts.shape[0]
    for nn in nns:
        km = capi.kmesh_get(synth.lattice(LATTICE_OF_NN[nn]), kpts)
        if km.nntot != nn:
            raise RuntimeError(f"lattice {LATTICE_OF_NN[nn]} gives nntot={km.nntot}, wanted {nn}")
        for nw in nws:
            if wann_class(nw) == 0:
                continue
            eye = np.eye(nw, dtype=np.complex128)
            m = np.asfortranarray(np.broadcast_to(eye[:, :, None, None], (nw, nw, nn, nk)))
            u = np.asfortranarray(np.broadcast_to(eye[:, :, None], (nw, nw, nk)))
            s = ctx.session(km, nw, nk, capi.control(num_iter=1))
            s.load(m, u)
            s.solve()
            s.close()
            print(f"tuned nw={nw} nn={nn}", file=sys.stderr, flush=True)
    ctx.close()


def read_choices(path):
    """{(nw, nn): (kernel, {candidate: us})} from a tuning file."""
    out = {}
    with open(path) as f:
        for line in f:
            parts = line.split()
            if not parts or parts[0].startswith("#") or len(parts) < 3:
                continue
            nw, nn, k = int(parts[0]), int(parts[1]), parts[2]
            times = {}
            for p in parts[3:]:
                name, _, v = p.partition("=")
                if name in KERNELS:
                    times[name] = float(v)
            out[(nw, nn)] = (k, times)
    return out


def nn_ranges(nns, nnmax=12):
    """Each measured nntot stands for the nntot nearest to it."""
    nns = sorted(nns)
    rng = {}
    for i, n in enumerate(nns):
        lo = 1 if i == 0 else (nns[i - 1] + n) // 2 + 1
        hi = nnmax if i == len(nns) - 1 else (n + nns[i + 1]) // 2
        rng[n] = (lo, hi)
    return rng


def table_rows(choices):
    """Rows (nw_lo, nw_hi, nn_lo, nn_hi, kernel) for the non-default choices:
    runs of consecutive measured num_wann of one class with one choice."""
    nns = sorted({nn for _, nn in choices})
    nr = nn_ranges(nns)
    rows = []
    for nn in nns:
        keys = sorted(nw for (nw, n) in choices if n == nn)
        run = None
        for nw in keys:
            k = choices[(nw, nn)][0]
            if k == DEFAULTS[wann_class(nw)]:
                k = None
            if run and k == run[2] and nw == run[1] + 1 and wann_class(nw) == wann_class(run[0]):
                run[1] = nw
                continue
            if run and run[2]:
                rows.append((run[0], run[1], *nr[nn], run[2]))
            run = [nw, nw, k]
        if run and run[2]:
            rows.append((run[0], run[1], *nr[nn], run[2]))
    return rows


def emit(choices, header, source):
    rows = table_rows(choices)
    lines = [
        "// Rotation kernel by (num_wann, nntot) on MI355X: the choices of the",
        "// autotuner in w90gpu.hip (tune_member), measured over a sweep of synthetic",
        "// problems by wannier90_amd/tune.py, which writes this file.  A key that no",
        "// row covers keeps its class default.",
        f"// Source: {os.path.basename(source)}, {len(choices)} keys, {len(rows)} rows.",
        "#pragma once",
        "",
        "struct TuneRow {",
        "  int nw_lo, nw_hi, nn_lo, nn_hi, rk;",
        "};",
        "",
        "constexpr TuneRow kTuneTable[] = {",
    ]
    if not rows:
        lines.append("    {0, -1, 0, -1, 0},  // no measured rows")
    for r in rows:
        lines.append(f"    {{{r[0]}, {r[1]}, {r[2]}, {r[3]}, RK_{r[4].upper()}}},")
    lines += ["};", ""]
    with open(header, "w") as f:
        f.write("\n".join(lines))
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", help="run the sweep on the GPU, appending choices to this file")
    ap.add_argument("--nw", default="17:128")
    ap.add_argument("--nn", default="6,8,12")
    ap.add_argument("--emit", help="write the table header from this tuning file")
    ap.add_argument("--header", default=HEADER)
    a = ap.parse_args(argv)
    if a.out:
        os.environ["W90GPU_TUNE"] = "1"
        os.environ["W90GPU_TUNE_FILE"] = a.out
        sweep(a.out, list(parse_range(a.nw)), [int(x) for x in a.nn.split(",")])
    if a.emit:
        rows = emit(read_choices(a.emit), a.header, a.emit)
        for r in rows:
            print(*r)


if __name__ == "__main__":
    main()
