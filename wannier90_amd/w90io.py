"""wannier90 file formats used on the wannierise path.

Writers: .win, .mmn, .amn, .eig (inputs), .chk (unformatted checkpoint),
_centres.xyz, _u.mat.  Readers: .mmn, .amn, .chk, .nnkp, _u.mat, _u_dis.mat and the
final state of .wout.  Formats follow src/overlap.F90:152-430 (.mmn/.amn),
src/readwrite.F90:583-635 (.eig), src/wannier90_readwrite.F90:2278-2373
(.chk), src/plot.F90:2224-2288 (_u.mat) and :2790-2862 (_centres.xyz).
"""
import re
import struct
import time

import numpy as np

BOHR = 0.5291772108  # constants.F90 bohr (CODATA 2006)


# ---------------------------------------------------------------- inputs

def write_win(path, real_lattice, mp, kpts, num_wann, num_bands=None, params=None, atoms=None):
    lines = []
    lines.append(f"num_wann = {num_wann}")
    lines.append(f"num_bands = {num_bands or num_wann}")
    for k, v in (params or {}).items():
        if isinstance(v, bool):
            v = ".true." if v else ".false."
        lines.append(f"{k} = {v}")
    lines.append("begin unit_cell_cart")
    lines.append("ang")
    for r in real_lattice:
        lines.append("  " + "  ".join(f"{x:.16f}" for x in r))
    lines.append("end unit_cell_cart")
    if atoms:
        lines.append("begin atoms_frac")
        for sym, pos in atoms:
            lines.append(f"{sym} " + " ".join(f"{x:.12f}" for x in pos))
        lines.append("end atoms_frac")
    lines.append(f"mp_grid = {mp[0]} {mp[1]} {mp[2]}")
    lines.append("begin kpoints")
    for k in kpts:
        lines.append("  " + "  ".join(f"{x:.16f}" for x in k))
    lines.append("end kpoints")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def _fmt_pairs(z):
    z = np.asarray(z).ravel(order="F")
    out = np.empty((z.size, 2))
    out[:, 0] = z.real
    out[:, 1] = z.imag
    return "\n".join(f"{a:.17e} {b:.17e}" for a, b in out)


def write_mmn(path, m, kmesh, comment="written by wannier90_amd.synth"):
    """m (nb,nb,nn,nk) Fortran layout; headers from nnlist/nncell."""
    nb, _, nn, nk = m.shape
    with open(path, "w") as f:
        f.write(comment + "\n")
        f.write(f"{nb:12d}{nk:12d}{nn:12d}\n")
        for k in range(nk):
            for b in range(nn):
                g = kmesh.nncell[:, k, b]
                f.write(f"{k + 1:5d}{kmesh.nnlist[k, b]:5d}{g[0]:5d}{g[1]:5d}{g[2]:5d}\n")
                f.write(_fmt_pairs(m[:, :, b, k]) + "\n")


def write_amn(path, amat_knm, comment="written by wannier90_amd.synth"):
    """amat (nk, nb, nproj) with A_mn(k)."""
    nk, nb, npj = amat_knm.shape
    with open(path, "w") as f:
        f.write(comment + "\n")
        f.write(f"{nb:12d}{nk:12d}{npj:12d}\n")
        rows = []
        for k in range(nk):
            for n in range(npj):
                for m in range(nb):
                    z = amat_knm[k, m, n]
                    rows.append(f"{m + 1:5d}{n + 1:5d}{k + 1:5d} {z.real:.17e} {z.imag:.17e}")
        f.write("\n".join(rows) + "\n")


def write_eig(path, evals_kn):
    nk, nb = evals_kn.shape
    with open(path, "w") as f:
        for k in range(nk):
            for n in range(nb):
                f.write(f"{n + 1:5d}{k + 1:5d}{evals_kn[k, n]:22.14f}\n")


def _sizes(path):
    """The three sizes on the second line of a .mmn or .amn file."""
    with open(path) as f:
        f.readline()
        for line in f:
            if line.strip():
                return tuple(int(x) for x in line.split()[:3])
    raise ValueError(f"{path}: no sizes line")


def read_mmn(path, kmesh, nthreads=0):
    """Reads .mmn into (nb,nb,nn,nk) with b ordered as kmesh (overlap_read),
    through the library's parallel reader (w90gpu_read_mmn)."""
    from . import capi
    nb = _sizes(path)[0]
    return capi.read_mmn(path, nb, kmesh, nthreads)


def read_amn(path, nthreads=0):
    """Reads .amn into (nk, nb, nproj) with A_mn(k) (overlap_read), through
    the library's parallel reader (w90gpu_read_amn)."""
    from . import capi
    nb, nk, npj = _sizes(path)
    a = capi.read_amn(path, nb, nk, npj, nthreads)
    return np.ascontiguousarray(a.transpose(2, 0, 1))


# ------------------------------------------------------------ checkpoint

def _rec(payload):
    n = len(payload)
    return struct.pack("<i", n) + payload + struct.pack("<i", n)


def write_chk(path, real_lattice, kpts, mp, nntot, u, m, centres, spreads, num_bands=None,
              checkpoint="postwann"):
    """Unformatted .chk (w90_wannier90_readwrite_write_chkpt, no disentanglement)."""
    nw = u.shape[0]
    nk = u.shape[2]
    nb = num_bands or nw
    recip = 2.0 * np.pi * np.linalg.inv(real_lattice).T
    header = time.strftime("written on %d%b%Y at %H:%M:%S").ljust(33)[:33]
    out = bytearray()
    out += _rec(header.encode())
    out += _rec(struct.pack("<i", nb))
    out += _rec(struct.pack("<i", 0))
    out += _rec(b"")
    out += _rec(np.asarray(real_lattice, dtype="<f8").tobytes(order="F"))
    out += _rec(np.asarray(recip, dtype="<f8").tobytes(order="F"))
    out += _rec(struct.pack("<i", nk))
    out += _rec(struct.pack("<3i", *mp))
    out += _rec(np.asarray(kpts, dtype="<f8").T.tobytes(order="F"))
    out += _rec(struct.pack("<i", nntot))
    out += _rec(struct.pack("<i", nw))
    out += _rec(checkpoint.ljust(20)[:20].encode())
    out += _rec(struct.pack("<i", 0))  # have_disentangled = .false.
    out += _rec(np.asarray(u, dtype="<c16").tobytes(order="F"))
    out += _rec(np.asarray(m, dtype="<c16").tobytes(order="F"))
    out += _rec(np.asarray(centres, dtype="<f8").tobytes(order="F"))
    out += _rec(np.asarray(spreads, dtype="<f8").tobytes(order="F"))
    with open(path, "wb") as f:
        f.write(out)


def read_chk(path):
    """Reads an unformatted .chk (w90_wannier90_readwrite_write_chkpt,
    src/wannier90_readwrite.F90:2336-2366), with the disentanglement records
    (omega_invariant, lwindow, ndimwin, u_matrix_opt) when present."""
    with open(path, "rb") as f:
        buf = f.read()
    recs = []
    pos = 0
    while pos < len(buf):
        (n,) = struct.unpack_from("<i", buf, pos)
        recs.append(buf[pos + 4: pos + 4 + n])
        pos += 8 + n
    r = iter(recs)
    out = {"header": next(r).decode(errors="replace")}
    out["num_bands"] = struct.unpack("<i", next(r))[0]
    nex = struct.unpack("<i", next(r))[0]
    next(r)
    out["num_exclude_bands"] = nex
    out["real_lattice"] = np.frombuffer(next(r), "<f8").reshape((3, 3), order="F")
    out["recip_lattice"] = np.frombuffer(next(r), "<f8").reshape((3, 3), order="F")
    nk = struct.unpack("<i", next(r))[0]
    out["num_kpts"] = nk
    out["mp_grid"] = struct.unpack("<3i", next(r))
    out["kpt_latt"] = np.frombuffer(next(r), "<f8").reshape((3, nk), order="F")
    nn = struct.unpack("<i", next(r))[0]
    nw = struct.unpack("<i", next(r))[0]
    out["nntot"], out["num_wann"] = nn, nw
    out["checkpoint"] = next(r).decode().strip()
    dis = struct.unpack("<i", next(r))[0]
    out["have_disentangled"] = bool(dis)
    if dis:
        nb = out["num_bands"]
        out["omega_invariant"] = struct.unpack("<d", next(r))[0]
        out["lwindow"] = np.frombuffer(next(r), "<i4").reshape((nb, nk), order="F") != 0
        out["ndimwin"] = np.frombuffer(next(r), "<i4").copy()
        out["u_matrix_opt"] = np.frombuffer(next(r), "<c16").reshape((nb, nw, nk), order="F").copy(order="F")
    out["u_matrix"] = np.frombuffer(next(r), "<c16").reshape((nw, nw, nk), order="F").copy(order="F")
    out["m_matrix"] = np.frombuffer(next(r), "<c16").reshape((nw, nw, nn, nk), order="F").copy(order="F")
    out["centres"] = np.frombuffer(next(r), "<f8").reshape((3, nw), order="F").copy()
    out["spreads"] = np.frombuffer(next(r), "<f8").copy()
    return out


# --------------------------------------------------------------- outputs

def write_centres_xyz(path, centres, real_lattice=None, translate_home_cell=False, atoms=()):
    """_centres.xyz (plot_write_xyz, src/plot.F90:2790-2862)."""
    wc = np.array(centres, dtype=float).T.copy()  # (nw,3)
    if translate_home_cell and real_lattice is not None:
        inv = np.linalg.inv(real_lattice)
        frac = wc @ inv
        frac -= np.floor(frac)
        wc = frac @ real_lattice
    date = time.strftime("%d%b%Y at %H:%M:%S")
    with open(path, "w") as f:
        f.write(f"{len(wc) + len(atoms):6d}\n")
        f.write(f" Wannier centres, written by Wannier90 on{date}\n")
        for c in wc:
            f.write("X      " + "".join(f"{x:14.8f}   " for x in c) + "\n")
        for sym, pos in atoms:
            f.write(f"{sym:<2s}     " + "".join(f"{x:14.8f}   " for x in pos) + "\n")


def write_u_mat(path, u, kpts):
    """_u.mat (plot_u_matrices, src/plot.F90:2224-2288)."""
    nw = u.shape[0]
    nk = u.shape[2]
    date = time.strftime("%d%b%Y at %H:%M:%S")
    with open(path, "w") as f:
        f.write(f" written on {date}\n")
        f.write(f"{nk:12d}{nw:12d}{nw:12d}\n")
        for k in range(nk):
            f.write("\n")
            f.write("".join(f"{x:15.10f}" for x in kpts[k]) + "\n")
            for j in range(nw):
                for i in range(nw):
                    z = u[i, j, k]
                    f.write(f"{z.real:15.10f}{z.imag:15.10f}\n")


def read_u_mat(path):
    with open(path) as f:
        f.readline()
        nk, nw1, nw2 = (int(x) for x in f.readline().split())
        toks = f.read().split()
    u = np.zeros((nw1, nw2, nk), dtype=np.complex128, order="F")
    kpts = np.zeros((nk, 3))
    p = 0
    for k in range(nk):
        kpts[k] = [float(x) for x in toks[p:p + 3]]
        p += 3
        vals = np.array(toks[p:p + 2 * nw1 * nw2], dtype=float).reshape(-1, 2)
        p += 2 * nw1 * nw2
        u[:, :, k] = (vals[:, 0] + 1j * vals[:, 1]).reshape((nw1, nw2), order="F")
    return u, kpts


def read_u_dis_mat(path):
    """_u_dis.mat (plot_u_matrices, src/plot.F90:2266-2285): U_opt as
    (num_bands, num_wann, num_kpts), window rows at their band positions;
    the header gives num_kpts, num_wann, num_bands."""
    with open(path) as f:
        f.readline()
        nk, nw, nb = (int(x) for x in f.readline().split())
        toks = f.read().split()
    u = np.zeros((nb, nw, nk), dtype=np.complex128, order="F")
    kpts = np.zeros((nk, 3))
    p = 0
    for k in range(nk):
        kpts[k] = [float(x) for x in toks[p:p + 3]]
        p += 3
        vals = np.array(toks[p:p + 2 * nb * nw], dtype=float).reshape(-1, 2)
        p += 2 * nb * nw
        u[:, :, k] = (vals[:, 0] + 1j * vals[:, 1]).reshape((nb, nw), order="F")
    return u, kpts


def read_nnkp(path):
    """nnlist/nncell from the nnkps block of a .nnkp file (kmesh_write)."""
    with open(path) as f:
        txt = f.read()
    blk = re.search(r"begin nnkpts\s*\n(.*?)end nnkpts", txt, re.S).group(1).split("\n")
    nn = int(blk[0])
    rows = np.array([l.split() for l in blk[1:] if l.strip()], dtype=int)
    nk = rows.shape[0] // nn
    rows = rows.reshape(nk, nn, 5)
    nnlist = rows[:, :, 1]
    nncell = np.transpose(rows[:, :, 2:5], (2, 0, 1))
    return nn, nnlist, nncell


_FLOAT = r"[-+]?\d*\.?\d+(?:[eEdD][-+]?\d+)?"


def read_wout_final(path):
    """Final centres/spreads and Omega decomposition printed by wann_main."""
    with open(path) as f:
        txt = f.read()
    i = txt.rfind("Final State")
    if i < 0:
        raise ValueError("no Final State in " + path)
    tail = txt[i:]
    cen = re.findall(r"WF centre and spread\s+\d+\s+\(\s*(" + _FLOAT + r"),\s*(" + _FLOAT + r"),\s*("
                     + _FLOAT + r")\s*\)\s+(" + _FLOAT + ")", tail)
    c = np.array([[float(x) for x in row[:3]] for row in cen]).T
    s = np.array([float(row[3]) for row in cen])
    om = {}
    for key, name in (("Omega I", "omega_i"), ("Omega D", "omega_d"), ("Omega OD", "omega_od"),
                      ("Omega Total", "omega_total")):
        mm = re.search(re.escape(key) + r"\s*=\s*(" + _FLOAT + ")", tail)
        if mm:
            om[name] = float(mm.group(1))
    conv = re.findall(r"^\s*(\d+)\s+(" + _FLOAT + r")\s+(" + _FLOAT + r")\s+(" + _FLOAT + r")\s+("
                      + _FLOAT + r")\s+<-- CONV", txt, re.M)
    om["iterations"] = int(conv[-1][0]) if conv else 0
    om["conv_iter"] = [int(r[0]) for r in conv]
    om["conv_omega"] = [float(r[3]) for r in conv]
    om["conv_time"] = [float(r[4]) for r in conv]  # io_wallclocktime, seconds
    return c, s, om


def read_wout_dis(path):
    """The DIS lines of dis_extract (src/disentangle.F90:2209-2212): per
    iteration Omega_I(i-1), Omega_I(i) and the fractional change, as an
    (iterations, 3) array, and the final Omega_I line's value (or None)."""
    with open(path) as f:
        txt = f.read()
    rows = re.findall(r"^\s*(\d+)\s+(" + _FLOAT + r")\s+(" + _FLOAT + r")\s+(" + _FLOAT + r")\s+("
                      + _FLOAT + r")\s+<-- DIS", txt, re.M)
    hist = np.array([[float(r[1]), float(r[2]), float(r[3])] for r in rows]).reshape(-1, 3)
    m = re.search(r"Final Omega_I\s+(" + _FLOAT + ")", txt)
    return hist, (float(m.group(1)) if m else None)
