"""Seeded synthetic wannierisation problems (tight-binding models).

A problem is a tight-binding model with num_wann orbitals at fractional
positions tau in one cell, Hermitian hoppings H(R) for R in [-2,2]^3
(on-site eps ~ N(0,1) eV, hoppings t ~ CN(0,1) exp(-|d|/decay)), diagonalised
on a Gamma-centred Monkhorst-Pack grid in wannier90's k order (first index
slowest).  With eigenvectors C(k) in the periodic gauge, random per-k band
phases, and b the Cartesian b-vectors of the k-mesh:

    M_mn(k,b) = sum_a conj(C_am(k)) exp(-i b.tau_a) C_an(k+b)
    A_mn(k)   = sum_a conj(C_am(k)) g_an,     g ~ CN(0,1)

Because M(k,b) is unitary, Omega_I = 0, and U(k) = C(k)^+ (up to per-WF
phases) is a global minimum with Omega = 0 and centres at tau (Cartesian)
when |b.tau| < pi.  This closed form pins the solver independently of any
reference run.

Seeding: base seed 20261202; config c uses seed + c; item i of the ragged
batch uses splitmix64(seed + 4, i).
"""
import math

import numpy as np

from . import capi

BASE_SEED = 20261202
MASK64 = (1 << 64) - 1


def splitmix64(seed, i):
    """i-th output of the splitmix64 generator started at `seed`."""
    z = (seed + (i + 1) * 0x9E3779B97F4A7C15) & MASK64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


def lattice(kind, a=5.0, c_over_a=1.6, rng=None):
    """Rows are lattice vectors in Angstrom."""
    if kind == "sc":
        return a * np.eye(3)
    if kind == "fcc":
        return 0.5 * a * np.array([[0, 1, 1], [1, 0, 1], [1, 1, 0]], dtype=float)
    if kind == "bcc":
        return 0.5 * a * np.array([[-1, 1, 1], [1, -1, 1], [1, 1, -1]], dtype=float)
    if kind == "hex":
        return np.array([[a, 0, 0], [-0.5 * a, 0.5 * math.sqrt(3) * a, 0], [0, 0, c_over_a * a]])
    if kind == "triclinic":
        lens = rng.uniform(4.0, 7.0, size=3)
        al, be, ga = np.deg2rad(rng.uniform(75.0, 105.0, size=3))
        a1 = np.array([lens[0], 0.0, 0.0])
        a2 = lens[1] * np.array([math.cos(ga), math.sin(ga), 0.0])
        cx = math.cos(be)
        cy = (math.cos(al) - math.cos(be) * math.cos(ga)) / math.sin(ga)
        cz = math.sqrt(max(1.0 - cx * cx - cy * cy, 1e-12))
        a3 = lens[2] * np.array([cx, cy, cz])
        return np.array([a1, a2, a3])
    raise ValueError(kind)


def mp_kpoints(mp):
    """Gamma-centred MP grid, fractional, first index slowest."""
    n1, n2, n3 = mp
    pts = [(i / n1, j / n2, k / n3) for i in range(n1) for j in range(n2) for k in range(n3)]
    return np.array(pts, dtype=float)


class Problem:
    """A generated problem: lattice, k-mesh, model, M, A and the oracle U*."""

    def __init__(self, **kw):
        self.__dict__.update(kw)

    @property
    def num_kpts(self):
        return self.kpts.shape[0]


def make_problem(seed, kind="sc", a=5.0, nw=4, mp=(4, 4, 4), decay=1.5, c_over_a=1.6,
                 rmax=2, with_m=True, kmesh=None):
    rng = np.random.default_rng(seed)
    real_lattice = lattice(kind, a=a, c_over_a=c_over_a, rng=rng)
    kpts = mp_kpoints(mp)
    nk = kpts.shape[0]
    tau = rng.uniform(0.0, 1.0, size=(nw, 3))
    tau_cart = tau @ real_lattice
    # hoppings on half of the R vectors, mirrored for Hermiticity
    rs = [np.array(r) for r in np.ndindex(2 * rmax + 1, 2 * rmax + 1, 2 * rmax + 1)]
    rs = [r - rmax for r in rs]
    hr = {}
    eps = rng.normal(size=nw)
    for r in rs:
        key = tuple(int(x) for x in r)
        neg = tuple(-x for x in key)
        if neg in hr:
            hr[key] = hr[neg].conj().T
            continue
        d = (r @ real_lattice)[None, None, :] + tau_cart[None, :, :] - tau_cart[:, None, :]
        dist = np.linalg.norm(d, axis=2)
        t = (rng.normal(size=(nw, nw)) + 1j * rng.normal(size=(nw, nw))) / math.sqrt(2.0)
        t = t * np.exp(-dist / decay)
        if key == (0, 0, 0):
            t = 0.5 * (t + t.conj().T)
            np.fill_diagonal(t, eps)
        hr[key] = t
    rvec = np.array(list(hr.keys()), dtype=float)
    hmat = np.array(list(hr.values()))
    phase = np.exp(2j * math.pi * (kpts @ rvec.T))           # (nk, nR)
    hk = (phase @ hmat.reshape(len(hmat), nw * nw)).reshape(nk, nw, nw)
    hk = 0.5 * (hk + np.conj(np.transpose(hk, (0, 2, 1))))
    evals, cvec = np.linalg.eigh(hk)                           # cvec[k][:, n]
    gauge = np.exp(2j * math.pi * rng.uniform(size=(nk, nw)))
    cvec = cvec * gauge[:, None, :]
    g = (rng.normal(size=(nw, nw)) + 1j * rng.normal(size=(nw, nw))) / math.sqrt(2.0)
    amat = np.einsum("kam,an->kmn", cvec.conj(), g)           # A(k) = C^+ g
    if kmesh is None:
        kmesh = capi.kmesh_get(real_lattice, kpts)
    prob = Problem(seed=seed, kind=kind, real_lattice=real_lattice, mp=tuple(mp), kpts=kpts,
                   tau=tau, tau_cart=tau_cart, evals=evals, cvec=cvec, amat=amat, kmesh=kmesh,
                   num_wann=nw, decay=decay)
    if with_m:
        prob.m_orig = overlaps(prob)
    return prob


def overlaps(prob):
    """M(k,b) in Fortran layout (nw,nw,nn,nk)."""
    km = prob.kmesh
    nk, nw, nn = prob.num_kpts, prob.num_wann, km.nntot
    m = np.empty((nw, nw, nn, nk), dtype=np.complex128, order="F")
    c = prob.cvec
    for b in range(nn):
        bvec = km.bk[:, b, :].T                                # (nk,3)
        ph = np.exp(-1j * (bvec @ prob.tau_cart.T))            # (nk, nw): e^{-i b.tau_a}
        k2 = km.nnlist[:, b] - 1
        # C(k)^+ diag(e^{-i b.tau}) C(k+b), one batched product per b
        blk = np.matmul(np.conj(np.transpose(c, (0, 2, 1))), ph[:, :, None] * c[k2])
        m[:, :, b, :] = np.transpose(blk, (1, 2, 0))
    return m


def lowdin(amat):
    """U(k) = Z V^+ from the SVD of A(k) (overlap_project, src/overlap.F90:758-952)."""
    z, _, vh = np.linalg.svd(amat)
    return z @ vh


def initial_state(prob):
    """Initial (U, M) of wann_main: Lowdin U and M <- U^+(k) M U(k+b)."""
    u = lowdin(prob.amat)                                     # (nk, nw, nw)
    km = prob.kmesh
    m0 = prob.m_orig
    nw, nn, nk = prob.num_wann, km.nntot, prob.num_kpts
    m = np.empty_like(m0)
    for b in range(nn):
        k2 = km.nnlist[:, b] - 1
        blk = np.transpose(m0[:, :, b, :], (2, 0, 1))
        r = np.conj(np.transpose(u, (0, 2, 1))) @ blk @ u[k2]
        m[:, :, b, :] = np.transpose(r, (1, 2, 0))
    uf = np.asfortranarray(np.transpose(u, (1, 2, 0)))
    return uf, np.asfortranarray(m)


def make_dis_problem(seed, kind="sc", num_bands=12, num_wann=8, mp=(3, 3, 3), decay=1.5, a=5.0, c_over_a=1.6,
                     frozen=0, mix=0.3):
    """A disentanglement problem: the tight-binding model of make_problem
    with num_bands orbitals (its bands, eigenvalues and overlaps M(k,b) of
    all num_bands), projections A(k) = C(k)^+ g onto num_wann trial
    orbitals g = the first num_wann orbitals mixed with the others by
    `mix`, and an inner window that freezes, where frozen > 0, the bands
    below the lowest energy of band `frozen` over k (at most `frozen` bands
    at any k).  Returns (Problem, dis control dict)."""
    p = make_problem(seed, kind=kind, a=a, nw=num_bands, mp=mp, decay=decay, c_over_a=c_over_a)
    rng = np.random.default_rng(seed + 1)
    g = np.eye(num_bands)[:, :num_wann].astype(complex)
    g += mix * (rng.normal(size=(num_bands, num_wann)) + 1j * rng.normal(size=(num_bands, num_wann)))
    p.amat_dis = np.asfortranarray(np.transpose(np.einsum("kam,an->kmn", p.cvec.conj(), g), (1, 2, 0)))
    p.eig = np.asfortranarray(p.evals.T)
    p.num_bands, p.num_wann_dis = num_bands, num_wann
    ctl = dict(num_iter=200, mix_ratio=0.5, conv_tol=1e-10, conv_window=3)
    if frozen > 0:
        ctl.update(frozen_states=1, froz_max=float(p.evals[:, frozen - 1].min()))
    return p, ctl


# ---------------------------------------------------------------- configs

# The three single-problem workloads: 1 = sc a = 5 Ang, 4 bands, 4^3 mesh
# (64 k, the small CPU-sized case); 2 = fcc a = 5.43 Ang, 16 bands, 8^3 mesh
# (512 k, the benchmark problem); 3 = triclinic, 64 bands, 10x10x8 mesh
# (800 k, nntot 12, the MFMA-bound case).
CONFIGS = {
    1: dict(kind="sc", a=5.0, nw=4, mp=(4, 4, 4), decay=1.5),
    2: dict(kind="fcc", a=5.43, nw=16, mp=(8, 8, 8), decay=2.0),
    3: dict(kind="triclinic", nw=64, mp=(10, 10, 8), decay=2.5),
}


def config(c, with_m=True, **over):
    kw = dict(CONFIGS[c])
    kw.update(over)
    return make_problem(BASE_SEED + c, with_m=with_m, **kw)


def batch_params(i, seed=BASE_SEED + 4, nw_range=(4, 48), meshes=(4, 6, 8)):
    """Parameters of item i of the ragged batch (lattice, num_wann, mesh)."""
    s = splitmix64(seed, i)
    rng = np.random.default_rng(s)
    kind = ["sc", "fcc", "bcc", "hex"][int(rng.integers(0, 4))]
    nw = int(rng.integers(nw_range[0], nw_range[1] + 1))
    n = int(meshes[int(rng.integers(0, len(meshes)))])
    a = float(rng.uniform(4.0, 6.0))
    coa = float(rng.uniform(1.2, 1.8))
    return dict(seed=s % (1 << 63), kind=kind, nw=nw, mp=(n, n, n), a=a, c_over_a=coa,
                decay=1.5)


def batch_item(i, **kw):
    p = batch_params(i, **kw)
    seed = p.pop("seed")
    return make_problem(seed, **p)
