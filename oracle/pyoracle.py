"""ctypes access to the C restatement (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  Build with `make -C oracle`.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
REF_EXE = os.path.join(HERE, "_ref", "wannier90.x")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OProblem(C.Structure):
    _fields_ = [("nw", C.c_int), ("nk", C.c_int), ("nn", C.c_int), ("nnlist", _ip), ("bk", _dp),
                ("wb", _dp), ("wbtot", C.c_double)]


class OControl(C.Structure):
    _fields_ = [("num_iter", C.c_int), ("num_cg_steps", C.c_int), ("conv_window", C.c_int),
                ("lfixstep", C.c_int), ("conv_tol", C.c_double), ("trial_step", C.c_double),
                ("fixed_step", C.c_double)]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise OSError(f"{LIB} not built: run `make -C oracle`")
        # MKL selects its CPU kernels by dlopen; they need the core symbols global
        lib = C.CDLL(LIB, mode=C.RTLD_GLOBAL)
        P = C.POINTER(OProblem)
        lib.oracle_omega.argtypes = [P, _dp, C.c_int, C.c_double, _dp, _dp, _dp, _dp, _dp]
        lib.oracle_omega.restype = None
        lib.oracle_domega.argtypes = [P, _dp, _dp, _dp]
        lib.oracle_domega.restype = None
        lib.oracle_new_u_and_m.argtypes = [P, _dp, _dp, _dp]
        lib.oracle_new_u_and_m.restype = C.c_int
        lib.oracle_check_unitarity.argtypes = [P, _dp]
        lib.oracle_check_unitarity.restype = C.c_double
        lib.oracle_lowdin.argtypes = [P, _dp, _dp, _dp]
        lib.oracle_lowdin.restype = C.c_int
        lib.oracle_wann_main.argtypes = [P, C.POINTER(OControl), _dp, _dp, _dp, _dp, _dp, _dp, _ip]
        lib.oracle_wann_main.restype = C.c_int
        _lib = lib
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


class Problem:
    def __init__(self, nw, nk, kmesh):
        self.nnl = np.asfortranarray(kmesh.nnlist, dtype=np.int32)
        self.bk = np.asfortranarray(kmesh.bk, dtype=np.float64)
        self.wb = np.ascontiguousarray(kmesh.wb, dtype=np.float64)
        self.s = OProblem(nw, nk, kmesh.nntot, self.nnl.ctypes.data_as(_ip), _d(self.bk),
                          _d(self.wb), float(kmesh.wbtot))
        self.nw, self.nk, self.nn = nw, nk, kmesh.nntot


def omega(prob, m, first_pass=True, omega_invariant=0.0):
    lib = load()
    nw, nk, nn = prob.nw, prob.nk, prob.nn
    m = np.asfortranarray(m, dtype=np.complex128)
    rave = np.zeros((3, nw), order="F")
    r2ave = np.zeros(nw)
    rave2 = np.zeros(nw)
    ln = np.zeros((nw, nn, nk), order="F")
    om = np.zeros(4)
    lib.oracle_omega(C.byref(prob.s), _d(m), 1 if first_pass else 0, omega_invariant, _d(rave),
                     _d(r2ave), _d(rave2), _d(ln), _d(om))
    return rave, r2ave, om, ln


def domega(prob, m):
    lib = load()
    m = np.asfortranarray(m, dtype=np.complex128)
    g = np.zeros((prob.nw, prob.nw, prob.nk), dtype=np.complex128, order="F")
    rave = np.zeros((3, prob.nw), order="F")
    lib.oracle_domega(C.byref(prob.s), _d(m), _d(g), _d(rave))
    return g, rave


def new_u_and_m(prob, cdq, u, m):
    """Returns rotated copies (U, M)."""
    lib = load()
    cdq = np.array(cdq, dtype=np.complex128, order="F")
    u = np.array(u, dtype=np.complex128, order="F")
    m = np.array(m, dtype=np.complex128, order="F")
    info = lib.oracle_new_u_and_m(C.byref(prob.s), _d(cdq), _d(u), _d(m))
    if info:
        raise RuntimeError(f"zheev info={info}")
    return u, m


def lowdin(prob, a_nwk, m):
    """a (nw,nw,nk) Fortran; returns (U, rotated M)."""
    lib = load()
    a = np.array(a_nwk, dtype=np.complex128, order="F")
    u = np.zeros_like(a)
    m = np.array(m, dtype=np.complex128, order="F")
    info = lib.oracle_lowdin(C.byref(prob.s), _d(a), _d(u), _d(m))
    if info:
        raise RuntimeError(f"zgesvd info={info}")
    return u, m


def unitarity(prob, u):
    return load().oracle_check_unitarity(C.byref(prob.s), _d(np.asfortranarray(u)))


def wann_main(prob, ctl, m, u):
    """Returns dict with final m, u, centres, spreads, omega, history, iterations."""
    lib = load()
    m = np.array(m, dtype=np.complex128, order="F")
    u = np.array(u, dtype=np.complex128, order="F")
    c = OControl(int(ctl["num_iter"]), int(ctl["num_cg_steps"]), int(ctl["conv_window"]),
                 int(ctl.get("lfixstep", 0)), float(ctl["conv_tol"]), float(ctl["trial_step"]),
                 float(ctl.get("fixed_step", -999.0)))
    cen = np.zeros((3, prob.nw), order="F")
    spr = np.zeros(prob.nw)
    om = np.zeros(4)
    hist = np.zeros((c.num_iter + 1, 4))
    stats = np.zeros(2, dtype=np.int32)
    info = lib.oracle_wann_main(C.byref(prob.s), C.byref(c), _d(m), _d(u), _d(cen), _d(spr), _d(om),
                                _d(hist), stats.ctypes.data_as(_ip))
    if info:
        raise RuntimeError(f"LAPACK info={info}")
    it = int(stats[0])
    return dict(m=m, u=u, centres=cen, spreads=spr, omega=om, history=hist[: it + 1],
                iterations=it, converged=bool(stats[1]))


# --------------------------------------------------------- disentanglement

class ODisProblem(C.Structure):
    _fields_ = [("nb", C.c_int), ("nw", C.c_int), ("nk", C.c_int), ("nn", C.c_int), ("nnlist", _ip),
                ("wb", _dp), ("wbtot", C.c_double)]


class ODisControl(C.Structure):
    _fields_ = [("win_min", C.c_double), ("win_max", C.c_double), ("froz_min", C.c_double),
                ("froz_max", C.c_double), ("frozen_states", C.c_int), ("num_iter", C.c_int),
                ("mix_ratio", C.c_double), ("conv_tol", C.c_double), ("conv_window", C.c_int)]


# dis_control / dis_manifold defaults (src/wannier90_types.F90:129-135,
# src/types.F90:183-189)
DIS_DEFAULTS = dict(win_min=-1.7976931348623157e308, win_max=1.7976931348623157e308,
                    froz_min=-1.7976931348623157e308, froz_max=1.7976931348623157e308, frozen_states=0,
                    num_iter=200, mix_ratio=0.5, conv_tol=1e-10, conv_window=3)
DIS_ERRORS = {1: "outer window contains no eigenvalues", 2: "window has fewer states than num_wann",
              3: "more frozen states than num_wann", 4: "unitarity of initial U in dis_project",
              5: "LAPACK failure", 6: "dis_proj_froz error", 7: "orthonormality of the trial orbitals"}


def dis_main(num_bands, num_wann, kmesh, ctl, eig, a, m):
    """dis_main + setup_m_loc.  eig (nb,nk), a (nb,nw,nk), m (nb,nb,nn,nk).
    Returns a dict: u_opt (nb,nw,nk), u (nw,nw,nk), m (nw,nw,nn,nk) in the
    gauge of u (the wannierise input), eigval_opt (nb,nk), nfirstwin and
    ndimwin (1-based first band, size), ndimfroz, omega_invariant, history
    (iterations,3): Omega_I estimate, Omega_I, delta; iterations, converged."""
    lib = load()
    if not hasattr(lib, "_dis"):
        lib.oracle_dis_main.argtypes = [C.POINTER(ODisProblem), C.POINTER(ODisControl), _dp, _dp, _dp, _dp, _dp,
                                        _dp, _dp, _ip, _ip, _ip, _dp, _dp, _ip]
        lib.oracle_dis_main.restype = C.c_int
        lib._dis = True
    nb, nw = num_bands, num_wann
    nk = eig.shape[1]
    nn = kmesh.nntot
    nnl = np.asfortranarray(kmesh.nnlist, dtype=np.int32)
    wb = np.ascontiguousarray(kmesh.wb, dtype=np.float64)
    p = ODisProblem(nb, nw, nk, nn, nnl.ctypes.data_as(_ip), _d(wb), float(kmesh.wbtot))
    c = dict(DIS_DEFAULTS)
    c.update(ctl)
    oc = ODisControl(*(c[f] for f, _ in ODisControl._fields_))
    eig = np.asfortranarray(eig, dtype=np.float64)
    a = np.array(a, dtype=np.complex128, order="F")
    m = np.array(m, dtype=np.complex128, order="F")
    uopt = np.zeros((nb, nw, nk), dtype=np.complex128, order="F")
    u = np.zeros((nw, nw, nk), dtype=np.complex128, order="F")
    mw = np.zeros((nw, nw, nn, nk), dtype=np.complex128, order="F")
    eigopt = np.zeros((nb, nk), order="F")
    nfirst = np.zeros(nk, dtype=np.int32)
    ndim = np.zeros(nk, dtype=np.int32)
    nfroz = np.zeros(nk, dtype=np.int32)
    om = np.zeros(1)
    hist = np.zeros((max(int(c["num_iter"]), 1), 3))
    conv = np.zeros(1, dtype=np.int32)
    ip = lambda v: v.ctypes.data_as(_ip)  # noqa: E731
    rc = lib.oracle_dis_main(C.byref(p), C.byref(oc), _d(eig), _d(a), _d(m), _d(uopt), _d(u), _d(mw), _d(eigopt),
                             ip(nfirst), ip(ndim), ip(nfroz), _d(om), _d(hist), ip(conv))
    if rc < 0:
        raise RuntimeError(f"oracle_dis_main: {DIS_ERRORS.get(-rc, -rc)}")
    return dict(u_opt=uopt, u=u, m=mw, eigval_opt=eigopt, nfirstwin=nfirst, ndimwin=ndim, ndimfroz=nfroz,
                omega_invariant=float(om[0]), history=hist[:rc], iterations=rc, converged=bool(conv[0]))
