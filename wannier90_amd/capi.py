"""ctypes binding of libw90gpu.so (include/w90gpu.h).

Arrays follow wannier90's Fortran layout: complex128 numpy arrays in
order='F' with shapes m(nw,nw,nn,nk), u(nw,nw,nk), cdq(nw,nw,nk);
nnlist int32 (nk,nn) 1-based; bk float64 (3,nn,nk); wb float64 (nn).

The library is loaded from wannier90_amd/lib; there is no fallback: if the
library is missing or has no GPU, calls raise.  The host-side entry points
(k-mesh search, .mmn/.amn readers) are also built without any HIP code into
libw90io.so, which kmesh_get/read_mmn/read_amn use: a process that only
generates or reads problems never starts the GPU runtime.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libw90gpu.so")
HOST_LIB_PATH = os.path.join(HERE, "lib", "libw90io.so")

OK = 0
ERR_FATAL = 1
ERR_ALLOC = 2
ERR_INPUT = 5
ERR_FILE = 6
WARN_UNCONVERGED = -1
NUM_NNMAX = 12
MAX_WANN = 256
PREC_FP64 = 0
PREC_MIXED = 1


class W90Error(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"w90gpu error {code}: {msg}")
        self.code = code


class Control(C.Structure):
    _fields_ = [
        ("num_iter", C.c_int), ("num_cg_steps", C.c_int), ("conv_window", C.c_int),
        ("lfixstep", C.c_int), ("conv_tol", C.c_double), ("trial_step", C.c_double),
        ("fixed_step", C.c_double), ("conv_noise_amp", C.c_double), ("precond", C.c_int),
        ("guiding_centres", C.c_int), ("selective_loc", C.c_int), ("precision", C.c_int),
    ]


class Problem(C.Structure):
    _fields_ = [
        ("num_wann", C.c_int), ("num_kpts", C.c_int), ("nntot", C.c_int),
        ("nnlist", C.POINTER(C.c_int)), ("bk", C.POINTER(C.c_double)),
        ("wb", C.POINTER(C.c_double)), ("wbtot", C.c_double),
    ]


class DisControl(C.Structure):
    """w90gpu_dis_control (dis_control_type / dis_manifold_type)."""
    _fields_ = [("win_min", C.c_double), ("win_max", C.c_double), ("froz_min", C.c_double),
                ("froz_max", C.c_double), ("frozen_states", C.c_int), ("num_iter", C.c_int),
                ("mix_ratio", C.c_double), ("conv_tol", C.c_double), ("conv_window", C.c_int)]


class DisResult(C.Structure):
    _fields_ = [("omega_invariant", C.c_double), ("seconds", C.c_double), ("sweeps_mean", C.c_double),
                ("iterations", C.c_int), ("converged", C.c_int), ("status", C.c_int), ("sweeps_max", C.c_int)]


class Result(C.Structure):
    _fields_ = [
        ("omega_total", C.c_double), ("omega_invariant", C.c_double), ("omega_d", C.c_double),
        ("omega_od", C.c_double), ("omega_tilde", C.c_double), ("gcnorm1", C.c_double),
        ("delta_omega", C.c_double), ("solve_seconds", C.c_double), ("iterations", C.c_int),
        ("converged", C.c_int), ("unitarity_ok", C.c_int), ("status", C.c_int),
        ("bulk_iterations", C.c_int), ("held_second_steps", C.c_int),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None
_host_lib = None

# (name, restype, argtypes) of every entry point of include/w90gpu.h
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p
SIGNATURES = [
    ("w90gpu_control_defaults", None, [C.POINTER(Control)]),
    ("w90gpu_create", C.c_int, [C.c_int, C.POINTER(_vp)]),
    ("w90gpu_destroy", C.c_int, [_vp]),
    ("w90gpu_last_error", C.c_char_p, [_vp]),
    ("w90gpu_version", C.c_char_p, []),
    ("w90gpu_kmesh_get", C.c_int, [_dp, C.c_int, _dp, C.c_int, C.c_double, C.c_int, _ip, C.c_int,
                                   _ip, _ip, _ip, _dp, _dp, _dp, C.c_char_p, C.c_int]),
    ("w90gpu_read_mmn", C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, _ip, _ip, _dp, C.c_int,
                                  C.c_char_p, C.c_int]),
    ("w90gpu_read_amn", C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, _dp, C.c_int, C.c_char_p,
                                  C.c_int]),
    ("w90gpu_wann_main", C.c_int, [_vp, C.POINTER(Problem), C.POINTER(Control), _dp, _dp, _dp, _dp,
                                   C.POINTER(Result)]),
    ("w90gpu_omega", C.c_int, [_vp, C.POINTER(Problem), _dp, C.c_int, C.c_double, _dp, _dp, _dp]),
    ("w90gpu_omega_noalloc", C.c_int, [_vp, C.POINTER(Problem), _dp, C.c_int, C.c_double, _dp, _dp, _dp]),
    ("w90gpu_dis_control_defaults", None, [C.POINTER(DisControl)]),
    ("w90gpu_disentangle", C.c_int, [_vp, C.POINTER(Problem), C.c_int, C.POINTER(DisControl), _dp, _dp, _dp, _dp,
                                     _dp, _dp, _dp, _ip, _ip, _dp, C.c_int, C.POINTER(DisResult)]),
    ("w90gpu_domega", C.c_int, [_vp, C.POINTER(Problem), _dp, _dp]),
    ("w90gpu_new_u_and_m", C.c_int, [_vp, C.POINTER(Problem), _dp, _dp, _dp]),
    ("w90gpu_project", C.c_int, [_vp, C.POINTER(Problem), _dp, _dp, _dp]),
    ("w90gpu_session_create", C.c_int, [_vp, C.POINTER(Problem), C.POINTER(Control), C.POINTER(_vp)]),
    ("w90gpu_session_destroy", C.c_int, [_vp]),
    ("w90gpu_session_load", C.c_int, [_vp, _dp, _dp]),
    ("w90gpu_session_solve", C.c_int, [_vp, C.POINTER(Result)]),
    ("w90gpu_session_launch", C.c_int, [_vp]),
    ("w90gpu_session_wait", C.c_int, [_vp, C.POINTER(Result)]),
    ("w90gpu_session_fetch", C.c_int, [_vp, _dp, _dp, _dp, _dp]),
    ("w90gpu_session_bytes", C.c_size_t, [C.c_int, C.c_int, C.c_int]),
    ("w90gpu_session_history", C.c_int, [_vp, _dp, C.c_int]),
    ("w90gpu_session_rotation_time", C.c_int, [_vp, C.c_int, _dp]),
    ("w90gpu_session_phase_time", C.c_int, [_vp, C.c_int, C.c_int, _dp]),
    ("w90gpu_problem_bytes", C.c_size_t, [C.c_int, C.c_int, C.c_int, C.c_int]),
    ("w90gpu_batch_create", C.c_int, [_vp, C.c_size_t, C.POINTER(_vp)]),
    ("w90gpu_batch_destroy", C.c_int, [_vp]),
    ("w90gpu_batch_add", C.c_int, [_vp, C.POINTER(Problem), C.POINTER(Control), _dp, _dp, _ip]),
    ("w90gpu_batch_step", C.c_int, [_vp, C.c_int, _ip, C.c_int, _ip]),
    ("w90gpu_batch_run", C.c_int, [_vp]),
    ("w90gpu_batch_restart", C.c_int, [_vp]),
    ("w90gpu_batch_result", C.c_int, [_vp, C.c_int, C.POINTER(Result)]),
    ("w90gpu_batch_fetch", C.c_int, [_vp, C.c_int, _dp, _dp, _dp, _dp]),
    ("w90gpu_batch_remove", C.c_int, [_vp, C.c_int]),
    ("w90gpu_batch_info", C.c_int, [_vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t), _ip, _ip]),
    ("w90gpu_tune_get", C.c_int, [C.c_int, C.c_int, C.c_char_p, C.c_size_t]),
    ("w90gpu_tune_set", C.c_int, [C.c_int, C.c_int, C.c_char_p]),
]


HOST_ENTRY_POINTS = ("w90gpu_kmesh_get", "w90gpu_read_mmn", "w90gpu_read_amn")


def load_host():
    """Load libw90io.so (host code only: no HIP runtime is loaded)."""
    global _host_lib
    if _host_lib is None:
        if not os.path.exists(HOST_LIB_PATH):
            raise OSError(f"{HOST_LIB_PATH} not built: run `python -m wannier90_amd.build`")
        lib = C.CDLL(HOST_LIB_PATH)
        for name, res, args in SIGNATURES:
            if name in HOST_ENTRY_POINTS:
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
        _host_lib = lib
    return _host_lib


def load(path=None):
    """Load libw90gpu.so and declare every entry point; raise if absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise OSError(f"{p} not built: run `python -m wannier90_amd.build`")
    lib = C.CDLL(p)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def dis_control(**kw):
    c = DisControl()
    load().w90gpu_dis_control_defaults(C.byref(c))
    for k, v in kw.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    return c


def control(**kw):
    c = Control()
    load().w90gpu_control_defaults(C.byref(c))
    for k, v in kw.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    return c


class KMesh:
    """k-mesh tables in the reference's layout (kmesh_info_type)."""

    def __init__(self, nntot, nnlist, nncell, bk, wb, wbtot):
        self.nntot = nntot
        self.nnlist = nnlist  # (nk, nn) int32, 1-based, order='F'
        self.nncell = nncell  # (3, nk, nn) int32, order='F'
        self.bk = bk          # (3, nn, nk) float64, order='F'
        self.wb = wb          # (nn,)
        self.wbtot = wbtot


ROT_KERNELS = ("w16", "fused32", "fused64", "direct64", "gemm64", "gemm32", "fused48")


def tune_get(num_wann, nntot):
    """Rotation kernel a problem of this shape admitted now would use."""
    lib = load()
    buf = C.create_string_buffer(32)
    rc = lib.w90gpu_tune_get(num_wann, nntot, buf, len(buf))
    if rc:
        raise W90Error(rc, lib.w90gpu_last_error(None).decode())
    return buf.value.decode()


def tune_set(num_wann, nntot, name):
    """Fix the rotation kernel of (num_wann, nntot) for problems admitted
    from now on; None drops a fixed choice."""
    lib = load()
    rc = lib.w90gpu_tune_set(num_wann, nntot, name.encode() if name else None)
    if rc:
        raise W90Error(rc, lib.w90gpu_last_error(None).decode())


def kmesh_get(real_lattice, kpt_latt, search_shells=36, tol=1e-6, shell_list=None,
              skip_b1_tests=False):
    """Shells, weights and neighbour lists (kmesh_get, src/kmesh.F90:61).

    real_lattice: (3,3) rows are lattice vectors in Angstrom.
    kpt_latt: (nk,3) fractional k-points.
    """
    lib = load_host()
    rl = np.asfortranarray(np.asarray(real_lattice, dtype=np.float64))
    kp = np.ascontiguousarray(np.asarray(kpt_latt, dtype=np.float64))  # (nk,3) C == (3,nk) F
    nk = kp.shape[0]
    nnmax = NUM_NNMAX
    nnlist = np.zeros(nk * nnmax, dtype=np.int32)
    nncell = np.zeros(3 * nk * nnmax, dtype=np.int32)
    bk = np.zeros(3 * nnmax * nk, dtype=np.float64)
    wb = np.zeros(nnmax, dtype=np.float64)
    wbtot = C.c_double(0.0)
    nntot = C.c_int(0)
    err = C.create_string_buffer(256)
    if shell_list:
        sl = np.asarray(shell_list, dtype=np.int32)
        nsh, slp = len(sl), _i(sl)
    else:
        nsh, slp = 0, None
    rc = lib.w90gpu_kmesh_get(_d(rl), nk, _d(kp), search_shells, tol, nsh, slp,
                              1 if skip_b1_tests else 0, C.byref(nntot), _i(nnlist), _i(nncell),
                              _d(bk), _d(wb), C.byref(wbtot), err, 256)
    if rc != OK:
        raise W90Error(rc, err.value.decode())
    nn = nntot.value
    return KMesh(
        nn,
        np.asfortranarray(nnlist[: nk * nn].reshape((nk, nn), order="F")),
        np.asfortranarray(nncell[: 3 * nk * nn].reshape((3, nk, nn), order="F")),
        np.asfortranarray(bk[: 3 * nn * nk].reshape((3, nn, nk), order="F")),
        wb[:nn].copy(),
        wbtot.value,
    )


def read_mmn(path, num_bands, kmesh, nthreads=0):
    """.mmn file -> M (num_bands, num_bands, nntot, nk) complex, order='F'
    (w90gpu_read_mmn; overlap_read, src/overlap.F90:234-308): each block
    goes to the b of the k-mesh with the block's k' and G.  Host threads
    parse the file; nthreads <= 0 takes OMP_NUM_THREADS or all cores."""
    lib = load_host()
    nnl = np.asfortranarray(kmesh.nnlist, dtype=np.int32)
    nnc = np.asfortranarray(kmesh.nncell, dtype=np.int32)
    nk, nn = nnl.shape
    m = np.empty((num_bands, num_bands, nn, nk), dtype=np.complex128, order="F")
    err = C.create_string_buffer(512)
    rc = lib.w90gpu_read_mmn(os.fsencode(path), num_bands, nk, nn, _i(nnl), _i(nnc), m.ctypes.data_as(_dp),
                             nthreads, err, 512)
    if rc != OK:
        raise W90Error(rc, err.value.decode())
    return m


def read_amn(path, num_bands, num_kpts, num_proj, nthreads=0):
    """.amn file -> A (num_bands, num_proj, nk) complex, order='F'
    (w90gpu_read_amn; overlap_read, src/overlap.F90:323-366)."""
    lib = load_host()
    a = np.empty((num_bands, num_proj, num_kpts), dtype=np.complex128, order="F")
    err = C.create_string_buffer(512)
    rc = lib.w90gpu_read_amn(os.fsencode(path), num_bands, num_kpts, num_proj, a.ctypes.data_as(_dp), nthreads,
                             err, 512)
    if rc != OK:
        raise W90Error(rc, err.value.decode())
    return a


class Context:
    """A device context (w90gpu_create)."""

    def __init__(self, device=0):
        self.lib = load()
        h = _vp()
        rc = self.lib.w90gpu_create(device, C.byref(h))
        if rc != OK:
            raise W90Error(rc, self.lib.w90gpu_last_error(None).decode())
        self.h = h

    def close(self):
        if self.h:
            self.lib.w90gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def error(self):
        return self.lib.w90gpu_last_error(self.h).decode()

    def _check(self, rc):
        if rc > 0:
            raise W90Error(rc, self.error())
        return rc

    # -- problem marshalling -------------------------------------------
    @staticmethod
    def problem(num_wann, kmesh, num_kpts):
        nnl = np.asfortranarray(kmesh.nnlist, dtype=np.int32)
        bk = np.asfortranarray(kmesh.bk, dtype=np.float64)
        wb = np.ascontiguousarray(kmesh.wb, dtype=np.float64)
        p = Problem(num_wann, num_kpts, kmesh.nntot, _i(nnl), _d(bk), _d(wb), float(kmesh.wbtot))
        p._keep = (nnl, bk, wb)
        return p

    # -- entry points ---------------------------------------------------
    def wann_main(self, kmesh, m, u, ctl=None):
        """Spread minimisation in place on m (nw,nw,nn,nk) and u (nw,nw,nk)."""
        _check_fortran(m, 4)
        _check_fortran(u, 3)
        nw, nk = u.shape[0], u.shape[2]
        prob = self.problem(nw, kmesh, nk)
        ctl = ctl or control()
        centres = np.zeros((3, nw), order="F")
        spreads = np.zeros(nw)
        res = Result()
        rc = self.lib.w90gpu_wann_main(self.h, C.byref(prob), C.byref(ctl), _d(m), _d(u),
                                       _d(centres), _d(spreads), C.byref(res))
        self._check(rc)
        return centres, spreads, res

    def omega(self, kmesh, m, first_pass=True, omega_invariant=0.0):
        _check_fortran(m, 4)
        nw, nk = m.shape[0], m.shape[3]
        prob = self.problem(nw, kmesh, nk)
        rave = np.zeros((3, nw), order="F")
        r2ave = np.zeros(nw)
        om = np.zeros(4)
        self._check(self.lib.w90gpu_omega(self.h, C.byref(prob), _d(m), 1 if first_pass else 0,
                                          omega_invariant, _d(rave), _d(r2ave), _d(om)))
        return rave, r2ave, om

    def disentangle(self, kmesh, num_bands, ctl, eig, a, m):
        """w90gpu_disentangle: eig (nb,nk), a (nb,nw,nk), m (nb,nb,nn,nk), ctl
        a dict of w90gpu_dis_control fields (the rest at their defaults).
        Returns u_opt, u, m (nw,nw,nn,nk), eigval_opt, nfirstwin, ndimwin,
        history (iterations,3), omega_invariant, iterations, converged,
        seconds."""
        _check_fortran(m, 4)
        _check_fortran(a, 3)
        nb, nw, nk = num_bands, a.shape[1], a.shape[2]
        nn = kmesh.nntot
        prob = self.problem(nw, kmesh, nk)
        c = dis_control(**ctl)
        eig = np.asfortranarray(eig, dtype=np.float64)
        uopt = np.zeros((nb, nw, nk), dtype=np.complex128, order="F")
        u = np.zeros((nw, nw, nk), dtype=np.complex128, order="F")
        mw = np.zeros((nw, nw, nn, nk), dtype=np.complex128, order="F")
        eigopt = np.zeros((nb, nk), order="F")
        nfirst = np.zeros(nk, dtype=np.int32)
        ndim = np.zeros(nk, dtype=np.int32)
        rows = max(int(c.num_iter), 1)
        hist = np.zeros((rows, 3))
        res = DisResult()
        self._check(self.lib.w90gpu_disentangle(self.h, C.byref(prob), nb, C.byref(c), _d(eig), _d(a), _d(m),
                                                _d(uopt), _d(u), _d(mw), _d(eigopt), _i(nfirst), _i(ndim),
                                                _d(hist), rows, C.byref(res)))
        return dict(u_opt=uopt, u=u, m=mw, eigval_opt=eigopt, nfirstwin=nfirst, ndimwin=ndim,
                    history=hist[: res.iterations], omega_invariant=res.omega_invariant,
                    iterations=res.iterations, converged=bool(res.converged), seconds=res.seconds,
                    sweeps_mean=res.sweeps_mean, sweeps_max=res.sweeps_max)

    def omega_noalloc(self, kmesh, m, first_pass=True, omega_invariant=0.0):
        """omega() through the allocation-free entry point (static device
        scratch)."""
        _check_fortran(m, 4)
        nw, nk = m.shape[0], m.shape[3]
        prob = self.problem(nw, kmesh, nk)
        rave = np.zeros((3, nw), order="F")
        r2ave = np.zeros(nw)
        om = np.zeros(4)
        self._check(self.lib.w90gpu_omega_noalloc(self.h, C.byref(prob), _d(m), 1 if first_pass else 0,
                                                  omega_invariant, _d(rave), _d(r2ave), _d(om)))
        return rave, r2ave, om

    def domega(self, kmesh, m):
        _check_fortran(m, 4)
        nw, nk = m.shape[0], m.shape[3]
        prob = self.problem(nw, kmesh, nk)
        g = np.zeros((nw, nw, nk), dtype=np.complex128, order="F")
        self._check(self.lib.w90gpu_domega(self.h, C.byref(prob), _d(m), _d(g)))
        return g

    def new_u_and_m(self, kmesh, cdq, u, m):
        _check_fortran(m, 4)
        _check_fortran(u, 3)
        _check_fortran(cdq, 3)
        nw, nk = u.shape[0], u.shape[2]
        prob = self.problem(nw, kmesh, nk)
        self._check(self.lib.w90gpu_new_u_and_m(self.h, C.byref(prob), _d(cdq), _d(u), _d(m)))

    def project(self, kmesh, a, m):
        """Lowdin projection: returns U from A (nw,nw,nk); rotates m in place."""
        _check_fortran(a, 3)
        _check_fortran(m, 4)
        nw, nk = a.shape[0], a.shape[2]
        prob = self.problem(nw, kmesh, nk)
        u = np.zeros_like(a, order="F")
        self._check(self.lib.w90gpu_project(self.h, C.byref(prob), _d(a), _d(u), _d(m)))
        return u

    def session(self, kmesh, num_wann, num_kpts, ctl=None):
        return Session(self, kmesh, num_wann, num_kpts, ctl or control())

    def batch(self, pool_bytes):
        return Batch(self, pool_bytes)


class Session:
    """Device-resident problem (w90gpu_session_*)."""

    def __init__(self, ctx, kmesh, num_wann, num_kpts, ctl):
        self.ctx = ctx
        self.lib = ctx.lib
        self.nw, self.nk, self.nn = num_wann, num_kpts, kmesh.nntot
        self.prob = ctx.problem(num_wann, kmesh, num_kpts)
        self.ctl = ctl
        h = _vp()
        ctx._check(self.lib.w90gpu_session_create(ctx.h, C.byref(self.prob), C.byref(ctl), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.w90gpu_session_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load(self, m, u):
        _check_fortran(m, 4)
        _check_fortran(u, 3)
        self._keep = (m, u)
        self.ctx._check(self.lib.w90gpu_session_load(self.h, _d(m), _d(u)))

    def solve(self):
        res = Result()
        self.ctx._check(self.lib.w90gpu_session_solve(self.h, C.byref(res)))
        return res

    def launch(self):
        self.ctx._check(self.lib.w90gpu_session_launch(self.h))

    def wait(self):
        res = Result()
        self.ctx._check(self.lib.w90gpu_session_wait(self.h, C.byref(res)))
        return res

    def fetch(self, want_m=True):
        nw, nk, nn = self.nw, self.nk, self.nn
        m = np.zeros((nw, nw, nn, nk), dtype=np.complex128, order="F") if want_m else None
        u = np.zeros((nw, nw, nk), dtype=np.complex128, order="F")
        centres = np.zeros((3, nw), order="F")
        spreads = np.zeros(nw)
        self.ctx._check(self.lib.w90gpu_session_fetch(self.h, _d(m) if want_m else None, _d(u),
                                                      _d(centres), _d(spreads)))
        return m, u, centres, spreads

    def history(self, rows=None):
        rows = rows if rows is not None else self.ctl.num_iter + 1
        out = np.zeros((rows, 4))
        n = self.lib.w90gpu_session_history(self.h, _d(out), rows)
        if n < 0:
            raise W90Error(-n, self.ctx.error())
        return out[:n]

    def rotation_time(self, reps=50):
        """Average duration (ms) of the M-rotation kernel, `reps` launches
        back to back on the final state of the last solve."""
        ms = C.c_double(0.0)
        self.ctx._check(self.lib.w90gpu_session_rotation_time(self.h, reps, C.byref(ms)))
        return ms.value

    PHASES = {"rotation": 0, "gradient": 1, "retraction": 2, "reduction": 3}

    def phase_time(self, phase, reps=50):
        """Average duration (ms) of one phase of an iteration ("rotation",
        "gradient", "retraction" or "reduction"), `reps` launches back to
        back on the final state of the last solve; the solve's results are
        unchanged."""
        ms = C.c_double(0.0)
        self.ctx._check(self.lib.w90gpu_session_phase_time(self.h, self.PHASES[phase], reps, C.byref(ms)))
        return ms.value


class Batch:
    """Many problems solved together on one device (w90gpu_batch_*)."""

    def __init__(self, ctx, pool_bytes):
        self.ctx = ctx
        self.lib = ctx.lib
        h = _vp()
        ctx._check(self.lib.w90gpu_batch_create(ctx.h, int(pool_bytes), C.byref(h)))
        self.h = h
        self.shapes = {}

    def close(self):
        if self.h:
            self.lib.w90gpu_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, kmesh, m, u, ctl=None):
        """Admit a problem; returns its id, or None when the pool has no room."""
        _check_fortran(m, 4)
        _check_fortran(u, 3)
        nw, nk = u.shape[0], u.shape[2]
        prob = self.ctx.problem(nw, kmesh, nk)
        ctl = ctl or control()
        i = C.c_int(-1)
        rc = self.lib.w90gpu_batch_add(self.h, C.byref(prob), C.byref(ctl), _d(m), _d(u), C.byref(i))
        if rc == ERR_ALLOC:
            return None
        self.ctx._check(rc)
        self.shapes[i.value] = (nw, nk, kmesh.nntot)
        return i.value

    def step(self, blocks=1, cap=4096):
        """Start added problems and run `blocks` blocks; ids finished since
        the last call."""
        out = np.zeros(cap, dtype=np.int32)
        n = C.c_int(0)
        self.ctx._check(self.lib.w90gpu_batch_step(self.h, blocks, _i(out), cap, C.byref(n)))
        return [int(x) for x in out[: n.value]]

    def run(self):
        self.ctx._check(self.lib.w90gpu_batch_run(self.h))

    def restart(self):
        self.ctx._check(self.lib.w90gpu_batch_restart(self.h))

    def result(self, i):
        res = Result()
        self.ctx._check(self.lib.w90gpu_batch_result(self.h, i, C.byref(res)))
        return res

    def fetch(self, i, want_m=False):
        nw, nk, nn = self.shapes[i]
        m = np.zeros((nw, nw, nn, nk), dtype=np.complex128, order="F") if want_m else None
        u = np.zeros((nw, nw, nk), dtype=np.complex128, order="F")
        centres = np.zeros((3, nw), order="F")
        spreads = np.zeros(nw)
        self.ctx._check(self.lib.w90gpu_batch_fetch(self.h, i, _d(m) if want_m else None, _d(u),
                                                    _d(centres), _d(spreads)))
        return m, u, centres, spreads

    def remove(self, i):
        self.ctx._check(self.lib.w90gpu_batch_remove(self.h, i))
        self.shapes.pop(i, None)

    def info(self):
        pool = C.c_size_t(0)
        used = C.c_size_t(0)
        nr = C.c_int(0)
        nq = C.c_int(0)
        self.ctx._check(self.lib.w90gpu_batch_info(self.h, C.byref(pool), C.byref(used), C.byref(nr),
                                                   C.byref(nq)))
        return {"pool_bytes": pool.value, "used_bytes": used.value, "running": nr.value, "queued": nq.value}


def problem_bytes(num_wann, num_kpts, nntot, num_iter):
    return int(load().w90gpu_problem_bytes(num_wann, num_kpts, nntot, num_iter))


def session_bytes(num_wann, num_kpts, nntot):
    return int(load().w90gpu_session_bytes(num_wann, num_kpts, nntot))


def _check_fortran(a, ndim):
    if not isinstance(a, np.ndarray) or a.dtype != np.complex128 or a.ndim != ndim:
        raise TypeError(f"expected complex128 array with {ndim} dims")
    if not a.flags.f_contiguous:
        raise TypeError("array must be Fortran-contiguous (order='F')")
