"""ctypes binding of the CPU oracle (oracle/build/liborc.so).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "build", "liborc.so")

CEDA, MMS = 0, 1
RKC, RKL, DIRK = 0, 1, 2
LSOLVE_FAIL = -7  # ORC_LSOLVE_FAIL, the value of ARKODE's ARK_LSOLVE_FAIL
ST = dict(nst=0, attempts=1, nfe=2, netf=3, nstlim=4, ndomeig=5, ndee_iters=6,
          stage_max=7, last_stages=8, nfe_dee=9, nli=10, npsetup=11, npsolve=12,
          nni=13, ncfn=14, nlsfail=15, njtimes=16, nfe_ls=17)


class Problem(C.Structure):
    _fields_ = [("kind", C.c_int), ("nx", C.c_int64), ("ny", C.c_int64), ("n", C.c_int64),
                ("xl", C.c_double), ("xu", C.c_double), ("yl", C.c_double), ("yu", C.c_double),
                ("dx", C.c_double), ("dy", C.c_double), ("kx", C.c_double), ("ky", C.c_double),
                ("tf", C.c_double), ("inhomogeneous", C.c_int)] + \
               [(f, C.POINTER(C.c_double)) for f in ("fx", "fy", "px", "py", "sx", "sy", "cx", "cy", "diag")] + \
               [("nrhs", C.c_long)]


# the -O3 / OpenMP build of the same sources, for timing the CPU baseline
TIMING_LIB_PATH = os.path.join(ROOT, "oracle", "build", "liborc_omp.so")

_libs = {}


def lib(path=LIB_PATH):
    """The checker (default), or the timing build (path=TIMING_LIB_PATH)."""
    if path not in _libs:
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run make -C oracle)")
        L = C.CDLL(path)
        P = C.POINTER(Problem)
        d, i, vp, dp = C.c_double, C.c_int, C.c_void_p, C.POINTER(C.c_double)
        sig = {
            "orc_problem_defaults": (None, [P, i]),
            "orc_problem_setup": (i, [P]),
            "orc_problem_free": (None, [P]),
            "orc_initial": (None, [P, dp]),
            "orc_exact": (None, [P, d, dp]),
            "orc_rhs": (i, [P, d, dp, dp]),
            "orc_stencil": (None, [P, dp, dp]),
            "orc_psetup": (None, [P, d]),
            "orc_psolve": (None, [P, dp, dp]),
            "orc_dom_eig": (None, [P, dp, dp]),
            "orc_seed_vector": (None, [C.c_int64, C.c_uint64, dp]),
            "orc_create": (vp, [P, i, d, dp]),
            "orc_free": (None, [vp]),
            "orc_set_tolerances": (None, [vp, d, d]),
            "orc_set_fixed_step": (None, [vp, d]),
            "orc_set_init_step": (None, [vp, d]),
            "orc_set_controller": (None, [vp, i]),
            "orc_set_max_num_steps": (None, [vp, C.c_long]),
            "orc_set_internal_eig": (None, [vp, i, i, d]),
            "orc_set_dom_eig_frequency": (None, [vp, C.c_long]),
            "orc_set_max_num_stages": (None, [vp, i]),
            "orc_set_dom_eig_safety": (None, [vp, d]),
            "orc_set_dee_preprocess": (None, [vp, i, i]),
            "orc_set_order": (None, [vp, i]),
            "orc_set_linear": (None, [vp, i]),
            "orc_set_lin_solver": (None, [vp, i, d, i, i]),
            "orc_set_ls_reltol": (None, [vp, d]),
            "orc_set_jtimes_analytic": (None, [vp, C.c_int]),
            "orc_evolve": (i, [vp, d, dp, dp]),
            "orc_get_stats": (i, [vp, C.POINTER(C.c_long), i]),
            "orc_get_current_h": (d, [vp]),
            "orc_get_spectral_radius": (d, [vp]),
            "orc_test_linear_solve": (i, [vp, d, d, dp, dp, d, i, C.POINTER(C.c_long)]),
            "orc_dirk_table": (i, [i, dp, dp, dp, dp]),
            "orc_vlinearcombination": (None, [i, dp, C.POINTER(dp), dp, C.c_int64]),
            "orc_vwrmsnorm": (d, [dp, dp, C.c_int64]),
            "orc_vdot": (d, [dp, dp, C.c_int64]),
            "orc_vmaxnorm": (d, [dp, C.c_int64]),
            "orc_set_threads": (i, [i]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _libs[path] = L
    return _libs[path]


def ptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OracleProblem:
    """Owns an orc_problem; kwargs override the reference defaults."""

    def __init__(self, kind=CEDA, lib_path=LIB_PATH, **kw):
        self.L = lib(lib_path)
        self.p = Problem()
        self.L.orc_problem_defaults(C.byref(self.p), kind)
        for k, v in kw.items():
            setattr(self.p, k, v)
        if self.L.orc_problem_setup(C.byref(self.p)) != 0:
            raise ValueError("oracle problem setup failed")
        self.n = self.p.n
        self.shape = (self.p.ny, self.p.nx)

    def __del__(self):
        try:
            self.L.orc_problem_free(C.byref(self.p))
        except Exception:
            pass

    def table(self, name, length):
        return np.ctypeslib.as_array(getattr(self.p, name), shape=(length,)).copy()

    def initial(self):
        u = np.zeros(self.n)
        self.L.orc_initial(C.byref(self.p), ptr(u))
        return u

    def exact(self, t):
        u = np.zeros(self.n)
        self.L.orc_exact(C.byref(self.p), t, ptr(u))
        return u

    def rhs(self, t, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        f = np.zeros(self.n)
        self.L.orc_rhs(C.byref(self.p), t, ptr(u), ptr(f))
        return f

    def psetup(self, gamma):
        self.L.orc_psetup(C.byref(self.p), gamma)
        return np.ctypeslib.as_array(self.p.diag, shape=(self.n,)).copy()

    def psolve(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros(self.n)
        self.L.orc_psolve(C.byref(self.p), ptr(r), ptr(z))
        return z

    def mg_apply(self, gamma, r, relax=2, nu=2):
        """One V-cycle of the multigrid preconditioner (orc_mg.c): smoother
        relax type 2 (symmetric red/black Gauss-Seidel) or 3, nu sweep pairs."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros(self.n)
        f = self.L.orc_mg_apply_opts
        f.restype = C.c_int
        f.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_double),
                      C.POINTER(C.c_double)]
        nl = f(C.cast(C.pointer(self.p), C.c_void_p), gamma, relax, nu, ptr(r), ptr(z))
        return z, nl

    def dom_eig(self):
        lr, li = C.c_double(), C.c_double()
        self.L.orc_dom_eig(C.byref(self.p), C.byref(lr), C.byref(li))
        return lr.value, li.value


def seed_vector(n, seed):
    v = np.zeros(n)
    lib().orc_seed_vector(n, seed, ptr(v))
    return v


class OracleIntegrator:
    def __init__(self, prob, method, y0, t0=0.0):
        self.prob = prob
        y0 = np.ascontiguousarray(y0, dtype=np.float64)
        self.L = prob.L
        self.h = self.L.orc_create(C.byref(prob.p), method, t0, ptr(y0))
        self.n = prob.n

    def __del__(self):
        try:
            self.L.orc_free(self.h)
        except Exception:
            pass

    def __getattr__(self, name):
        f = getattr(self.L, "orc_" + name)
        return lambda *a: f(self.h, *a)

    def evolve(self, tout):
        y = np.zeros(self.n)
        t = C.c_double()
        flag = self.L.orc_evolve(self.h, tout, ptr(y), C.byref(t))
        return flag, t.value, y

    def stats(self):
        out = (C.c_long * 32)()
        m = self.L.orc_get_stats(self.h, out, 32)
        return {k: out[v] for k, v in ST.items() if v < m}

    def linear_solve(self, t, gamma, b, reltol, maxl):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n)
        it = C.c_long()
        ret = self.L.orc_test_linear_solve(self.h, t, gamma, ptr(b), ptr(x), reltol, maxl, C.byref(it))
        return ret, x, it.value


def dirk_table(order):
    A = np.zeros(16)
    b = np.zeros(4)
    bh = np.zeros(4)
    c = np.zeros(4)
    s = lib().orc_dirk_table(order, ptr(A), ptr(b), ptr(bh), ptr(c))
    return A.reshape(4, 4)[:s, :s], b[:s], bh[:s], c[:s]


def lp_num_params(dil):
    d = np.ascontiguousarray(dil, dtype=np.int32)
    f = lib().orc_lp_num_params
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.POINTER(C.c_int)]
    return f(len(d), d.ctypes.data_as(C.POINTER(C.c_int)))


def lp_num_params_sym(dil):
    d = np.ascontiguousarray(dil, dtype=np.int32)
    f = lib().orc_lp_sym_num_params
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.POINTER(C.c_int)]
    return f(len(d), d.ctypes.data_as(C.POINTER(C.c_int)))


def lp_lossgrad(dil, params, out_scale, nb, nx, ny, ratio, r, e, emu=1, want_grad=True, sym=False, ratio_y=None):
    """Loss, parameter gradient and normalised network output of the learned
    preconditioner restated on the CPU (emu=1: BF16 roundings of the GPU path
    emulated; emu=0: exact FP64; sym: the symmetric network K^T K), for the
    operator I - ratio Dxx - ratio_y Dyy (ratio_y defaults to ratio)."""
    d = np.ascontiguousarray(dil, dtype=np.int32)
    p = np.ascontiguousarray(params, dtype=np.float32)
    r = np.ascontiguousarray(r, dtype=np.float64)
    e = np.ascontiguousarray(e, dtype=np.float64)
    g = np.zeros(len(p)) if want_grad else None
    u = np.zeros(nb * nx * ny)
    f = lib().orc_lp_sym_lossgrad if sym else lib().orc_lp_lossgrad
    f.restype = C.c_double
    f.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_float), C.c_double, C.c_int, C.c_int,
                  C.c_int, C.c_double, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int,
                  C.POINTER(C.c_double), C.POINTER(C.c_double)]
    ry = ratio if ratio_y is None else ratio_y
    loss = f(len(d), d.ctypes.data_as(C.POINTER(C.c_int)), p.ctypes.data_as(C.POINTER(C.c_float)),
             out_scale, nb, nx, ny, ratio, ry, ptr(r), ptr(e), emu, ptr(g) if g is not None else None, ptr(u))
    return loss, g, u
