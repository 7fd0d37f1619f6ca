"""ctypes binding of the MI355X diffusion_2D library (libceda_hip.so).

The compute path is the HIP C-ABI library built from ``csrc/``; this module
only declares its signatures and wraps handles (SUNContext, N_Vector,
DiffusionProblem, integrator memory) in small Python objects for tests and the
benchmark.  There is no fallback: if the library is missing, ``lib()`` raises.
The entry points keep the reference's names (``diffusion``, ``PSetup``,
``PSolve``, ``Initial``, ``dom_eig`` of diffusion_2D/diffusion_2D.hpp:246-271 and
main.cpp:536) and the SUNDIALS API names the reference's main.cpp calls.
"""
import ctypes as C
import os
import re
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libceda_hip.so")
INCLUDE = os.path.join(ROOT, "include")

CEDA_PERIODIC, MMS_DIRICHLET = 0, 1
RKC, RKL = 0, 1
ARK_NORMAL, ARK_ONE_STEP = 1, 2
SUN_PREC_NONE, SUN_PREC_RIGHT = 0, 2
STATS = ["nst", "attempts", "nfe", "netf", "nstlim", "ndomeig", "ndee_iters", "stage_max",
         "last_stages", "nfe_dee", "nli", "npsetup", "npsolve", "nni", "ncfn", "nlsfail",
         "njtimes", "nfe_ls"]


class Params(C.Structure):
    _fields_ = [("kind", C.c_int), ("nx", C.c_int64), ("ny", C.c_int64),
                ("xl", C.c_double), ("xu", C.c_double), ("yl", C.c_double), ("yu", C.c_double),
                ("kx", C.c_double), ("ky", C.c_double), ("inhomogeneous", C.c_int),
                ("tf", C.c_double)]


_lib = None
vp, d, i32, i64, dp = C.c_void_p, C.c_double, C.c_int, C.c_int64, C.POINTER(C.c_double)
RHSFN = C.CFUNCTYPE(C.c_int, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p)

_SIGS = {
    "SUNContext_Create": (i32, [i32, C.POINTER(vp)]),
    "SUNContext_Free": (i32, [C.POINTER(vp)]),
    "N_VNew_HipNative": (vp, [i64, vp]),
    "N_VMake_HipNative": (vp, [i64, vp, vp]),
    "N_VDestroy": (None, [vp]),
    "N_VClone": (vp, [vp]),
    "N_VGetLength": (i64, [vp]),
    "N_VGetArrayPointer": (dp, [vp]),
    "N_VGetDeviceArrayPointer": (dp, [vp]),
    "N_VCopyToDevice_HipNative": (i32, [vp]),
    "N_VCopyFromDevice_HipNative": (i32, [vp]),
    "N_VSeedRandom_HipNative": (i32, [vp, C.c_uint64]),
    "N_VLinearSum": (None, [d, vp, d, vp, vp]),
    "N_VConst": (None, [d, vp]),
    "N_VProd": (None, [vp, vp, vp]),
    "N_VDiv": (None, [vp, vp, vp]),
    "N_VScale": (None, [d, vp, vp]),
    "N_VAbs": (None, [vp, vp]),
    "N_VInv": (None, [vp, vp]),
    "N_VAddConst": (None, [vp, d, vp]),
    "N_VDotProd": (d, [vp, vp]),
    "N_VMaxNorm": (d, [vp]),
    "N_VWrmsNorm": (d, [vp, vp]),
    "N_VWrmsNormMask": (d, [vp, vp, vp]),
    "N_VMin": (d, [vp]),
    "N_VWL2Norm": (d, [vp, vp]),
    "N_VL1Norm": (d, [vp]),
    "N_VCompare": (None, [d, vp, vp]),
    "N_VInvTest": (i32, [vp, vp]),
    "N_VConstrMask": (i32, [vp, vp, vp]),
    "N_VMinQuotient": (d, [vp, vp]),
    "N_VLinearCombination": (i32, [i32, dp, C.POINTER(vp), vp]),
    "N_VScaleAddMulti": (i32, [i32, dp, vp, C.POINTER(vp), C.POINTER(vp)]),
    "N_VDotProdMulti": (i32, [i32, vp, C.POINTER(vp), dp]),
    "DiffusionParams_Default": (None, [C.POINTER(Params), i32]),
    "DiffusionProblem_Create": (vp, [C.POINTER(Params), vp]),
    "DiffusionProblem_Destroy": (None, [vp]),
    "DiffusionProblem_GetGrid": (None, [vp, C.POINTER(i64), C.POINTER(i64), dp, dp]),
    "DiffusionProblem_GetNumRhsEvals": (C.c_long, [vp]),
    "DiffusionProblem_SetStencilRows": (i32, [vp, i32]),
    "DiffusionProblem_TimeKernel": (i32, [vp, i32, vp, vp, vp, vp, vp, i32, dp]),
    "DiffusionProblem_SetStencilConfig": (i32, [vp, i32, i32, i32, i32]),
    "DiffusionProblem_GetStencilConfig": (i32, [vp, i32, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]),
    "DiffusionProblem_Autotune": (i32, [vp, i32]),
    "DiffusionProblem_StagePass": (i32, [vp, i32, dp, vp, vp, vp, vp, vp, vp]),
    "Initial": (i32, [d, vp, vp]),
    "Exact": (i32, [d, vp, vp]),
    "laplacian": (i32, [d, vp, vp, vp]),
    "diffusion": (i32, [d, vp, vp, vp]),
    "PSetup": (i32, [d, vp, vp, i32, C.POINTER(i32), d, vp]),
    "PSolve": (i32, [d, vp, vp, vp, vp, d, d, i32, vp]),
    "MGPSetup": (i32, [d, vp, vp, i32, C.POINTER(i32), d, vp]),
    "MGPSolve": (i32, [d, vp, vp, vp, vp, d, d, i32, vp]),
    "DiffusionProblem_MGLevels": (i32, [vp]),
    "DiffusionProblem_SetMGOptions": (i32, [vp, i32, i32]),
    "DiffusionProblem_MGNumCycles": (C.c_long, [vp]),
    "dom_eig": (i32, [d, vp, vp, dp, dp, vp, vp, vp, vp]),
    "LSRKStepCreateSTS": (vp, [vp, d, vp, vp]),
    "ARKStepCreate": (vp, [vp, vp, d, vp, vp]),
    "ARKodeFree": (None, [C.POINTER(vp)]),
    "LSRKStepSetSTSMethod": (i32, [vp, i32]),
    "LSRKStepSetDomEigFn": (i32, [vp, vp]),
    "LSRKStepSetDomEigEstimator": (i32, [vp, vp]),
    "LSRKStepSetDomEigFrequency": (i32, [vp, C.c_long]),
    "LSRKStepSetMaxNumStages": (i32, [vp, i32]),
    "LSRKStepSetDomEigSafetyFactor": (i32, [vp, d]),
    "LSRKStepSetNumDomEigEstInitPreprocessIters": (i32, [vp, i32]),
    "LSRKStepSetNumDomEigEstPreprocessIters": (i32, [vp, i32]),
    "LSRKStepSetDiffusionFastPath": (i32, [vp, vp, i32]),
    "ARKodeSetDiffusionFastPCG": (i32, [vp, vp]),
    "ARKodeSetFusedPCGBulk": (i32, [vp, d]),
    "LSRKStepGetSpectralRadius": (i32, [vp, dp]),
    "SUNDomEigEstimator_Power": (vp, [vp, C.c_long, d, vp]),
    "SUNDomEigEstimator_Destroy": (i32, [C.POINTER(vp)]),
    "ARKodeSStolerances": (i32, [vp, d, d]),
    "ARKodeSetUserData": (i32, [vp, vp]),
    "ARKodeSetOrder": (i32, [vp, i32]),
    "ARKodeSetFixedStep": (i32, [vp, d]),
    "ARKodeSetInitStep": (i32, [vp, d]),
    "ARKodeSetAdaptController": (i32, [vp, vp]),
    "ARKodeSetMaxNumSteps": (i32, [vp, C.c_long]),
    "ARKodeSetStopTime": (i32, [vp, d]),
    "ARKodeEvolve": (i32, [vp, d, vp, dp, i32]),
    "ARKodeGetStatsArray": (i32, [vp, C.POINTER(C.c_long), i32]),
    "ARKodeSetLinearSolver": (i32, [vp, vp, vp]),
    "ARKodeSetPreconditioner": (i32, [vp, vp, vp]),
    "ARKodeSetLSetupFrequency": (i32, [vp, i32]),
    "ARKodeSetEpsLin": (i32, [vp, d]),
    "ARKodeSetLinear": (i32, [vp, i32]),
    "ARKodeSetNonlinear": (i32, [vp]),
    "ARKodeSetLSFallbackPCG": (i32, [vp, i32, vp, vp]),
    "ARKodeSetJacTimes": (i32, [vp, vp, vp]),
    "ARKodeGetNumLSFallbacks": (i32, [vp, C.POINTER(C.c_long)]),
    "LearnedPrec_Create": (vp, [i32, C.POINTER(i32), vp]),
    "LearnedPrec_CreateSymmetric": (vp, [i32, C.POINTER(i32), vp]),
    "LearnedPrec_IsSymmetric": (i32, [vp]),
    "LearnedPrec_Destroy": (None, [vp]),
    "LearnedPrec_NumParams": (i32, [vp]),
    "LearnedPrec_SetParams": (i32, [vp, C.POINTER(C.c_float)]),
    "LearnedPrec_GetParams": (i32, [vp, C.POINTER(C.c_float)]),
    "LearnedPrec_InitIdentity": (i32, [vp, d, C.c_ulong]),
    "LearnedPrec_SetOutputScale": (i32, [vp, d]),
    "LearnedPrec_GetOutputScale": (d, [vp]),
    "LearnedPrec_Apply": (i32, [vp, i32, i32, i32, vp, vp, i32]),
    "LearnedPrec_ApplyNoAlloc": (i32, [vp, i32, i32, i32, vp, vp, i32]),
    "LearnedPrec_SetConvKernel": (i32, [i32]),
    "LearnedPrec_SetPairTuning": (i32, [i32, i32]),
    "LearnedPrec_LossGrad": (i32, [vp, i32, i32, i32, d, vp, vp, dp, C.POINTER(C.c_float)]),
    "LearnedPrec_MakeTrainingSet": (i32, [i32, i32, d, d, C.c_ulong, vp, vp, C.POINTER(i32), vp]),
    "LearnedPrec_Train": (i32, [vp, i32, i32, d, vp, vp, i32, i32, d, C.c_ulong, dp]),
    "LearnedPrec_LossGradXY": (i32, [vp, i32, i32, i32, d, d, vp, vp, dp, C.POINTER(C.c_float)]),
    "LearnedPrec_MakeTrainingSetXY": (i32, [i32, i32, d, d, d, C.c_ulong, vp, vp, C.POINTER(i32), vp]),
    "LearnedPrec_TrainXY": (i32, [vp, i32, i32, d, d, vp, vp, i32, i32, d, C.c_ulong, dp]),
    "LearnedPrec_SetChainTasks": (i32, [i32]),
    "DiffusionProblem_SetLearnedPrec": (i32, [vp, vp]),
    "ARKodeSetLSRelativeTolerance": (i32, [vp, d]),
    "SUNLinSol_PCG": (vp, [vp, i32, i32, vp]),
    "SUNLinSol_FPCG": (vp, [vp, i32, i32, vp]),
    "SUNLinSol_SPGMR": (vp, [vp, i32, i32, vp]),
    "SUNLinSol_SPFGMR": (vp, [vp, i32, i32, vp]),
    "SUNLinSol_SPFGMRSetMaxRestartDim": (i32, [vp, i32]),
    "SUNLinSolFree": (i32, [vp]),
    "SUNAdaptController_PID": (vp, [vp]),
    "SUNAdaptController_PI": (vp, [vp]),
    "SUNAdaptController_I": (vp, [vp]),
    "SUNAdaptController_ExpGus": (vp, [vp]),
    "SUNAdaptController_ImpGus": (vp, [vp]),
    "SUNAdaptController_ImExGus": (vp, [vp]),
    "SUNAdaptController_Destroy": (i32, [vp]),
    "SUNContext_GetHipStream": (vp, [vp]),
    "DiffusionBatch_Create": (vp, [i32, C.POINTER(Params), i32, vp]),
    "DiffusionBatch_Destroy": (None, [vp]),
    "DiffusionBatch_SetTolerances": (i32, [vp, d, d]),
    "DiffusionBatch_SetController": (i32, [vp, i32]),
    "DiffusionBatch_SetMaxNumStages": (i32, [vp, i32]),
    "DiffusionBatch_SetMaxNumSteps": (i32, [vp, C.c_long]),
    "DiffusionBatch_SetFixedStep": (i32, [vp, d]),
    "DiffusionBatch_SetTileRows": (i32, [vp, i32]),
    "DiffusionBatch_SetThinPhases": (i32, [vp, i32, i32]),
    "DiffusionBatch_SetStagesPerPass": (i32, [vp, i32]),
    "DiffusionBatch_Evolve": (i32, [vp, d]),
    "DiffusionBatch_GetSolution": (i32, [vp, i32, vp]),
    "DiffusionBatch_GetCurrentTime": (i32, [vp, i32, dp]),
    "DiffusionBatch_GetStatsArray": (i32, [vp, i32, C.POINTER(C.c_long), i32]),
    "DiffusionBatch_GetNumGroups": (i32, [vp, C.POINTER(C.c_long), C.POINTER(C.c_long)]),
    "DiffusionBatch_GetAlgorithmicBytes": (i32, [vp, C.POINTER(d)]),
}


def lib():
    """Load libceda_hip.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            if hasattr(L, name):
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
        _lib = L
    return _lib


def header_functions():
    """Every function declared in include/*.h (name list)."""
    names = set()
    pat = re.compile(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", re.M)
    for fn in sorted(os.listdir(INCLUDE)):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for line in text.split(";"):
            line = line.strip()
            if not line or line.startswith("#") or line.startswith("typedef") or "(*" in line:
                continue
            m = pat.search(line)
            if m and not line.startswith("struct") and not line.startswith("enum"):
                names.add(m.group(1))
    return sorted(names)


class Context:
    """SUNContext.  Objects created from it register here; close() destroys
    the live ones first (vectors, problems and integrators must be freed
    before their context, as in SUNDIALS)."""

    def __init__(self):
        h = vp()
        if lib().SUNContext_Create(0, C.byref(h)) != 0:
            raise RuntimeError("SUNContext_Create failed")
        self.h = h
        self.children = weakref.WeakSet()

    def register(self, obj):
        self.children.add(obj)

    def close(self):
        if self.h:
            kids = list(self.children)
            for kind in (Integrator, Batch, LearnedPrec, Vector, Problem):
                for k in kids:
                    if isinstance(k, kind):
                        k.release()
            lib().SUNContext_Free(C.byref(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Vector:
    """N_Vector handle of the HIP-native vector."""

    def __init__(self, ctx, n=None, handle=None, owner=True):
        self.ctx = ctx
        self.h = handle if handle is not None else lib().N_VNew_HipNative(n, ctx.h)
        if not self.h:
            raise RuntimeError("N_VNew_HipNative failed")
        self.n = lib().N_VGetLength(self.h)
        self.owner = owner
        ctx.register(self)

    @classmethod
    def from_numpy(cls, ctx, a):
        v = cls(ctx, n=len(a))
        v.set(a)
        return v

    def set(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == self.n
        host = lib().N_VGetArrayPointer(self.h)
        C.memmove(host, a.ctypes.data, 8 * self.n)
        lib().N_VCopyToDevice_HipNative(self.h)

    def get(self):
        lib().N_VCopyFromDevice_HipNative(self.h)
        host = lib().N_VGetArrayPointer(self.h)
        return np.ctypeslib.as_array(host, shape=(self.n,)).copy()

    def clone(self):
        return Vector(self.ctx, handle=lib().N_VClone(self.h))

    def destroy(self):
        if self.h and self.owner and self.ctx.h:
            lib().N_VDestroy(self.h)
        self.h = None

    release = destroy

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def vec_array(vs):
    return (vp * len(vs))(*[v.h for v in vs])


def dbl_array(xs):
    return (C.c_double * len(xs))(*xs)


class Problem:
    """DiffusionProblem: the reference's UserData on one GPU."""

    def __init__(self, ctx, kind=CEDA_PERIODIC, **kw):
        self.ctx = ctx
        p = Params()
        lib().DiffusionParams_Default(C.byref(p), kind)
        for k, v in kw.items():
            setattr(p, k, v)
        self.params = p
        self.h = lib().DiffusionProblem_Create(C.byref(p), ctx.h)
        if not self.h:
            raise ValueError("DiffusionProblem_Create rejected the parameters")
        nx, ny, dx, dy = i64(), i64(), d(), d()
        lib().DiffusionProblem_GetGrid(self.h, C.byref(nx), C.byref(ny), C.byref(dx), C.byref(dy))
        self.nx, self.ny, self.dx, self.dy = nx.value, ny.value, dx.value, dy.value
        self.n = self.nx * self.ny
        ctx.register(self)

    def vector(self, a=None):
        v = Vector(self.ctx, n=self.n)
        if a is not None:
            v.set(a)
        return v

    def initial(self):
        v = self.vector()
        assert lib().Initial(0.0, v.h, self.h) == 0
        return v

    def exact(self, t):
        v = self.vector()
        assert lib().Exact(t, v.h, self.h) == 0
        return v

    def rhs(self, t, u, f):
        return lib().diffusion(t, u.h, f.h, self.h)

    def laplacian(self, u, f):
        return lib().laplacian(0.0, u.h, f.h, self.h)

    def psetup(self, gamma):
        jcur = i32()
        return lib().PSetup(0.0, None, None, 0, C.byref(jcur), gamma, self.h)

    def psolve(self, r, z):
        return lib().PSolve(0.0, None, None, r.h, z.h, 0.0, 0.0, 1, self.h)

    def mg_psetup(self, gamma):
        jcur = i32()
        return lib().MGPSetup(0.0, None, None, 0, C.byref(jcur), gamma, self.h)

    def mg_psolve(self, r, z):
        return lib().MGPSolve(0.0, None, None, r.h, z.h, 0.0, 0.0, 1, self.h)

    def mg_options(self, relax, nrelax):
        """V-cycle smoother: relax type 2 (symmetric) or 3, nrelax sweep pairs."""
        return lib().DiffusionProblem_SetMGOptions(self.h, relax, nrelax)

    def mg_cycles(self):
        return lib().DiffusionProblem_MGNumCycles(self.h)

    def dom_eig(self):
        lr, li = d(), d()
        lib().dom_eig(0.0, None, None, C.byref(lr), C.byref(li), self.h, None, None, None)
        return lr.value, li.value

    def set_rows(self, rows):
        return lib().DiffusionProblem_SetStencilRows(self.h, rows)

    def set_config(self, which, variant, rows, nt=0):
        return lib().DiffusionProblem_SetStencilConfig(self.h, which, variant, rows, nt)

    def stage_pass(self, K, coef, yjm1, yjm2, yn, fn, outkm1, outk):
        """K fused STS stages in one multi-stage pass; coef is K rows of
        {c0..c4, T1, T2, unused}."""
        import numpy as np
        cf = np.ascontiguousarray(np.asarray(coef, dtype=np.float64).reshape(-1))
        return lib().DiffusionProblem_StagePass(
            self.h, K, cf.ctypes.data_as(C.POINTER(C.c_double)), yjm1.h, yjm2.h, yn.h, fn.h,
            outkm1.h if outkm1 is not None else None, outk.h)

    def get_config(self, which):
        v, r, n = i32(), i32(), i32()
        lib().DiffusionProblem_GetStencilConfig(self.h, which, C.byref(v), C.byref(r), C.byref(n))
        return v.value, r.value, n.value

    def autotune(self, reps=3):
        return lib().DiffusionProblem_Autotune(self.h, reps)

    def destroy(self):
        if self.h and self.ctx.h:
            lib().DiffusionProblem_Destroy(self.h)
        self.h = None

    release = destroy

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Batch:
    """DiffusionBatch: many problems, one step group at a time."""

    def __init__(self, ctx, params_list, method="rkc"):
        self.ctx = ctx
        arr = (Params * len(params_list))()
        for k, kw in enumerate(params_list):
            p = Params()
            lib().DiffusionParams_Default(C.byref(p), kw.get("kind", MMS_DIRICHLET))
            for key, v in kw.items():
                setattr(p, key, v)
            arr[k] = p
        self.params = arr
        self.nprob = len(params_list)
        self.h = lib().DiffusionBatch_Create(self.nprob, arr, RKC if method == "rkc" else RKL, ctx.h)
        if not self.h:
            raise ValueError("DiffusionBatch_Create rejected the parameters")
        ctx.register(self)

    def __getattr__(self, name):
        f = getattr(lib(), "DiffusionBatch_" + name)
        return lambda *a: f(self.h, *a)

    def evolve(self, tout):
        return lib().DiffusionBatch_Evolve(self.h, tout)

    def solution(self, p):
        n = self.params[p].nx * self.params[p].ny
        v = Vector(self.ctx, n=n)
        assert lib().DiffusionBatch_GetSolution(self.h, p, v.h) == 0
        return v.get()

    def time(self, p):
        t = d()
        lib().DiffusionBatch_GetCurrentTime(self.h, p, C.byref(t))
        return t.value

    def stats(self, p):
        out = (C.c_long * 18)()
        lib().DiffusionBatch_GetStatsArray(self.h, p, out, 18)
        return dict(zip(STATS, list(out)))

    def groups(self):
        g, l = C.c_long(), C.c_long()
        lib().DiffusionBatch_GetNumGroups(self.h, C.byref(g), C.byref(l))
        return g.value, l.value

    def algorithmic_bytes(self):
        v = C.c_double()
        lib().DiffusionBatch_GetAlgorithmicBytes(self.h, C.byref(v))
        return v.value

    def release(self):
        if self.h and self.ctx.h:
            lib().DiffusionBatch_Destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def fn_ptr(name):
    return C.cast(getattr(lib(), name), vp)


def dptr_of(v):
    """Device address of a Vector's data."""
    return C.cast(lib().N_VGetDeviceArrayPointer(v.h), vp)


class Integrator:
    """LSRKStep (rkc/rkl) or implicit ARKStep (dirk) over a Problem."""

    def __init__(self, prob, method, y0, t0=0.0, ls="cg", maxl=20, precond=True):
        L = lib()
        self.prob = prob
        self.method = method
        self.ctrl = None
        self.dee = None
        self.LS = None
        rhs = fn_ptr("diffusion")
        if method in ("rkc", "rkl"):
            self.h = L.LSRKStepCreateSTS(rhs, t0, y0.h, prob.ctx.h)
            L.LSRKStepSetSTSMethod(self.h, RKC if method == "rkc" else RKL)
            L.LSRKStepSetDomEigFn(self.h, fn_ptr("dom_eig"))
        else:
            self.h = L.ARKStepCreate(None, rhs, t0, y0.h, prob.ctx.h)
            pre = SUN_PREC_RIGHT if precond else SUN_PREC_NONE
            make = {"cg": L.SUNLinSol_PCG, "gmres": L.SUNLinSol_SPGMR, "fgmres": L.SUNLinSol_SPFGMR,
                    "fcg": L.SUNLinSol_FPCG}[ls]
            self.LS = make(y0.h, pre, maxl, prob.ctx.h)
            L.ARKodeSetLinearSolver(self.h, self.LS, None)
            if precond == "learned":
                # network on the problem (DiffusionProblem_SetLearnedPrec), Jacobi-PCG fallback
                L.ARKodeSetPreconditioner(self.h, fn_ptr("LearnedPSetup"), fn_ptr("LearnedPSolve"))
                L.ARKodeSetLSFallbackPCG(self.h, 100000, fn_ptr("PSolve"), prob.h)
            elif precond == "mg":
                L.ARKodeSetPreconditioner(self.h, fn_ptr("MGPSetup"), fn_ptr("MGPSolve"))
            elif precond:
                L.ARKodeSetPreconditioner(self.h, fn_ptr("PSetup"), fn_ptr("PSolve"))
        if not self.h:
            raise RuntimeError("integrator creation failed")
        L.ARKodeSetUserData(self.h, prob.h)
        self.y = y0.clone()
        prob.ctx.register(self)

    def __getattr__(self, name):
        f = getattr(lib(), name)
        return lambda *a: f(self.h, *a)

    def internal_eig(self, maxiters=100, rtol=0.01):
        self.dee = lib().SUNDomEigEstimator_Power(self.y.h, maxiters, rtol, self.prob.ctx.h)
        return lib().LSRKStepSetDomEigEstimator(self.h, self.dee)

    def controller(self, kind):
        names = ["PID", "PI", "I", "ExpGus", "ImpGus", "ImExGus"]
        self.ctrl = getattr(lib(), "SUNAdaptController_" + names[kind])(self.prob.ctx.h)
        return lib().ARKodeSetAdaptController(self.h, self.ctrl)

    def evolve(self, tout, itask=ARK_NORMAL):
        t = d()
        flag = lib().ARKodeEvolve(self.h, tout, self.y.h, C.byref(t), itask)
        return flag, t.value

    def stats(self):
        out = (C.c_long * 18)()
        lib().ARKodeGetStatsArray(self.h, out, 18)
        return dict(zip(STATS, list(out)))

    def analytic_jtimes(self):
        """J v from the stencil (JTimes) instead of the difference quotient."""
        return lib().ARKodeSetJacTimes(self.h, None, fn_ptr("JTimes"))

    def fused_pcg(self, on=True):
        """PCG in fused kernels (two per iteration with Jacobi; needs
        analytic_jtimes and the Jacobi or multigrid preconditioner; otherwise
        the N_Vector PCG)."""
        return lib().ARKodeSetDiffusionFastPCG(self.h, self.prob.h if on else None)

    def fallbacks(self):
        n = C.c_long()
        lib().ARKodeGetNumLSFallbacks(self.h, C.byref(n))
        return n.value

    def spectral_radius(self):
        r = d()
        lib().LSRKStepGetSpectralRadius(self.h, C.byref(r))
        return r.value

    def release(self):
        self.free()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def free(self):
        L = lib()
        if not self.prob.ctx.h:
            self.h = self.LS = self.dee = self.ctrl = None
            return
        if self.h:
            h = vp(self.h)
            L.ARKodeFree(C.byref(h))
            self.h = None
        if self.LS:
            L.SUNLinSolFree(self.LS)
            self.LS = None
        if self.dee:
            h = vp(self.dee)
            L.SUNDomEigEstimator_Destroy(C.byref(h))
            self.dee = None
        if self.ctrl:
            L.SUNAdaptController_Destroy(self.ctrl)
            self.ctrl = None


class LearnedPrec:
    """The learned preconditioner network (learned_prec.h): 3x3 conv layers
    with the given dilations, 16 hidden channels, BF16 on the matrix cores.
    symmetric=True: the linear network K^T K (K the layers, applied and then
    its adjoint), symmetric positive semi-definite."""

    def __init__(self, ctx, dil=(1, 1, 1, 1), symmetric=False):
        d = (i32 * len(dil))(*dil)
        self.ctx = ctx
        self.dil = tuple(dil)
        self.symmetric = bool(symmetric)
        create = lib().LearnedPrec_CreateSymmetric if symmetric else lib().LearnedPrec_Create
        self.h = create(len(dil), d, ctx.h)
        if not self.h:
            raise ValueError("bad network shape")
        ctx.register(self)

    @property
    def nparams(self):
        return lib().LearnedPrec_NumParams(self.h)

    def params(self):
        import numpy as np
        p = np.zeros(self.nparams, dtype=np.float32)
        lib().LearnedPrec_GetParams(self.h, p.ctypes.data_as(C.POINTER(C.c_float)))
        return p

    def set_params(self, p):
        import numpy as np
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        return lib().LearnedPrec_SetParams(self.h, p.ctypes.data_as(C.POINTER(C.c_float)))

    def init_identity(self, noise=1e-2, seed=0):
        return lib().LearnedPrec_InitIdentity(self.h, noise, seed)

    def set_output_scale(self, s):
        return lib().LearnedPrec_SetOutputScale(self.h, s)

    def apply(self, nb, nx, ny, vin, vout, addto=0):
        return lib().LearnedPrec_Apply(self.h, nb, nx, ny, dptr_of(vin), dptr_of(vout), addto)

    def loss_grad(self, nb, nx, ny, ratio, r, e, grad=True, ratio_y=None):
        """Training loss and gradient on (I - ratio Dxx - ratio_y Dyy) e = r
        (ratio_y defaults to ratio: the isotropic operator)."""
        import numpy as np
        loss = C.c_double()
        g = np.zeros(self.nparams, dtype=np.float32) if grad else None
        ry = ratio if ratio_y is None else ratio_y
        rc = lib().LearnedPrec_LossGradXY(self.h, nb, nx, ny, ratio, ry, dptr_of(r), dptr_of(e), C.byref(loss),
                                          g.ctypes.data_as(C.POINTER(C.c_float)) if grad else None)
        assert rc == 0
        return loss.value, g

    def train(self, nsamp, n, ratio, r, e, steps, batch, lr, seed=0, ratio_y=None):
        """Adam on the pairs for (I - ratio Dxx - ratio_y Dyy) (ratio_y
        defaults to ratio); returns the loss of every step."""
        import numpy as np
        hist = np.zeros(steps)
        ry = ratio if ratio_y is None else ratio_y
        rc = lib().LearnedPrec_TrainXY(self.h, nsamp, n, ratio, ry, dptr_of(r), dptr_of(e), steps, batch, lr,
                                       seed, hist.ctypes.data_as(C.POINTER(C.c_double)))
        assert rc == 0
        return hist

    def attach(self, prob):
        return lib().DiffusionProblem_SetLearnedPrec(prob.h, self.h)

    def release(self):
        if getattr(self, "h", None) and self.ctx.h:
            lib().LearnedPrec_Destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def training_set(ctx, nsamp, n, ratio, reltol=1e-12, seed=1234, ratio_y=None):
    """(r, e) pairs on the GPU for (I - ratio Dxx - ratio_y Dyy) e = r
    (ratio_y defaults to ratio): two Vectors of nsamp*n*n values and the
    largest Jacobi-PCG iteration count."""
    r = Vector(ctx, n=nsamp * n * n)
    e = Vector(ctx, n=nsamp * n * n)
    it = i32()
    ry = ratio if ratio_y is None else ratio_y
    rc = lib().LearnedPrec_MakeTrainingSetXY(nsamp, n, ratio, ry, reltol, seed, dptr_of(r), dptr_of(e),
                                            C.byref(it), ctx.h)
    return r, e, it.value, rc
