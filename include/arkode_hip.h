/* ---------------------------------------------------------------------------
 * Integrator layer: the SUNDIALS ARKODE subset the reference's driver
 * configures (diffusion_2D/main.cpp:176-400), restated for MI355X.
 *
 * Every function below keeps the SUNDIALS name, signature and return
 * convention of the call site it replaces in the reference, so main.cpp's
 * sequence of calls runs unchanged against this library:
 *   LSRKStepCreateSTS                     main.cpp:257
 *   LSRKStepSetSTSMethod                  main.cpp:336-339
 *   SUNDomEigEstimator_Power              main.cpp:344
 *   LSRKStepSetDomEigEstimator            main.cpp:347
 *   LSRKStepSetDomEigFn                   main.cpp:352
 *   ARKStepCreate                         main.cpp:242
 *   ARKodeSStolerances / SetUserData      main.cpp:262-267
 *   ARKodeSetOrder                        main.cpp:272
 *   SUNLinSol_PCG / SUNLinSol_SPGMR       main.cpp:211,216
 *   ARKodeSetLinearSolver                 main.cpp:305
 *   ARKodeSetPreconditioner               main.cpp:311
 *   ARKodeSetLSetupFrequency              main.cpp:315
 *   ARKodeSetEpsLin                       main.cpp:320
 *   ARKodeSetLinear                       main.cpp:326
 *   ARKodeSetFixedStep                    main.cpp:360
 *   SUNAdaptController_{PID,...}          main.cpp:367-372
 *   ARKodeSetAdaptController              main.cpp:374
 *   ARKodeSetMaxNumSteps / SetStopTime    main.cpp:380,384
 *   ARKodeEvolve                          main.cpp:442
 *   ARKodePrintAllStats                   main.cpp:486
 *   ARKodeFree / SUNLinSolFree            main.cpp:506,509
 * (and LSRKStepSetDomEigFrequency, LSRKStepSetMaxNumStages,
 * LSRKStepSetDomEigSafetyFactor, LSRKStepSetNumDomEigEst*PreprocessIters as
 * called at gkeyll_diffusion/gk_diffusion_1x1v_p1.c:873-906.)
 * Stop-time semantics: ARKodeEvolve in ARK_NORMAL mode steps exactly to
 * min(tout, tstop) (no dense-output interpolation).
 *
 * Extensions with no reference counterpart are marked [ext].
 * ------------------------------------------------------------------------- */
#ifndef ARKODE_HIP_H
#define ARKODE_HIP_H

#include "ceda_sundials.h"
#include "diffusion_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct SUNDomEigEstimator_* SUNDomEigEstimator;
typedef struct SUNLinearSolver_* SUNLinearSolver;
typedef struct SUNAdaptController_* SUNAdaptController;
typedef struct SUNMatrix_* SUNMatrix;

typedef enum {
  ARKODE_LSRK_RKC_2 = 0,
  ARKODE_LSRK_RKL_2 = 1
} ARKODE_LSRKMethodType;

#define ARK_NORMAL 1
#define ARK_ONE_STEP 2

#define ARK_SUCCESS 0
#define ARK_TSTOP_RETURN 1
#define ARK_TOO_MUCH_WORK (-1)
#define ARK_ERR_FAILURE (-3)
#define ARK_CONV_FAILURE (-4)
#define ARK_LSETUP_FAIL (-6)
#define ARK_LSOLVE_FAIL (-7)
#define ARK_RHSFUNC_FAIL (-8)
#define ARK_MEM_NULL (-21)
#define ARK_ILL_INPUT (-22)
#define ARK_VECTOROP_ERR (-28)
#define ARK_MAX_STAGE_LIMIT_FAIL (-51)

#define ARK_ADAPT_PID 0
#define ARK_ADAPT_PI 1
#define ARK_ADAPT_I 2
#define ARK_ADAPT_EXP_GUS 3
#define ARK_ADAPT_IMP_GUS 4
#define ARK_ADAPT_IMEX_GUS 5

#define SUN_PREC_NONE 0
#define SUN_PREC_LEFT 1
#define SUN_PREC_RIGHT 2
#define SUN_PREC_BOTH 3

#define SUN_OUTPUTFORMAT_TABLE 0
#define SUN_OUTPUTFORMAT_CSV 1

/* ---- creation ---- */
void* LSRKStepCreateSTS(ARKRhsFn rhs, sunrealtype t0, N_Vector y0, SUNContext sunctx);
/* implicit-only: fe must be NULL (the reference's DIRK path, main.cpp:242) */
void* ARKStepCreate(ARKRhsFn fe, ARKRhsFn fi, sunrealtype t0, N_Vector y0, SUNContext sunctx);
void ARKodeFree(void** arkode_mem);

/* ---- LSRKStep ---- */
int LSRKStepSetSTSMethod(void* arkode_mem, ARKODE_LSRKMethodType method);
int LSRKStepSetDomEigFn(void* arkode_mem, ARKDomEigFn dom_eig);
int LSRKStepSetDomEigEstimator(void* arkode_mem, SUNDomEigEstimator DEE);
int LSRKStepSetDomEigFrequency(void* arkode_mem, long nsteps);
int LSRKStepSetMaxNumStages(void* arkode_mem, int stage_max_limit);
int LSRKStepSetDomEigSafetyFactor(void* arkode_mem, sunrealtype dom_eig_safety);
int LSRKStepSetNumDomEigEstInitPreprocessIters(void* arkode_mem, int num_iters);
int LSRKStepSetNumDomEigEstPreprocessIters(void* arkode_mem, int num_iters);
int LSRKStepGetNumDomEigUpdates(void* arkode_mem, long* num_dom_eig_updates);
int LSRKStepGetMaxNumStages(void* arkode_mem, int* stage_max);
/* [ext] fused stencil+stage kernels for a DiffusionProblem RHS; bitwise equal
   to the N_Vector path; use_graphs replays each stage loop as a HIP graph */
int LSRKStepSetDiffusionFastPath(void* arkode_mem, DiffusionProblem* prob, int use_graphs);
/* [ext] PCG (SUNLinSol_PCG, main.cpp:305) in fused kernels for a
   DiffusionProblem that is also the user data, when J v is the analytic
   JTimes (otherwise the N_Vector PCG runs).  With Jacobi (PSolve) or no
   preconditioner an iteration is two kernels with the scalars kept on the
   device (the host checks the status every few iterations); any other
   preconditioner is called through its callback between iterations.  Same algorithm and stopping test as
   the N_Vector PCG, not bitwise equal to it (the dot products are summed in
   another order).  prob = NULL switches it off. */
int ARKodeSetDiffusionFastPCG(void* arkode_mem, DiffusionProblem* prob);
/* [ext] with a preconditioner callback, the fused PCG enqueues this fraction
   of the previous solve's iteration count without testing the residual
   (the host waits only every few iterations), then tests it before every
   callback (default 0.75; 0..1, ARK_ILL_INPUT otherwise).  A larger
   fraction waits less but may run the callback for iterations past
   convergence. */
int ARKodeSetFusedPCGBulk(void* arkode_mem, sunrealtype fraction);

SUNDomEigEstimator SUNDomEigEstimator_Power(N_Vector q, long max_iters, sunrealtype rel_tol,
                                            SUNContext sunctx);
int SUNDomEigEstimator_Destroy(SUNDomEigEstimator* DEE);

/* ---- common ARKODE ---- */
int ARKodeSStolerances(void* arkode_mem, sunrealtype reltol, sunrealtype abstol);
int ARKodeSetUserData(void* arkode_mem, void* user_data);
int ARKodeSetOrder(void* arkode_mem, int maxord);
int ARKodeSetFixedStep(void* arkode_mem, sunrealtype hfixed);
int ARKodeSetInitStep(void* arkode_mem, sunrealtype hin);
int ARKodeSetAdaptController(void* arkode_mem, SUNAdaptController C);
int ARKodeSetMaxNumSteps(void* arkode_mem, long mxsteps);
int ARKodeSetStopTime(void* arkode_mem, sunrealtype tstop);
int ARKodeEvolve(void* arkode_mem, sunrealtype tout, N_Vector yout, sunrealtype* tret, int itask);
int ARKodePrintAllStats(void* arkode_mem, FILE* outfile, int fmt);
int ARKodeGetNumSteps(void* arkode_mem, long* nsteps);
int ARKodeGetNumStepAttempts(void* arkode_mem, long* step_attempts);
int ARKodeGetNumErrTestFails(void* arkode_mem, long* netfails);
int ARKodeGetNumRhsEvals(void* arkode_mem, int partition_index, long* num_rhs_evals);
int ARKodeGetLastStep(void* arkode_mem, sunrealtype* hlast);
int ARKodeGetCurrentStep(void* arkode_mem, sunrealtype* hcur);
int ARKodeGetCurrentTime(void* arkode_mem, sunrealtype* tcur);
/* [ext] the oracle-comparable statistics vector (same layout as the CPU
   restatement's orc_get_stats) */
int ARKodeGetStatsArray(void* arkode_mem, long* out, int nout);
/* [ext] spectral radius used by the last LSRK step */
int LSRKStepGetSpectralRadius(void* arkode_mem, sunrealtype* rho);

/* ---- implicit solve ---- */
int ARKodeSetLinearSolver(void* arkode_mem, SUNLinearSolver LS, SUNMatrix A);
int ARKodeSetPreconditioner(void* arkode_mem, ARKLsPrecSetupFn psetup, ARKLsPrecSolveFn psolve);
/* SUNDIALS ARKodeSetJacTimes: user J*v instead of the difference quotient
   (the reference sets none, main.cpp:302-329); jtsetup may be NULL */
int ARKodeSetJacTimes(void* arkode_mem, ARKLsJacTimesSetupFn jtsetup, ARKLsJacTimesVecFn jtimes);
int ARKodeSetLSetupFrequency(void* arkode_mem, int msbp);
int ARKodeSetEpsLin(void* arkode_mem, sunrealtype eplifac);
int ARKodeSetLinear(void* arkode_mem, int timedepend);
int ARKodeSetNonlinear(void* arkode_mem);
/* [ext] linear solves stop at ||r||_2 <= reltol ||b||_2 instead of the
   ARKODE weighted-norm test (used by the fixed-step implicit benchmark);
   0 restores the weighted test; ARK_ILL_INPUT unless 0 <= reltol < 1 */
int ARKodeSetLSRelativeTolerance(void* arkode_mem, sunrealtype reltol);
/* [ext] fallback of a (F)GMRES linear solve: when it stops at its iteration
   cap short of the tolerance, the same system is solved again by PCG with at
   most maxl iterations and the preconditioner solve psolve(user_data) (for
   example the Jacobi PSolve of preconditioner_jacobi.cpp:44-62, whose setup
   the primary preconditioner setup must have done); maxl = 0 removes it.  No
   reference counterpart (the reference uses a single Krylov method). */
int ARKodeSetLSFallbackPCG(void* arkode_mem, int maxl, ARKLsPrecSolveFn psolve, void* user_data);
int ARKodeGetNumLSFallbacks(void* arkode_mem, long* nfallbacks);
int ARKodeGetNumLinIters(void* arkode_mem, long* nliters);

SUNLinearSolver SUNLinSol_PCG(N_Vector y, int pretype, int maxl, SUNContext sunctx);
/* [ext] flexible PCG (Polak-Ribiere beta) for preconditioners that vary from
   one application to the next, such as the learned network; no reference
   counterpart.  With a fixed linear preconditioner it takes the PCG steps up
   to rounding.  The (F)GMRES fallback of ARKodeSetLSFallbackPCG applies to it
   too. */
SUNLinearSolver SUNLinSol_FPCG(N_Vector y, int pretype, int maxl, SUNContext sunctx);
SUNLinearSolver SUNLinSol_SPGMR(N_Vector y, int pretype, int maxl, SUNContext sunctx);
SUNLinearSolver SUNLinSol_SPFGMR(N_Vector y, int pretype, int maxl, SUNContext sunctx);
/* [ext] Krylov restart length of SPGMR/SPFGMR (default 30) */
int SUNLinSol_SPFGMRSetMaxRestartDim(SUNLinearSolver S, int mrestart);
int SUNLinSolFree(SUNLinearSolver S);

/* ---- step-size controllers ---- */
SUNAdaptController SUNAdaptController_PID(SUNContext sunctx);
SUNAdaptController SUNAdaptController_PI(SUNContext sunctx);
SUNAdaptController SUNAdaptController_I(SUNContext sunctx);
SUNAdaptController SUNAdaptController_ExpGus(SUNContext sunctx);
SUNAdaptController SUNAdaptController_ImpGus(SUNContext sunctx);
SUNAdaptController SUNAdaptController_ImExGus(SUNContext sunctx);
SUNErrCode SUNAdaptController_Destroy(SUNAdaptController C);

#ifdef __cplusplus
}
#endif
#endif
