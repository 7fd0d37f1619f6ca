/* ---------------------------------------------------------------------------
 * diffusion_2D problem layer on one MI355X.
 *
 * DiffusionProblem replaces the reference's UserData (diffusion_2D.hpp:66-211)
 * for one GPU: the grid, coefficients and the Jacobi workspace live in HBM,
 * and the one-rank periodic halo exchange (diffusion_2D.cpp:400-584, every
 * neighbour is the rank itself) is folded into the stencil's index math.
 *
 * Entry points keep the reference's names and callback signatures:
 *   diffusion  — ARKRhsFn, diffusion_2D.cpp:23-35 (diffusion_2D.hpp:250)
 *   laplacian  — the stencil, diffusion.cpp:9-209 (diffusion_2D.hpp:247)
 *   PSetup     — ARKLsPrecSetupFn, preconditioner_jacobi.cpp:9-44
 *   PSolve     — ARKLsPrecSolveFn, preconditioner_jacobi.cpp:47-62
 *   Initial    — initial.cpp:20-48
 *   dom_eig    — ARKDomEigFn, main.cpp:536-550 (static there)
 * Return convention as in the reference: 0 success, -1 unrecoverable (a HIP
 * launch error or an argument mismatch is reported as -1, never as >0).
 * Two problems are provided: DIFFUSION_CEDA_PERIODIC (the reference code's
 * periodic Gaussian problem) and DIFFUSION_MMS_DIRICHLET (the manufactured
 * problem of diffusion_2D/README.md:11-32, interior nodes only).
 * ------------------------------------------------------------------------- */
#ifndef DIFFUSION_HIP_H
#define DIFFUSION_HIP_H

#include "ceda_sundials.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { DIFFUSION_CEDA_PERIODIC = 0, DIFFUSION_MMS_DIRICHLET = 1 };

typedef struct {
  int kind;
  int64_t nx, ny;           /* nodes (periodic) or interior nodes (MMS) */
  double xl, xu, yl, yu;    /* domain (diffusion_2D.hpp:79-82 defaults) */
  double kx, ky;            /* diffusion coefficients                    */
  int inhomogeneous;        /* k(1+0.99 sin s) (diffusion_2D.cpp:887-897) */
  double tf;                /* final time                                */
} DiffusionParams;

typedef struct DiffusionProblem DiffusionProblem;

/* reference defaults (diffusion_2D.hpp:66-93) or README.md defaults (MMS) */
void DiffusionParams_Default(DiffusionParams* p, int kind);
/* allocates the HBM tables and the Jacobi diagonal; NULL on bad input */
DiffusionProblem* DiffusionProblem_Create(const DiffusionParams* p, SUNContext ctx);
void DiffusionProblem_Destroy(DiffusionProblem* prob);
void DiffusionProblem_GetGrid(const DiffusionProblem* prob, int64_t* nx, int64_t* ny,
                              double* dx, double* dy);
long DiffusionProblem_GetNumRhsEvals(const DiffusionProblem* prob);
/* stencil launch shape of the row-interleaved variant for both kernels:
   rows per wavefront (1,2,4,8; 0 = 1) */
int DiffusionProblem_SetStencilRows(DiffusionProblem* prob, int rows);
/* launch shape of one kernel kind (which: 0 RHS stencil, 1 fused STS stage):
   variant 0 = row-interleaved tiles (rows per wavefront), variant 1 =
   y-marching (rows = consecutive rows per wavefront, nt = non-temporal
   streamed operands).  which = 2 is the multi-stage pass of the fused path:
   variant = stages per pass K (2..6; 0 = one kernel per stage), rows = output
   rows per wavefront tile (1..1024), nt a bit mask: 1 non-temporal streamed
   operands, 2 the stage coefficients read from LDS at every row step instead
   of held in scalar registers.  Every shape computes bitwise-identical
   results. */
int DiffusionProblem_SetStencilConfig(DiffusionProblem* prob, int which, int variant, int rows,
                                      int nt);
int DiffusionProblem_GetStencilConfig(const DiffusionProblem* prob, int which, int* variant,
                                      int* rows, int* nt);
/* times every candidate launch shape of the three kernels on scratch buffers
   of the problem's size (reps launches each, HIP events) and keeps the
   fastest; multi-stage passes are enabled only when their time per stage
   beats the best single-stage kernel.  Results are cached per (nx, ny, kind)
   for the life of the process */
int DiffusionProblem_Autotune(DiffusionProblem* prob, int reps);

/* [ext] times `reps` back-to-back launches of one hot-path kernel with HIP
   events on the problem's stream: which = 0 the RHS stencil (u -> out),
   which = 1 the fused STS stage (u = Y_{j-1}, a = Y_{j-2}, b = Y_0, c = F_0),
   which = 2 the configured multi-stage pass (same operands; the last stage
   goes to out).  Writes the mean launch duration in milliseconds. */
int DiffusionProblem_TimeKernel(DiffusionProblem* prob, int which, N_Vector u, N_Vector a,
                                N_Vector b, N_Vector c, N_Vector out, int reps,
                                double* avg_ms);
/* [ext] K consecutive STS stages (K = 1..6) in one pass of the multi-stage
   kernel, with the configured rows/nt of which = 2: stage m (0-based) uses
   coef[8m .. 8m+4] as {c0..c4} and coef[8m+5], coef[8m+6] as the MMS time
   factors (T1, T2) of its RHS; reads yjm1 = Y_{j-1}, yjm2 = Y_{j-2}, yn, fn;
   writes stage K-1 to outkm1 (K >= 2; ignored for K = 1) and stage K to
   outk.  The outputs must not alias an input.  Bitwise equal to K fused
   single-stage launches.  Returns 0, or -1 for invalid arguments. */
int DiffusionProblem_StagePass(DiffusionProblem* prob, int K, const double* coef, N_Vector yjm1,
                               N_Vector yjm2, N_Vector yn, N_Vector fn, N_Vector outkm1,
                               N_Vector outk);

int Initial(sunrealtype t, N_Vector u, void* user_data);
int Exact(sunrealtype t, N_Vector u, void* user_data);
int laplacian(sunrealtype t, N_Vector u, N_Vector f, void* user_data);
/* [ext] geometric-multigrid preconditioner, the GPU counterpart of the
   reference's hypre PFMG option (preconditioner_hypre.cpp:51-160, settings
   diffusion_2D.hpp:187-193: one V-cycle, zero initial guess, symmetric
   red/black Gauss-Seidel, 2 pre + 2 post sweeps), same signatures and user
   data as PSetup/PSolve.  Both problems: the manufactured Dirichlet one
   (coarse node I = fine node 2I+1) and the reference's periodic one (even
   sides, coarse node I = fine node 2I, everything wrapping; -1 for an odd
   side).  Coarsening n -> n/2, rediscretised coarse operators, full
   weighting, bilinear prolongation; with the default smoother the cycle is
   symmetric for a symmetric operator, so it preconditions PCG. */
int MGPSetup(sunrealtype t, N_Vector u, N_Vector f, sunbooleantype jok, sunbooleantype* jcurPtr,
             sunrealtype gamma, void* user_data);
int MGPSolve(sunrealtype t, N_Vector u, N_Vector f, N_Vector r, N_Vector z, sunrealtype gamma,
             sunrealtype delta, int lr, void* user_data);
/* levels of the hierarchy built by MGPSetup (0 before the first setup) */
int DiffusionProblem_MGLevels(const DiffusionProblem* prob);
/* [ext] V-cycle smoother of MGPSetup/MGPSolve, the reference's --pfmg_relax
   and --pfmg_nrelax (diffusion_2D.cpp:137-151, preconditioner_hypre.cpp:91-98):
   relax 2 = symmetric red/black Gauss-Seidel (red, black before the coarse
   correction; black, red after), 3 = nonsymmetric (red, black both times);
   nrelax pre- and post-sweep pairs (0..16).  Default 2, 2.  Takes effect at
   the next MGPSetup.  0, or -1 for other values. */
int DiffusionProblem_SetMGOptions(DiffusionProblem* prob, int relax, int nrelax);
/* V-cycles applied so far (MGPSolve calls): hypre's PFMG iteration count,
   printed by main.cpp:489-492 as "Total PFMG iterations" */
long DiffusionProblem_MGNumCycles(const DiffusionProblem* prob);
/* [ext] ARKLsJacTimesVecFn: J v of the RHS, exactly the 5-point operator of
   laplacian() (diffusion.cpp:9-209; the manufactured forcing does not depend
   on u), for ARKodeSetJacTimes; the reference uses arkLs's difference
   quotient instead (main.cpp:302-329) */
int JTimes(N_Vector v, N_Vector Jv, sunrealtype t, N_Vector y, N_Vector fy, void* user_data,
           N_Vector tmp);
int diffusion(sunrealtype t, N_Vector u, N_Vector f, void* user_data);
int PSetup(sunrealtype t, N_Vector u, N_Vector f, sunbooleantype jok,
           sunbooleantype* jcurPtr, sunrealtype gamma, void* user_data);
int PSolve(sunrealtype t, N_Vector u, N_Vector f, N_Vector r, N_Vector z,
           sunrealtype gamma, sunrealtype delta, int lr, void* user_data);
int dom_eig(sunrealtype t, N_Vector y, N_Vector fn, sunrealtype* lambdaR,
            sunrealtype* lambdaI, void* user_data, N_Vector temp1,
            N_Vector temp2, N_Vector temp3);

#ifdef __cplusplus
}
#endif
#endif
