/* ---------------------------------------------------------------------------
 * [ext] Batched super-time-stepping: many independent diffusion problems of
 * uneven sizes and coefficients on one GPU, each with its own adaptive step
 * size and stage count.
 *
 * There is no reference counterpart (the reference integrates one problem
 * per run, diffusion_2D/main.cpp:431-470); each problem of a batch takes
 * exactly the steps LSRKStepCreateSTS + ARKodeEvolve would take for it alone
 * (same controller, stage-count and spectral-radius rules, closed-form
 * dom_eig of main.cpp:536-550), but the stages of all problems run together:
 * one kernel launch per stage index covers every problem still in its stage
 * loop, one segmented reduction gives all error norms, and the host makes one
 * device-to-host copy per step group.
 * ------------------------------------------------------------------------- */
#ifndef BATCH_HIP_H
#define BATCH_HIP_H

#include "arkode_hip.h"
#include "diffusion_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct DiffusionBatch DiffusionBatch;

/* nprob problems (params[k]); method ARKODE_LSRK_RKC_2 or ARKODE_LSRK_RKL_2;
   every problem starts from Initial() at t = 0 */
DiffusionBatch* DiffusionBatch_Create(int nprob, const DiffusionParams* params,
                                      ARKODE_LSRKMethodType method, SUNContext ctx);
void DiffusionBatch_Destroy(DiffusionBatch* B);
int DiffusionBatch_SetTolerances(DiffusionBatch* B, sunrealtype rtol, sunrealtype atol);
int DiffusionBatch_SetController(DiffusionBatch* B, int controller);    /* ARK_ADAPT_* */
int DiffusionBatch_SetMaxNumStages(DiffusionBatch* B, int smax);
int DiffusionBatch_SetMaxNumSteps(DiffusionBatch* B, long mxsteps);
int DiffusionBatch_SetFixedStep(DiffusionBatch* B, sunrealtype h);
/* tile height of the batched kernels: each wavefront tile is `rows` rows of a
   124-column strip (62 for odd nx); rows 1..256, 0 = choose per size class by
   the tile autotuner */
int DiffusionBatch_SetTileRows(DiffusionBatch* B, int rows);
/* thin phases: a phase launch that would have fewer than min_tiles wavefront
   tiles (late in a step group, when only the problems with the most stages
   are left) runs on tiles of at most `rows` rows instead (shorter column
   marches over more wavefronts); rows 1..256, 0 = never; defaults 12 and
   2048.  The results do not depend on it. */
int DiffusionBatch_SetThinPhases(DiffusionBatch* B, int rows, int min_tiles);
/* stages per pass of the batched multi-stage kernel (1..5, default 5): each
   phase launch advances every active problem by up to this many stages */
int DiffusionBatch_SetStagesPerPass(DiffusionBatch* B, int kpass);
/* advance every problem to tout; returns 0, or the first failing problem's
   ARKODE code */
int DiffusionBatch_Evolve(DiffusionBatch* B, sunrealtype tout);
/* copy problem p's current state into y (a HIP-native vector of its length) */
int DiffusionBatch_GetSolution(DiffusionBatch* B, int p, N_Vector y);
int DiffusionBatch_GetCurrentTime(DiffusionBatch* B, int p, sunrealtype* t);
/* same layout as ARKodeGetStatsArray */
int DiffusionBatch_GetStatsArray(DiffusionBatch* B, int p, long* out, int nout);
/* step groups and stage-phase launches issued so far */
int DiffusionBatch_GetNumGroups(DiffusionBatch* B, long* ngroups, long* nlaunches);
/* algorithmic HBM bytes the phase kernels have moved so far (FP64 cells:
   24 B per cell for a first stage, 48 for a pass of up to kpass stages, 16
   for the closing RHS, 32 for the error norm), for roofline accounting */
int DiffusionBatch_GetAlgorithmicBytes(DiffusionBatch* B, double* bytes);

#ifdef __cplusplus
}
#endif
#endif
