// diffusion_2D driver on one MI355X.
//
// Mirrors diffusion_2D/main.cpp: the same flags (UserData::parse_args
// diffusion_2D.cpp:42-163, UserOptions::parse_args main.cpp:556-687,
// UserOutput::parse_args diffusion_2D.cpp:651-677), the same sequence of
// integrator calls (main.cpp:176-400), the output loop with
// ||u||_rms = sqrt(u.u/nx/ny) (main.cpp:431-470, diffusion_2D.cpp:801), the
// tight-tolerance reference run of --error (main.cpp:387-465) and the final
// lines that runtests-diffusion2d.py:36-54 parses.  Added flags:
//   --problem ceda|mms   manufactured Dirichlet problem of README.md:11-32
//   --deefreq N --stagemax N --eigsafety X   LSRKStep settings
//   --fastpath 0|1|2     0: N_Vector API path, 1: fused stage kernels,
//                        2: fused kernels replayed as HIP graphs
//   --ls cg|gmres|fgmres --ls_reltol X --write_final FILE
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "arkode_hip.h"
#include "diffusion_hip.h"
#include "nvector_hipnative.h"

static double wall()
{
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int main(int argc, char** argv)
{
  int kind = DIFFUSION_CEDA_PERIODIC;
  for (int a = 1; a < argc - 1; a++)
    if (!strcmp(argv[a], "--problem"))
      kind = !strcmp(argv[a + 1], "mms") ? DIFFUSION_MMS_DIRICHLET : DIFFUSION_CEDA_PERIODIC;
  DiffusionParams prm;
  DiffusionParams_Default(&prm, kind);
  std::string integrator = "dirk", ls = "cg";
  double rtol = 1e-5, atol = 1e-10, hfixed = 0.0, epslin = 0.0, eigsafety = 1.01, ls_reltol = 0.0;
  int order = 2, controller = 0, maxsteps = 0, error = 0, linear = 1, precond = 1;
  int liniters = 20, msbp = 0, internaleig = 0, output = 1, nout = 20, fastpath = 0;
  long deefreq = 25;
  int stagemax = 200;
  const char* write_final = nullptr;
  for (int a = 1; a < argc; a++) {
    const char* k = argv[a];
    const char* v = (a + 1 < argc) ? argv[a + 1] : nullptr;
#define TAKE(name) (!strcmp(k, name) && v && (a++, true))
    if (TAKE("--problem")) continue;
    else if (TAKE("--nx")) prm.nx = strtol(v, nullptr, 10);
    else if (TAKE("--ny")) prm.ny = strtol(v, nullptr, 10);
    else if (TAKE("--xl")) prm.xl = atof(v);
    else if (TAKE("--xu")) prm.xu = atof(v);
    else if (TAKE("--yl")) prm.yu = atof(v);  // as written at diffusion_2D.cpp:95-100
    else if (TAKE("--yu")) prm.yu = atof(v);
    else if (TAKE("--kx")) prm.kx = atof(v);
    else if (TAKE("--ky")) prm.ky = atof(v);
    else if (!strcmp(k, "--inhomogeneous")) prm.inhomogeneous = 1;
    else if (TAKE("--tf")) prm.tf = atof(v);
    else if (TAKE("--npx") || TAKE("--npy")) continue;
    else if (TAKE("--integrator")) integrator = v;
    else if (TAKE("--rtol")) rtol = atof(v);
    else if (TAKE("--atol")) atol = atof(v);
    else if (TAKE("--fixedstep")) hfixed = atof(v);
    else if (TAKE("--order")) order = atoi(v);
    else if (TAKE("--controller")) controller = atoi(v);
    else if (TAKE("--maxsteps")) maxsteps = atoi(v);
    else if (!strcmp(k, "--error")) error = 1;
    else if (!strcmp(k, "--nonlinear")) linear = 0;
    else if (TAKE("--ls")) ls = v;
    else if (!strcmp(k, "--lsinfo")) continue;
    else if (TAKE("--liniters")) liniters = atoi(v);
    else if (TAKE("--msbp")) msbp = atoi(v);
    else if (TAKE("--epslin")) epslin = atof(v);
    else if (!strcmp(k, "--noprec")) precond = 0;
    else if (!strcmp(k, "--internaleig")) internaleig = 1;
    else if (TAKE("--output")) output = atoi(v);
    else if (TAKE("--nout")) nout = atoi(v);
    else if (TAKE("--deefreq")) deefreq = strtol(v, nullptr, 10);
    else if (TAKE("--stagemax")) stagemax = atoi(v);
    else if (TAKE("--eigsafety")) eigsafety = atof(v);
    else if (TAKE("--ls_reltol")) ls_reltol = atof(v);
    else if (TAKE("--fastpath")) fastpath = atoi(v);
    else if (TAKE("--write_final")) write_final = v;
    else {
      fprintf(stderr, "ERROR: Unknown inputs: %s\n", k);
      return 1;
    }
#undef TAKE
  }
  const bool impl = (integrator == "dirk");
  const bool sts = (integrator == "rkc" || integrator == "rkl");
  if (!impl && !sts) {
    fprintf(stderr, "ERROR: illegal integrator\n");
    return 1;
  }

  SUNContext ctx = nullptr;
  if (SUNContext_Create(SUN_COMM_NULL, &ctx)) return 1;
  DiffusionProblem* udata = DiffusionProblem_Create(&prm, ctx);
  if (!udata) {
    fprintf(stderr, "ERROR: problem setup\n");
    return 1;
  }
  int64_t nx, ny;
  double dx, dy;
  DiffusionProblem_GetGrid(udata, &nx, &ny, &dx, &dy);
  printf("\n Problem options:\n");
  printf("  problem        = %s\n", kind == DIFFUSION_MMS_DIRICHLET ? "mms" : "ceda");
  printf("  kx             = %g\n  ky             = %g\n", prm.kx, prm.ky);
  printf("  %s diffusion\n", prm.inhomogeneous ? "spatially-varying" : "homogeneous");
  printf("  tf             = %g\n", prm.tf);
  printf("  nx             = %ld\n  ny             = %ld\n", (long)nx, (long)ny);
  printf("  dx             = %.16g\n  dy             = %.16g\n", dx, dy);
  printf(" Integrator: %s\n  rtol        = %g\n  atol        = %g\n", integrator.c_str(), rtol, atol);

  N_Vector u = N_VNew_HipNative(nx * ny, ctx);
  if (Initial(0.0, u, udata)) return 1;
  N_Vector uref = nullptr, uerr = nullptr;
  if (error) {
    uref = N_VClone(u);
    uerr = N_VClone(u);
    N_VScale(1.0, u, uref);
    N_VConst(0.0, uerr);
  }

  SUNLinearSolver LS = nullptr;
  if (impl) {
    const int prectype = precond ? SUN_PREC_RIGHT : SUN_PREC_NONE;
    if (ls == "cg") LS = SUNLinSol_PCG(u, prectype, liniters, ctx);
    else if (ls == "gmres") LS = SUNLinSol_SPGMR(u, prectype, liniters, ctx);
    else LS = SUNLinSol_SPFGMR(u, prectype, liniters, ctx);
  }

  void* arkode_mem = impl ? ARKStepCreate(nullptr, diffusion, 0.0, u, ctx)
                          : LSRKStepCreateSTS(diffusion, 0.0, u, ctx);
  if (!arkode_mem) return 1;
  ARKodeSStolerances(arkode_mem, rtol, atol);
  ARKodeSetUserData(arkode_mem, udata);
  if (impl) {
    ARKodeSetOrder(arkode_mem, order);
    ARKodeSetLinearSolver(arkode_mem, LS, nullptr);
    if (precond) {
      ARKodeSetPreconditioner(arkode_mem, PSetup, PSolve);
      ARKodeSetLSetupFrequency(arkode_mem, msbp);
    }
    ARKodeSetEpsLin(arkode_mem, epslin);
    if (linear) ARKodeSetLinear(arkode_mem, 0);
    else ARKodeSetNonlinear(arkode_mem);
    if (ls_reltol > 0.0) ARKodeSetLSRelativeTolerance(arkode_mem, ls_reltol);
  }
  SUNDomEigEstimator DEE = nullptr;
  if (sts) {
    LSRKStepSetSTSMethod(arkode_mem, integrator == "rkc" ? ARKODE_LSRK_RKC_2 : ARKODE_LSRK_RKL_2);
    if (internaleig) {
      DEE = SUNDomEigEstimator_Power(u, 100, 0.01, ctx);
      LSRKStepSetDomEigEstimator(arkode_mem, DEE);
    } else LSRKStepSetDomEigFn(arkode_mem, dom_eig);
    LSRKStepSetDomEigFrequency(arkode_mem, deefreq);
    LSRKStepSetMaxNumStages(arkode_mem, stagemax);
    LSRKStepSetDomEigSafetyFactor(arkode_mem, eigsafety);
    if (fastpath > 0) LSRKStepSetDiffusionFastPath(arkode_mem, udata, fastpath > 1);
  }
  SUNAdaptController C = nullptr;
  if (hfixed > 0.0) ARKodeSetFixedStep(arkode_mem, hfixed);
  else {
    switch (controller) {
    case ARK_ADAPT_PID: C = SUNAdaptController_PID(ctx); break;
    case ARK_ADAPT_PI: C = SUNAdaptController_PI(ctx); break;
    case ARK_ADAPT_I: C = SUNAdaptController_I(ctx); break;
    case ARK_ADAPT_EXP_GUS: C = SUNAdaptController_ExpGus(ctx); break;
    case ARK_ADAPT_IMP_GUS: C = SUNAdaptController_ImpGus(ctx); break;
    case ARK_ADAPT_IMEX_GUS: C = SUNAdaptController_ImExGus(ctx); break;
    }
    ARKodeSetAdaptController(arkode_mem, C);
  }
  ARKodeSetMaxNumSteps(arkode_mem, maxsteps);
  ARKodeSetStopTime(arkode_mem, prm.tf);

  void* arkref_mem = nullptr;
  if (error) {
    arkref_mem = LSRKStepCreateSTS(diffusion, 0.0, u, ctx);
    ARKodeSStolerances(arkref_mem, 1.e-12, atol);
    ARKodeSetUserData(arkref_mem, udata);
    LSRKStepSetDomEigFn(arkref_mem, dom_eig);
    ARKodeSetMaxNumSteps(arkref_mem, 100000L * maxsteps);
  }

  double t = 0.0, t2 = 0.0, dTout = prm.tf / nout, tout = dTout, errtot = 0.0, simtime = 0.0;
  auto rms = [&](N_Vector v) { return sqrt(N_VDotProd(v, v) / (double)nx / (double)ny); };
  if (output > 0) {
    printf("\n          t                     ||u||_rms      \n");
    printf(" ----------------------------------------------\n");
    printf("%22.15e%25.15e\n", t, rms(u));
  }
  for (int iout = 0; iout < nout; iout++) {
    const double t0 = wall();
    if (error) ARKodeSetStopTime(arkode_mem, tout);
    int flag = ARKodeEvolve(arkode_mem, tout, u, &t, ARK_NORMAL);
    (void)hipDeviceSynchronize();
    simtime += wall() - t0;
    if (flag < 0) {
      fprintf(stderr, "\nERROR: ARKodeEvolve returned with flag = %d\n", flag);
      return 1;
    }
    if (error) {
      ARKodeSetStopTime(arkref_mem, t);
      if (ARKodeEvolve(arkref_mem, tout, uref, &t2, ARK_NORMAL) < 0) {
        fprintf(stderr, "\nERROR: ARKodeEvolve returned with flag < 0 (reference run)\n");
        break;
      }
      N_VLinearSum(1.0, uref, -1.0, u, uerr);
      const double e = N_VMaxNorm(uerr) / N_VMaxNorm(uref);
      if (e > errtot) errtot = e;
    }
    if (output > 0) printf("%22.15e%25.15e\n", t, rms(u));
    tout += dTout;
    if (tout > prm.tf) tout = prm.tf;
  }
  if (output > 0) printf(" ----------------------------------------------\n\n");
  if (error) printf("Maximum relative error = %.16e\n", errtot);
  if (kind == DIFFUSION_MMS_DIRICHLET) {
    N_Vector ue = N_VClone(u);
    Exact(t, ue, udata);
    N_VLinearSum(1.0, u, -1.0, ue, ue);
    const double num = N_VDotProd(ue, ue);
    Exact(t, ue, udata);
    const double den = N_VDotProd(ue, ue);
    printf("L2 error vs exact = %.16e\n", sqrt(num / (double)(nx * ny)));
    printf("Relative L2 error vs exact = %.16e\n", den > 0 ? sqrt(num / den) : -1.0);
    N_VDestroy(ue);
  }
  printf("Total simulation time = %.6f\n\n", simtime);
  printf("Final integrator statistics:\n");
  ARKodePrintAllStats(arkode_mem, stdout, SUN_OUTPUTFORMAT_TABLE);
  if (write_final) {
    N_VCopyFromDevice_HipNative(u);
    FILE* fp = fopen(write_final, "wb");
    if (fp) {
      fwrite(N_VGetArrayPointer(u), sizeof(double), (size_t)(nx * ny), fp);
      fclose(fp);
    }
  }
  SUNAdaptController_Destroy(C);
  ARKodeFree(&arkode_mem);
  if (arkref_mem) ARKodeFree(&arkref_mem);
  if (LS) SUNLinSolFree(LS);
  if (DEE) SUNDomEigEstimator_Destroy(&DEE);
  N_VDestroy(u);
  if (uref) N_VDestroy(uref);
  if (uerr) N_VDestroy(uerr);
  DiffusionProblem_Destroy(udata);
  SUNContext_Free(&ctx);
  return 0;
}
