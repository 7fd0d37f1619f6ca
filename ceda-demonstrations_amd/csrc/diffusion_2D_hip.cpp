// diffusion_2D driver on one MI355X.
//
// Mirrors diffusion_2D/main.cpp: the same flags (UserData::parse_args
// diffusion_2D.cpp:42-163, UserOptions::parse_args main.cpp:556-687,
// UserOutput::parse_args diffusion_2D.cpp:651-677), the same sequence of
// integrator calls (main.cpp:176-400), the output loop with
// ||u||_rms = sqrt(u.u/nx/ny) (main.cpp:431-470, diffusion_2D.cpp:801), the
// tight-tolerance reference run of --error (main.cpp:387-465), the one-step
// mode of --onestep (main.cpp:407-414), the solution files of --output 2
// (UserOutput::open/write, diffusion_2D.cpp:729-787, read by
// plot_data_2d.py:271-586) and the final lines that
// runtests-diffusion2d.py:36-54 parses.  The hypre build's --pfmg_relax and
// --pfmg_nrelax (diffusion_2D.cpp:137-151) select the GPU multigrid
// preconditioner in place of Jacobi (periodic grids: even sides) and print
// "Total PFMG iterations" (main.cpp:489-492).  Added flags:
//   --problem ceda|mms   manufactured Dirichlet problem of README.md:11-32
//   --deefreq N --stagemax N --eigsafety X   LSRKStep settings
//   --fastpath 0|1|2|3   0: N_Vector API path, 1: fused stage kernels,
//                        2: fused kernels replayed as HIP graphs, 3: as 2
//                        with autotuned launch shapes (multi-stage passes)
//   --ls cg|gmres|fgmres --ls_reltol X --write_final FILE
//   --learned 1          implicit path: the learned symmetric network in the
//                        preconditioner slot, trained on the GPU before the
//                        run (--learned_pairs N, --learned_grid n: training
//                        set size), Jacobi-PCG fallback for the GMRES solvers
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "arkode_hip.h"
#include "diffusion_hip.h"
#include "learned_prec.h"
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
  int onestep = 0, pfmg = 0, pfmg_relax = 2, pfmg_nrelax = 2;
  int learned = 0, learned_pairs = 20000, learned_grid = 192;
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
    else if (TAKE("--onestep")) onestep = atoi(v);
    else if (TAKE("--pfmg_relax")) {
      pfmg = 1;
      pfmg_relax = atoi(v);
    } else if (TAKE("--pfmg_nrelax")) {
      pfmg = 1;
      pfmg_nrelax = atoi(v);
    }
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
    else if (TAKE("--learned")) learned = atoi(v);
    else if (TAKE("--learned_pairs")) learned_pairs = atoi(v);
    else if (TAKE("--learned_grid")) learned_grid = atoi(v);
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
  if (learned && (!impl || !precond || pfmg || learned_pairs < 1 || learned_grid < 16 || learned_grid > 1024)) {
    fprintf(stderr, "ERROR: --learned needs --integrator dirk with a preconditioner (not --pfmg_*), "
                    "--learned_pairs >= 1 and --learned_grid 16..1024\n");
    return 1;
  }

  SUNContext ctx = nullptr;
  if (SUNContext_Create(SUN_COMM_NULL, &ctx)) return 1;
  DiffusionProblem* udata = DiffusionProblem_Create(&prm, ctx);
  if (!udata) {
    fprintf(stderr, "ERROR: problem setup\n");
    return 1;
  }
  const bool odd = (prm.nx % 2 != 0) || (prm.ny % 2 != 0);
  if (pfmg && ((kind == DIFFUSION_CEDA_PERIODIC && odd) ||
               DiffusionProblem_SetMGOptions(udata, pfmg_relax, pfmg_nrelax))) {
    fprintf(stderr, "ERROR: the multigrid preconditioner needs even --nx and --ny for the periodic "
                    "problem, --pfmg_relax 2|3 and --pfmg_nrelax 0..16\n");
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

  // learned preconditioner: the symmetric network K^T K (4 layers, dilations
  // 1, 3, 9, 27), trained on pairs from Jacobi-PCG solves with this run's
  // operator I - rx Dxx - ry Dyy, r = gamma*h*k/d^2 along each axis (h =
  // --fixedstep, or tf/nout)
  LearnedPrec* net = nullptr;
  if (learned) {
    const double gam = (order == 3) ? 1767732205903.0 / 4055673282236.0 : 1.0;  // DIRK diagonal
    const double h = (hfixed > 0.0) ? hfixed : prm.tf / nout;
    const double rx = gam * h * (prm.kx / (dx * dx)), ry = gam * h * (prm.ky / (dy * dy));
    const int dil[4] = {1, 3, 9, 27};
    const double tt = wall();
    net = LearnedPrec_CreateSymmetric(4, dil, ctx);
    double *r = nullptr, *e = nullptr;
    const size_t bytes = sizeof(double) * (size_t)learned_pairs * learned_grid * learned_grid;
    int its = 0;
    bool ok = net && hipMalloc(&r, bytes) == hipSuccess && hipMalloc(&e, bytes) == hipSuccess;
    ok = ok && LearnedPrec_MakeTrainingSetXY(learned_pairs, learned_grid, rx, ry, 1e-12, 1234, r, e, &its, ctx) == 0;
    ok = ok && LearnedPrec_InitIdentity(net, 8e-4, 0) == 0;
    ok = ok && LearnedPrec_SetOutputScale(net, 1.0 / (1.0 + 2.0 * rx + 2.0 * ry)) == 0;
    ok = ok && LearnedPrec_TrainXY(net, learned_pairs, learned_grid, rx, ry, r, e, 1000, 16, 3e-3, 0, nullptr) == 0;
    if (r) (void)hipFree(r);
    if (e) (void)hipFree(e);
    if (!ok || DiffusionProblem_SetLearnedPrec(udata, net)) {
      fprintf(stderr, "ERROR: learned preconditioner setup\n");
      return 1;
    }
    printf("  learned preconditioner: symmetric network trained in %.2f s (gamma*h*k/d^2 = %g, %g, %d pairs of "
           "%d^2)\n", wall() - tt, rx, ry, learned_pairs, learned_grid);
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
      if (pfmg) ARKodeSetPreconditioner(arkode_mem, MGPSetup, MGPSolve);
      else if (net) {
        ARKodeSetPreconditioner(arkode_mem, LearnedPSetup, LearnedPSolve);
        if (ls != "cg") ARKodeSetLSFallbackPCG(arkode_mem, 20000, PSolve, udata);
      } else ARKodeSetPreconditioner(arkode_mem, PSetup, PSolve);
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
    if (fastpath > 2 && DiffusionProblem_Autotune(udata, 3) != 0) {
      fprintf(stderr, "autotune failed\n");
      return 1;
    }
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

  int stepmode = ARK_NORMAL;
  if (onestep) {
    nout = onestep;
    stepmode = ARK_ONE_STEP;
  }
  double t = 0.0, t2 = 0.0, dTout = prm.tf / nout, tout = dTout, errtot = 0.0, simtime = 0.0;
  auto rms = [&](N_Vector v) { return sqrt(N_VDotProd(v, v) / (double)nx / (double)ny); };
  // --output 2: one file per rank (here rank 0), a header of "# key value"
  // lines in the default stream format, then per output time a row
  // "t u_0 u_1 ... " in %.15e (UserOutput::open/write)
  FILE* uout = nullptr;
  if (output == 2) {
    uout = fopen("diffusion_2d_solution.00000.txt", "w");
    if (!uout) {
      fprintf(stderr, "ERROR: cannot open diffusion_2d_solution.00000.txt\n");
      return 1;
    }
    // node coordinates of the stored nodes run from xl to xu (periodic: the
    // domain bounds; manufactured: the first and last interior nodes)
    const bool mms = (kind == DIFFUSION_MMS_DIRICHLET);
    const double xl = mms ? prm.xl + dx : prm.xl, xu = mms ? prm.xu - dx : prm.xu;
    const double yl = mms ? prm.yl + dy : prm.yl, yu = mms ? prm.yu - dy : prm.yu;
    fprintf(uout, "# title Diffusion 2D\n# nvar 1\n# vars u\n# nt  %d\n", nout + 1);
    fprintf(uout, "# nx  %ld\n# xl  %g\n# xu  %g\n", (long)nx, xl, xu);
    fprintf(uout, "# ny  %ld\n# yl  %g\n# yu  %g\n", (long)ny, yl, yu);
    fprintf(uout, "# px  1\n# py  1\n# np  1\n# is  0\n# ie  %ld\n# js  0\n# je  %ld\n", (long)(nx - 1),
            (long)(ny - 1));
  }
  auto write_solution = [&](double tt) {
    if (!uout) return;
    N_VCopyFromDevice_HipNative(u);  // synchronises the vector's stream
    const double* h = N_VGetArrayPointer(u);
    fprintf(uout, "%.15e ", tt);
    for (int64_t k = 0; k < nx * ny; k++) fprintf(uout, "%.15e ", h[k]);
    fprintf(uout, "\n");
  };
  if (output > 0) {
    printf("\n          t                     ||u||_rms      \n");
    printf(" ----------------------------------------------\n");
    printf("%22.15e%25.15e\n", t, rms(u));
  }
  write_solution(t);
  for (int iout = 0; iout < nout; iout++) {
    const double t0 = wall();
    if (error) ARKodeSetStopTime(arkode_mem, tout);
    int flag = ARKodeEvolve(arkode_mem, tout, u, &t, stepmode);
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
    write_solution(t);
    tout += dTout;
    if (tout > prm.tf) tout = prm.tf;
  }
  if (output > 0) printf(" ----------------------------------------------\n\n");
  if (uout) fclose(uout);
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
  if (impl && precond && pfmg) printf("Total PFMG iterations        = %ld\n", DiffusionProblem_MGNumCycles(udata));
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
  if (net) LearnedPrec_Destroy(net);
  DiffusionProblem_Destroy(udata);
  SUNContext_Free(&ctx);
  return 0;
}
