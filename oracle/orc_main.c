/* ---------------------------------------------------------------------------
 * CPU oracle driver — TEST INFRASTRUCTURE ONLY (see orc.h).
 *
 * Mirrors diffusion_2D/main.cpp: the command-line flags of
 * UserData::parse_args (diffusion_2D.cpp:42-163), UserOptions::parse_args
 * (main.cpp:556-687) and UserOutput::parse_args (diffusion_2D.cpp:651-677),
 * the output loop (main.cpp:431-470, ||u||_rms = sqrt(u.u/nx/ny) as at
 * diffusion_2D.cpp:801) and the final lines parsed by
 * runtests-diffusion2d.py:36-54.  Extra flags of this restatement:
 *   --problem ceda|mms   (mms = manufactured Dirichlet problem of README.md)
 *   --deefreq N --stagemax N --eigsafety X --ls_reltol X
 *   --write_final FILE   (raw little-endian FP64 final state)
 * ------------------------------------------------------------------------- */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "orc.h"

static double wall(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int main(int argc, char** argv)
{
  int kind = ORC_CEDA_PERIODIC;
  /* first pass: problem kind decides the defaults */
  for (int a = 1; a < argc - 1; a++)
    if (!strcmp(argv[a], "--problem"))
      kind = !strcmp(argv[a + 1], "mms") ? ORC_MMS_DIRICHLET : ORC_CEDA_PERIODIC;
  orc_problem P;
  orc_problem_defaults(&P, kind);
  const char* integrator = "dirk";
  double rtol = 1e-5, atol_ = 1e-10, hfixed = 0.0, epslin = 0.0, eigsafety = 1.01;
  double ls_reltol = 0.0;
  int order = 2, controller = 0, maxsteps = 0, error = 0, linear = 1, precond = 1;
  int liniters = 20, msbp = 0, internaleig = 0, output = 1, nout = 20;
  long deefreq = 25;
  int stagemax = 200;
  const char* write_final = NULL;
  for (int a = 1; a < argc; a++) {
    const char* k = argv[a];
    const char* v = (a + 1 < argc) ? argv[a + 1] : NULL;
#define TAKE(name) (!strcmp(k, name) && v && (a++, 1))
    if (TAKE("--problem")) continue;
    else if (TAKE("--nx")) P.nx = atol(v);
    else if (TAKE("--ny")) P.ny = atol(v);
    else if (TAKE("--xl")) P.xl = atof(v);
    else if (TAKE("--xu")) P.xu = atof(v);
    else if (TAKE("--yl")) P.yu = atof(v); /* as written at diffusion_2D.cpp:95-100 */
    else if (TAKE("--yu")) P.yu = atof(v);
    else if (TAKE("--kx")) P.kx = atof(v);
    else if (TAKE("--ky")) P.ky = atof(v);
    else if (!strcmp(k, "--inhomogeneous")) P.inhomogeneous = 1;
    else if (TAKE("--tf")) P.tf = atof(v);
    else if (TAKE("--npx") || TAKE("--npy")) continue;
    else if (TAKE("--integrator")) integrator = v;
    else if (TAKE("--rtol")) rtol = atof(v);
    else if (TAKE("--atol")) atol_ = atof(v);
    else if (TAKE("--fixedstep")) hfixed = atof(v);
    else if (TAKE("--order")) order = atoi(v);
    else if (TAKE("--controller")) controller = atoi(v);
    else if (TAKE("--maxsteps")) maxsteps = atoi(v);
    else if (!strcmp(k, "--error")) error = 1;
    else if (!strcmp(k, "--nonlinear")) linear = 0;
    else if (TAKE("--ls")) continue;
    else if (!strcmp(k, "--lsinfo")) continue;
    else if (TAKE("--liniters")) liniters = atoi(v);
    else if (TAKE("--msbp")) msbp = atoi(v);
    else if (TAKE("--epslin")) epslin = atof(v);
    else if (!strcmp(k, "--noprec")) precond = 0;
    else if (!strcmp(k, "--internaleig")) internaleig = 1;
    else if (TAKE("--output")) output = atoi(v);
    else if (TAKE("--nout")) nout = atoi(v);
    else if (TAKE("--deefreq")) deefreq = atol(v);
    else if (TAKE("--stagemax")) stagemax = atoi(v);
    else if (TAKE("--eigsafety")) eigsafety = atof(v);
    else if (TAKE("--ls_reltol")) ls_reltol = atof(v);
    else if (TAKE("--write_final")) write_final = v;
    else {
      fprintf(stderr, "ERROR: Unknown inputs: %s\n", k);
      return 1;
    }
#undef TAKE
  }
  int method;
  if (!strcmp(integrator, "rkc")) method = ORC_RKC;
  else if (!strcmp(integrator, "rkl")) method = ORC_RKL;
  else if (!strcmp(integrator, "dirk")) method = ORC_DIRK;
  else {
    fprintf(stderr, "ERROR: illegal integrator\n");
    return 1;
  }
  if (orc_problem_setup(&P)) {
    fprintf(stderr, "ERROR: problem setup\n");
    return 1;
  }
  printf("\n Problem options:\n");
  printf("  problem        = %s\n", kind == ORC_MMS_DIRICHLET ? "mms" : "ceda");
  printf("  kx             = %g\n  ky             = %g\n", P.kx, P.ky);
  printf("  %s diffusion\n", P.inhomogeneous ? "spatially-varying" : "homogeneous");
  printf("  tf             = %g\n", P.tf);
  printf("  nx             = %ld\n  ny             = %ld\n", (long)P.nx, (long)P.ny);
  printf("  dx             = %.16g\n  dy             = %.16g\n", P.dx, P.dy);
  printf(" Integrator: %s\n  rtol        = %g\n  atol        = %g\n", integrator, rtol, atol_);

  double* u = (double*)malloc(sizeof(double) * (size_t)P.n);
  double* uref = NULL;
  orc_initial(&P, u);
  orc_integrator* I = orc_create(&P, method, 0.0, u);
  orc_set_tolerances(I, rtol, atol_);
  orc_set_max_num_steps(I, maxsteps);
  if (method == ORC_DIRK) {
    orc_set_order(I, order);
    orc_set_linear(I, linear);
    orc_set_lin_solver(I, liniters, epslin, msbp, precond);
    if (ls_reltol > 0.0) orc_set_ls_reltol(I, ls_reltol);
  } else {
    if (internaleig) orc_set_internal_eig(I, 1, 100, 0.01);
    orc_set_dom_eig_frequency(I, deefreq);
    orc_set_max_num_stages(I, stagemax);
    orc_set_dom_eig_safety(I, eigsafety);
  }
  if (hfixed > 0.0) orc_set_fixed_step(I, hfixed);
  else orc_set_controller(I, controller);

  orc_integrator* R = NULL;
  orc_problem PR;
  if (error) {
    /* main.cpp:387-400: RKC reference at rtol 1e-12 with the closed-form
       dom_eig, evolved to each output time */
    PR = P;
    orc_problem_setup(&PR);
    uref = (double*)malloc(sizeof(double) * (size_t)P.n);
    memcpy(uref, u, sizeof(double) * (size_t)P.n);
    R = orc_create(&PR, ORC_RKC, 0.0, u);
    orc_set_tolerances(R, 1e-12, atol_);
    orc_set_max_num_steps(R, 100000L * maxsteps);
  }

  const double dTout = P.tf / nout;
  double tout = dTout, t = 0.0, errtot = 0.0, simtime = 0.0;
  if (output > 0) {
    printf("\n          t                     ||u||_rms      \n");
    printf(" ----------------------------------------------\n");
    printf("%22.15e%25.15e\n", t, sqrt(orc_vdot(u, u, P.n) / (double)P.nx / (double)P.ny));
  }
  int flag = 0;
  for (int iout = 0; iout < nout; iout++) {
    const double t0 = wall();
    flag = orc_evolve(I, tout, u, &t);
    simtime += wall() - t0;
    if (flag < 0) {
      fprintf(stderr, "\nERROR: ARKodeEvolve returned with flag = %d\n", flag);
      return 1;
    }
    if (error) {
      double t2;
      if (orc_evolve(R, t, uref, &t2) < 0) {
        fprintf(stderr, "\nERROR: ARKodeEvolve returned with flag < 0 (reference run)\n");
        break;
      }
      double num = 0.0, den = 0.0;
      for (int64_t i = 0; i < P.n; i++) {
        const double d = fabs(uref[i] - u[i]);
        if (d > num) num = d;
        if (fabs(uref[i]) > den) den = fabs(uref[i]);
      }
      if (num / den > errtot) errtot = num / den;
    }
    if (output > 0)
      printf("%22.15e%25.15e\n", t, sqrt(orc_vdot(u, u, P.n) / (double)P.nx / (double)P.ny));
    tout += dTout;
    if (tout > P.tf) tout = P.tf;
  }
  if (output > 0) printf(" ----------------------------------------------\n\n");
  if (error) printf("Maximum relative error = %.16e\n", errtot);
  if (kind == ORC_MMS_DIRICHLET) {
    double* ue = (double*)malloc(sizeof(double) * (size_t)P.n);
    orc_exact(&P, t, ue);
    double num = 0.0, den = 0.0;
    for (int64_t i = 0; i < P.n; i++) {
      num += (u[i] - ue[i]) * (u[i] - ue[i]);
      den += ue[i] * ue[i];
    }
    printf("L2 error vs exact = %.16e\n", sqrt(num / (double)P.n));
    printf("Relative L2 error vs exact = %.16e\n", den > 0 ? sqrt(num / den) : -1.0);
    free(ue);
  }
  printf("Total simulation time = %.6f\n\n", simtime);
  printf("Final integrator statistics:\n");
  orc_print_stats(I, stdout);
  if (write_final) {
    FILE* fp = fopen(write_final, "wb");
    if (fp) {
      fwrite(u, sizeof(double), (size_t)P.n, fp);
      fclose(fp);
    }
  }
  orc_free(I);
  if (R) {
    orc_free(R);
    orc_problem_free(&PR);
    free(uref);
  }
  orc_problem_free(&P);
  free(u);
  return 0;
}
