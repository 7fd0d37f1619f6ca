/* ---------------------------------------------------------------------------
 * CPU oracle for the diffusion_2D hot path — TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C, serial FP64 restatement of
 *   - the reference problem layer (diffusion_2D/diffusion.cpp:9-209 stencil,
 *     diffusion_2D/diffusion_2D.cpp:887-897 coefficients, initial.cpp:20-48,
 *     preconditioner_jacobi.cpp:9-62, main.cpp:536-550 dom_eig), and
 *   - the subset of SUNDIALS it is driven by (LSRKStep RKC_2/RKL_2, the power
 *     dominant-eigenvalue estimator, ARKODE step adaptivity, ARKStep DIRK with
 *     a Newton/PCG/Jacobi solve).  SUNDIALS is an un-vendored submodule of the
 *     reference (.gitmodules:1-3, pinned commit unrecoverable); its algorithms
 *     are restated from the published method papers and the ARKODE
 *     documentation, so agreement with a real SUNDIALS build is unpinned.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker or the timed CPU baseline.  The
 * product path never links it.
 *
 * Floating-point contract shared with the HIP kernels: every element-wise
 * formula is evaluated in the order written here, without FMA contraction
 * (built with -ffp-contract=off), so stencil and vector results are bitwise
 * reproducible; only reductions differ in summation order.
 * ------------------------------------------------------------------------- */
#ifndef CEDA_ORACLE_H
#define CEDA_ORACLE_H

#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_CEDA_PERIODIC = 0, ORC_MMS_DIRICHLET = 1 };
enum { ORC_RKC = 0, ORC_RKL = 1, ORC_DIRK = 2 };
enum { ORC_CTRL_PID = 0, ORC_CTRL_PI = 1, ORC_CTRL_I = 2, ORC_CTRL_EXPGUS = 3,
       ORC_CTRL_IMPGUS = 4, ORC_CTRL_IMEXGUS = 5 };

/* return codes (ARKODE convention: 0 ok, <0 fatal) */
#define ORC_SUCCESS 0
#define ORC_TOO_MUCH_WORK (-1)
#define ORC_ERR_FAILURE (-3)
#define ORC_CONV_FAILURE (-4)
#define ORC_LSOLVE_FAIL (-7)
#define ORC_RHSFUNC_FAIL (-8)
#define ORC_MAX_STAGE_LIMIT_FAIL (-51)
#define ORC_ILL_INPUT (-22)
#define ORC_RETRY_STEP 100

typedef struct {
  int kind;
  int64_t nx, ny, n;
  double xl, xu, yl, yu, dx, dy, kx, ky, tf;
  int inhomogeneous;
  double* fx;  /* stencil face coeffs Kx(xl+(k-0.5)dx)/(dx*dx), k=0..nx     */
  double* fy;  /* same in y, k=0..ny                                         */
  double* px;  /* Jacobi face coeffs Kx(k*dx)/(dx*dx) (reference quirk)     */
  double* py;
  double* sx;  /* MMS: sin^2(pi x_i)                                          */
  double* sy;
  double* cx;  /* MMS: kx * (-2 pi^2 (cos^2(pi x_i) - sin^2(pi x_i)))          */
  double* cy;
  double* diag; /* Jacobi inverse diagonal (PSetup output)                   */
  long nrhs;    /* RHS evaluations on this problem                           */
} orc_problem;

/* ---- problem layer ---- */
void orc_problem_defaults(orc_problem* p, int kind);
int orc_problem_setup(orc_problem* p);
void orc_problem_free(orc_problem* p);
void orc_initial(const orc_problem* p, double* u);
void orc_exact(const orc_problem* p, double t, double* u);
void orc_mms_time(double t, double* T1, double* T2);
int orc_rhs(orc_problem* p, double t, const double* u, double* f);
void orc_stencil(const orc_problem* p, const double* u, double* f);
void orc_psetup(orc_problem* p, double gamma);
void orc_psolve(const orc_problem* p, const double* r, double* z);
void orc_dom_eig(const orc_problem* p, double* lr, double* li);
uint64_t orc_splitmix64(uint64_t x);
void orc_seed_vector(int64_t n, uint64_t seed, double* v);

/* ---- integrator ---- */
typedef struct orc_integrator orc_integrator;

orc_integrator* orc_create(orc_problem* p, int method, double t0, const double* y0);
void orc_free(orc_integrator* I);
void orc_set_tolerances(orc_integrator* I, double rtol, double atol);
void orc_set_fixed_step(orc_integrator* I, double h);
void orc_set_init_step(orc_integrator* I, double h0);
void orc_set_controller(orc_integrator* I, int type);
void orc_set_max_num_steps(orc_integrator* I, long mxstep);
/* LSRK options */
void orc_set_internal_eig(orc_integrator* I, int on, int maxiters, double rtol);
void orc_set_dom_eig_frequency(orc_integrator* I, long freq);
void orc_set_max_num_stages(orc_integrator* I, int smax);
void orc_set_dom_eig_safety(orc_integrator* I, double safety);
void orc_set_dee_preprocess(orc_integrator* I, int init_iters, int iters);
/* DIRK options */
void orc_set_order(orc_integrator* I, int order);
void orc_set_linear(orc_integrator* I, int linear);
void orc_set_lin_solver(orc_integrator* I, int maxl, double epslin, int msbp, int precond);
void orc_set_ls_reltol(orc_integrator* I, double reltol);
void orc_set_jtimes_analytic(orc_integrator* I, int on);
int orc_evolve(orc_integrator* I, double tout, double* yout, double* tret);
int orc_get_stats(const orc_integrator* I, long* out, int nout);
double orc_get_current_h(const orc_integrator* I);
int orc_set_threads(int n);
double orc_get_h0_used(const orc_integrator* I);
double orc_get_last_dsm(const orc_integrator* I);
int orc_test_controller(int ctype, int p_order, double C, double q, double h0, double ubound,
                        int nattempts, double* h_out, double* dsm_out, int* acc_out);
double orc_get_spectral_radius(const orc_integrator* I);
void orc_print_stats(const orc_integrator* I, FILE* fp);
int orc_test_linear_solve(orc_integrator* I, double t, double gamma, const double* b,
                          double* x, double reltol, int maxl, long* iters);
/* DIRK tableau access for order-condition tests */
int orc_dirk_table(int order, double* A, double* b, double* bhat, double* c);

/* stats layout for orc_get_stats */
enum {
  ORC_ST_NST = 0, ORC_ST_ATTEMPTS, ORC_ST_NFE, ORC_ST_NETF, ORC_ST_NSTLIM,
  ORC_ST_NDOMEIG, ORC_ST_NDEE_ITERS, ORC_ST_STAGE_MAX, ORC_ST_LAST_STAGES,
  ORC_ST_NFE_DEE, ORC_ST_NLI, ORC_ST_NPSETUP, ORC_ST_NPSOLVE, ORC_ST_NNI,
  ORC_ST_NCFN, ORC_ST_NLSFAIL, ORC_ST_NJTIMES, ORC_ST_NFE_LS, ORC_ST_COUNT
};

/* ---- vector helpers exported for the known-answer tests ---- */
void orc_vlinearcombination(int nv, const double* c, const double* const* X,
                            double* z, int64_t n);
double orc_vwrmsnorm(const double* x, const double* w, int64_t n);
double orc_vdot(const double* x, const double* y, int64_t n);
double orc_vmaxnorm(const double* x, int64_t n);

/* geometric-multigrid V-cycle (orc_mg.c): z = M r for I - gamma J */
int orc_mg_apply(const orc_problem* p, double gamma, const double* r, double* z);
/* the same with smoother relax type 2 (symmetric) or 3 and nu sweep pairs */
int orc_mg_apply_opts(const orc_problem* p, double gamma, int relax, int nu, const double* r, double* z);

/* learned preconditioner network (orc_learned.c) */
int orc_lp_num_params(int nl, const int* dil);
double orc_lp_lossgrad(int nl, const int* dil, const float* params, double out_scale, int nb, int nx,
                       int ny, double ratio_x, double ratio_y, const double* r, const double* e, int emu,
                       double* grad, double* u_out);
/* the symmetric network N = K^T K (linear layers, no biases) */
int orc_lp_sym_num_params(int nl, const int* dil);
double orc_lp_sym_lossgrad(int nl, const int* dil, const float* params, double out_scale, int nb, int nx,
                           int ny, double ratio_x, double ratio_y, const double* r, const double* e, int emu,
                           double* grad, double* u_out);

#ifdef __cplusplus
}
#endif

#endif
