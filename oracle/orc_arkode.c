/* ---------------------------------------------------------------------------
 * CPU oracle, integrator — TEST INFRASTRUCTURE ONLY (see orc.h).
 *
 * Restates the SUNDIALS pieces the reference driver configures in
 * diffusion_2D/main.cpp:
 *   LSRKStepCreateSTS + LSRKStepSetSTSMethod(RKC_2|RKL_2)   main.cpp:257,336-339
 *   SUNDomEigEstimator_Power(u,100,0.01) / LSRKStepSetDomEigFn main.cpp:342-352
 *   ARKodeSStolerances, SUNAdaptController_{PID,PI,I,...}    main.cpp:262,357-377
 *   ARKStepCreate + ARKodeSetOrder (DIRK), SUNLinSol_PCG,
 *   ARKodeSetPreconditioner/LSetupFrequency/EpsLin/Linear     main.cpp:207-329
 * Methods: RKC — Sommeijer, Shampine & Verwer, J. Comput. Appl. Math. 88
 * (1998) 315-326 (eps = 2/13, error estimate (12(y_n-y_{n+1})+6h(F_n+F_{n+1}))/15);
 * RKL2 — Meyer, Balsara & Aslam, J. Comput. Phys. 257 (2014) 594-626;
 * ESDIRK ARK3(2)4L[2]SA — Kennedy & Carpenter, Appl. Numer. Math. 44 (2003);
 * power iteration with a difference-quotient Jacobian-vector product;
 * Soderlind/Gustafsson step controllers; initial step by the CVODE/ARKODE
 * second-derivative heuristic; PCG by Hestenes-Stiefel with a zero guess.
 * Step-sequence agreement with a real SUNDIALS build is unpinned.
 * ------------------------------------------------------------------------- */
#include <float.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdlib.h>
#include <string.h>

#include "orc.h"

#define DEE_SEED 20240917ull
#define TINY 1.0e-10
#define ONEMSM 0.999999
#define ONEPSM 1.000001
#define ORC_CONV_RECOVER 101

struct orc_integrator {
  orc_problem* P;
  int method;
  int64_t n;
  double rtol, atol;
  double tn, h, h0u, hfixed, hold, h0_used;
  int fixedstep, initialized;
  long mxstep;
  double *yn, *ycur, *fn, *tv1, *tv2, *ewt;
  int fn_is_current;
  /* controller */
  int ctype, p_order, firststeps;
  double k1, k2, k3, bias, ep, epp, hp, hpp;
  /* adaptivity */
  double safety, growth, etamx1, etamxf, etamin, etacf, lbound, ubound, eta;
  int small_nef, maxnef, maxncf;
  /* LSRK */
  double rho, rho_max, dom_eig_safety, dee_rtol;
  long dom_eig_freq, dom_eig_nst;
  int dom_eig_update, smax, internal_eig, dee_maxiters, dee_init_pre, dee_pre,
    dee_started;
  double *dee_V, *dee_q, *dee_work;
  /* DIRK */
  int order, stages, linear, maxl, msbp, precond, psetup_done, maxcor;
  int jtimes_analytic;
  double A[16], b[4], bh[4], c[4];
  double* F[4];
  double *sdata, *zcur, *fz, *rhs, *dz, *r, *pv, *zv, *Ap, *jwork, *jv;
  double epslin, ls_reltol, gammap, nlscoef;
  long nstlp;
  double dsm_last;  /* error estimate of the last accepted step */
  /* statistics */
  long nst, attempts, nfe, netf, nstlim, ndomeig, ndee_iters, stage_max,
    last_stages, nfe_dee, nli, npsetup, npsolve, nni, ncfn, nlsfail, njtimes,
    nfe_ls;
};

/* ---------------- serial vector kernels (fixed evaluation order) --------- */
static void vcopy(double* z, const double* x, int64_t n)
{
  memcpy(z, x, sizeof(double) * (size_t)n);
}
/* The OpenMP pragmas below only act in the CPU-baseline timing build
   (liborc_omp.so, -fopenmp); the checker build ignores them and stays
   serial, so its reductions keep their fixed serial order. */
static void vlinsum(double a, const double* x, double b, const double* y,
                    double* z, int64_t n)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) z[i] = a * x[i] + b * y[i];
}
void orc_vlinearcombination(int nv, const double* c, const double* const* X,
                            double* z, int64_t n)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) {
    double acc = c[0] * X[0][i];
    for (int k = 1; k < nv; k++) acc = acc + c[k] * X[k][i];
    z[i] = acc;
  }
}
double orc_vwrmsnorm(const double* x, const double* w, int64_t n)
{
  double s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s)
  for (int64_t i = 0; i < n; i++) {
    const double p = x[i] * w[i];
    s += p * p;
  }
  return sqrt(s / (double)n);
}
double orc_vdot(const double* x, const double* y, int64_t n)
{
  double s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s)
  for (int64_t i = 0; i < n; i++) s += x[i] * y[i];
  return s;
}
double orc_vmaxnorm(const double* x, int64_t n)
{
  double m = 0.0;
  for (int64_t i = 0; i < n; i++) {
    const double a = fabs(x[i]);
    if (a > m) m = a;
  }
  return m;
}
static double vwl2(const double* x, const double* w, int64_t n)
{
  double s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s)
  for (int64_t i = 0; i < n; i++) {
    const double p = x[i] * w[i];
    s += p * p;
  }
  return sqrt(s);
}

static double* vzalloc(int64_t n) { return (double*)calloc((size_t)n, sizeof(double)); }

static int rhs(orc_integrator* I, double t, const double* y, double* f)
{
  I->nfe++;
  return orc_rhs(I->P, t, y, f);
}

/* ARKODE arkEwtSetSS: ewt = 1/(rtol*|y| + atol) */
static int ewt_set(orc_integrator* I, const double* y)
{
  for (int64_t i = 0; i < I->n; i++) {
    const double d = I->rtol * fabs(y[i]) + I->atol;
    if (d <= 0.0) return -1;
    I->ewt[i] = 1.0 / d;
  }
  return 0;
}

/* ---------------- construction / options ---------------- */
int orc_dirk_table(int order, double* A, double* b, double* bhat, double* c)
{
  memset(A, 0, sizeof(double) * 16);
  if (order == 2) {
    /* SDIRK-2-1-2 (ARKODE default order-2 DIRK) */
    A[0] = 1.0;
    A[4] = -1.0; A[5] = 1.0;
    b[0] = 0.5; b[1] = 0.5;
    bhat[0] = 1.0; bhat[1] = 0.0;
    c[0] = 1.0; c[1] = 0.0;
    return 2;
  }
  if (order == 3) {
    /* ESDIRK ARK3(2)4L[2]SA implicit table (Kennedy & Carpenter 2003) */
    const double g = 1767732205903.0 / 4055673282236.0;
    A[4] = g; A[5] = g;
    A[8] = 2746238789719.0 / 10658868560708.0;
    A[9] = -640167445237.0 / 6845629431997.0;
    A[10] = g;
    A[12] = 1471266399579.0 / 7840856788654.0;
    A[13] = -4482444167858.0 / 7529755066697.0;
    A[14] = 11266239266428.0 / 11593286722821.0;
    A[15] = g;
    b[0] = A[12]; b[1] = A[13]; b[2] = A[14]; b[3] = A[15];
    bhat[0] = 2756255671327.0 / 12835298489170.0;
    bhat[1] = -10771552573575.0 / 22201958757719.0;
    bhat[2] = 9247589265047.0 / 10645013368117.0;
    bhat[3] = 2193209047091.0 / 5459859503100.0;
    c[0] = 0.0; c[1] = 1767732205903.0 / 2027836641118.0; c[2] = 0.6; c[3] = 1.0;
    return 4;
  }
  return -1;
}

static void set_ctrl_params(orc_integrator* I)
{
  I->k3 = 0.0;
  switch (I->ctype) {
  case ORC_CTRL_PID: I->k1 = 0.58; I->k2 = 0.21; I->k3 = 0.1; break;
  case ORC_CTRL_PI: I->k1 = 0.8; I->k2 = 0.31; break;
  case ORC_CTRL_I: I->k1 = 1.0; I->k2 = 0.0; break;
  case ORC_CTRL_EXPGUS: I->k1 = 0.367; I->k2 = 0.268; break;
  case ORC_CTRL_IMPGUS: I->k1 = 0.98; I->k2 = 0.95; break;
  default: I->k1 = 0.367; I->k2 = 0.268; break; /* ImExGus: explicit part */
  }
}

/* ARKODE's time-step adaptivity defaults: the PID controller, bias 1.5,
   safety 0.96, growth 20 (10^4 on the first step), etamxf 0.3 after 2
   error-test failures, etamin 0.1, etacf 0.25, fixed-step window [1, 1.5] */
static void adapt_defaults(orc_integrator* I)
{
  I->ctype = ORC_CTRL_PID;
  set_ctrl_params(I);
  I->bias = 1.5;
  I->safety = 0.96; I->growth = 20.0; I->etamx1 = 10000.0; I->etamxf = 0.3;
  I->etamin = 0.1; I->etacf = 0.25; I->lbound = 1.0; I->ubound = 1.5;
  I->small_nef = 2; I->maxnef = 7; I->maxncf = 10;
}

orc_integrator* orc_create(orc_problem* p, int method, double t0, const double* y0)
{
  orc_integrator* I = (orc_integrator*)calloc(1, sizeof(orc_integrator));
  if (!I) return NULL;
  I->P = p;
  I->method = method;
  I->n = p->n;
  I->rtol = 1.0e-5;
  I->atol = 1.0e-10;
  I->tn = t0;
  I->yn = vzalloc(I->n); I->ycur = vzalloc(I->n); I->fn = vzalloc(I->n);
  I->tv1 = vzalloc(I->n); I->tv2 = vzalloc(I->n); I->ewt = vzalloc(I->n);
  vcopy(I->yn, y0, I->n);
  adapt_defaults(I);
  I->mxstep = 500;
  /* LSRKStep defaults: safety 1.01, frequency 25, 200 stages */
  I->dom_eig_safety = 1.01; I->dom_eig_freq = 25; I->smax = 200;
  I->dom_eig_update = 1;
  I->dee_maxiters = 100; I->dee_rtol = 0.01;
  I->p_order = 2;
  /* DIRK defaults (main.cpp:27-46; ARKODE msbp 20, eplifac 0.05, nlscoef 0.1) */
  I->order = 3; I->linear = 1; I->maxl = 20; I->msbp = 20; I->precond = 1;
  I->maxcor = 3; I->nlscoef = 0.1;
  if (method == ORC_DIRK) orc_set_order(I, 3);
  return I;
}

void orc_free(orc_integrator* I)
{
  if (!I) return;
  free(I->yn); free(I->ycur); free(I->fn); free(I->tv1); free(I->tv2); free(I->ewt);
  free(I->dee_V); free(I->dee_q); free(I->dee_work);
  for (int k = 0; k < 4; k++) free(I->F[k]);
  free(I->sdata); free(I->zcur); free(I->fz); free(I->rhs); free(I->dz);
  free(I->r); free(I->pv); free(I->zv); free(I->Ap); free(I->jwork); free(I->jv);
  free(I);
}

void orc_set_tolerances(orc_integrator* I, double rtol, double atol)
{
  I->rtol = rtol;
  I->atol = atol;
}
void orc_set_fixed_step(orc_integrator* I, double h)
{
  I->hfixed = h;
  I->fixedstep = (h > 0.0);
}
void orc_set_init_step(orc_integrator* I, double h0) { I->h0u = h0; }
void orc_set_controller(orc_integrator* I, int type)
{
  I->ctype = type;
  set_ctrl_params(I);
}
void orc_set_max_num_steps(orc_integrator* I, long mxstep)
{
  I->mxstep = (mxstep == 0) ? 500 : mxstep;
}
void orc_set_internal_eig(orc_integrator* I, int on, int maxiters, double rtol)
{
  I->internal_eig = on;
  I->dee_maxiters = maxiters;
  I->dee_rtol = rtol;
  if (on && !I->dee_V) {
    I->dee_V = vzalloc(I->n); I->dee_q = vzalloc(I->n); I->dee_work = vzalloc(I->n);
  }
}
void orc_set_dom_eig_frequency(orc_integrator* I, long freq) { I->dom_eig_freq = freq; }
void orc_set_max_num_stages(orc_integrator* I, int smax) { I->smax = smax; }
void orc_set_dom_eig_safety(orc_integrator* I, double s) { I->dom_eig_safety = s; }
void orc_set_dee_preprocess(orc_integrator* I, int init_iters, int iters)
{
  I->dee_init_pre = init_iters;
  I->dee_pre = iters;
}
void orc_set_order(orc_integrator* I, int order)
{
  double A[16], b[4], bh[4], c[4];
  int s = orc_dirk_table(order, A, b, bh, c);
  if (s < 0) return;
  I->order = order;
  I->stages = s;
  memcpy(I->A, A, sizeof(A)); memcpy(I->b, b, sizeof(b));
  memcpy(I->bh, bh, sizeof(bh)); memcpy(I->c, c, sizeof(c));
  I->p_order = order - 1; /* embedding order drives the controller */
  for (int k = 0; k < 4; k++) {
    if (k < s && !I->F[k]) I->F[k] = vzalloc(I->n);
  }
  if (!I->sdata) {
    I->sdata = vzalloc(I->n); I->zcur = vzalloc(I->n); I->fz = vzalloc(I->n);
    I->rhs = vzalloc(I->n); I->dz = vzalloc(I->n); I->r = vzalloc(I->n);
    I->pv = vzalloc(I->n); I->zv = vzalloc(I->n); I->Ap = vzalloc(I->n);
    I->jwork = vzalloc(I->n); I->jv = vzalloc(I->n);
  }
}
void orc_set_linear(orc_integrator* I, int linear) { I->linear = linear; }
void orc_set_lin_solver(orc_integrator* I, int maxl, double epslin, int msbp, int precond)
{
  I->maxl = (maxl > 0) ? maxl : 5;
  I->epslin = epslin;
  I->msbp = (msbp == 0) ? 20 : msbp;
  I->precond = precond;
}
void orc_set_ls_reltol(orc_integrator* I, double reltol) { I->ls_reltol = reltol; }
void orc_set_jtimes_analytic(orc_integrator* I, int on) { I->jtimes_analytic = on ? 1 : 0; }

/* ---------------- step-size controllers ---------------- */
static double ctrl_estimate(const orc_integrator* I, double h, double dsm)
{
  const double ord = (double)(I->p_order + 1);
  const double e1 = fmax(I->bias * dsm, TINY);
  const double e2 = fmax(I->ep, TINY);
  const double e3 = fmax(I->epp, TINY);
  switch (I->ctype) {
  case ORC_CTRL_PID:
  case ORC_CTRL_PI:
  case ORC_CTRL_I:
    if (I->firststeps < 1) return h * pow(e1, -I->k1 / ord);
    if (I->firststeps < 2) return h * pow(e1, -I->k1 / ord) * pow(e2, I->k2 / ord);
    return h * pow(e1, -I->k1 / ord) * pow(e2, I->k2 / ord) * pow(e3, -I->k3 / ord);
  case ORC_CTRL_EXPGUS:
    if (I->firststeps < 1) return h * pow(e1, -1.0 / ord);
    return h * pow(e1, -I->k1 / ord) * pow(e1 / e2, -I->k2 / ord);
  case ORC_CTRL_IMPGUS:
    if (I->firststeps < 1) return h * pow(e1, -1.0 / ord);
    return h * (h / I->hp) * pow(e1, -I->k1 / ord) * pow(e1 / e2, -I->k2 / ord);
  default: {
    if (I->firststeps < 1) return h * pow(e1, -1.0 / ord);
    const double he = h * pow(e1, -0.367 / ord) * pow(e1 / e2, -0.268 / ord);
    const double hi = h * (h / I->hp) * pow(e1, -0.95 / ord) * pow(e1 / e2, -0.95 / ord);
    return fmin(he, hi);
  }
  }
}

static void ctrl_update(orc_integrator* I, double h, double dsm)
{
  I->epp = I->ep;
  I->ep = I->bias * dsm;
  I->hpp = I->hp;
  I->hp = h;
  I->firststeps++;
}

static double accept_eta(orc_integrator* I, double dsm)
{
  const double h = I->h;
  double hacc = I->safety * ctrl_estimate(I, h, dsm);
  const double etamax = (I->nst <= 1) ? I->etamx1 : I->growth;
  if (hacc > etamax * h) hacc = etamax * h;
  if (hacc > h * I->lbound * ONEMSM && hacc < h * I->ubound * ONEPSM) hacc = h;
  return hacc / h;
}

static double failure_eta(orc_integrator* I, double dsm, int nef)
{
  double eta = I->safety * ctrl_estimate(I, I->h, dsm) / I->h;
  eta = fmax(eta, I->etamin);
  if (nef >= I->small_nef) eta = fmin(eta, I->etamxf);
  return eta;
}

/* the step-size decisions of one attempt: an error-test failure shrinks h,
   an accepted step records the error in the controller's history and sets
   the next h (used by take_one_step and by the controller test hook) */
static void on_error_test_fail(orc_integrator* I, double dsm, int nef)
{
  I->h *= failure_eta(I, dsm, nef);
}

static void on_accept(orc_integrator* I, double dsm)
{
  const double eta = accept_eta(I, dsm);
  ctrl_update(I, I->h, dsm);
  I->dsm_last = dsm;
  I->h = I->h * eta;
}

/* test hook: the controller alone on the model error dsm = C h^q (q = p+1:
   the asymptotic local error of an embedding of order p).  Default
   adaptivity constants except the window [1, ubound] in which h is held;
   nattempts attempts from h0; per attempt the h tried, its dsm and whether
   it passed. */
int orc_test_controller(int ctype, int p_order, double C, double q, double h0, double ubound,
                        int nattempts, double* h_out, double* dsm_out, int* acc_out)
{
  orc_integrator I;
  memset(&I, 0, sizeof(I));
  adapt_defaults(&I);
  I.ctype = ctype;
  set_ctrl_params(&I);
  I.p_order = p_order;
  I.ubound = ubound;
  I.h = h0;
  int nef = 0;
  for (int k = 0; k < nattempts; k++) {
    const double dsm = C * pow(I.h, q);
    h_out[k] = I.h;
    dsm_out[k] = dsm;
    acc_out[k] = dsm <= 1.0;
    if (dsm > 1.0) {
      nef++;
      I.netf++;
      on_error_test_fail(&I, dsm, nef);
    } else {
      nef = 0;
      I.nst++;
      on_accept(&I, dsm);
    }
  }
  return 0;
}

/* ---------------- initial step (CVODE/ARKODE heuristic) ---------------- */
static int ydd_norm(orc_integrator* I, double hg, double* yddnrm)
{
  vlinsum(hg, I->fn, 1.0, I->yn, I->tv1, I->n);
  rhs(I, I->tn + hg, I->tv1, I->tv2);
  vlinsum(1.0, I->tv2, -1.0, I->fn, I->tv2, I->n);
  const double s = 1.0 / hg;
  for (int64_t i = 0; i < I->n; i++) I->tv2[i] = s * I->tv2[i];
  *yddnrm = orc_vwrmsnorm(I->tv2, I->ewt, I->n);
  return 0;
}

static double initial_step(orc_integrator* I, double tout)
{
  const double tdist = fabs(tout - I->tn);
  const double tround = DBL_EPSILON * fmax(fabs(I->tn), fabs(tout));
  const double hlb = 100.0 * tround;
  /* upper bound: hub_inv = max |f| / (0.1|y| + 1/ewt) */
  double hub_inv = 0.0;
  for (int64_t i = 0; i < I->n; i++) {
    const double d = 0.1 * fabs(I->yn[i]) + 1.0 * (1.0 / I->ewt[i]);
    const double q = fabs(I->fn[i]) / d;
    if (q > hub_inv) hub_inv = q;
  }
  double hub = 0.1 * tdist;
  if (hub * hub_inv > 1.0) hub = 1.0 / hub_inv;
  double hg = sqrt(hlb * hub);
  if (hub < hlb) return hg;
  int hnew_ok = 0;
  double hnew = hg, ydd = 0.0;
  for (int count = 1; count <= 4; count++) {
    ydd_norm(I, hg, &ydd);
    hnew = (ydd * hub * hub > 2.0) ? sqrt(2.0 / ydd) : sqrt(hg * hub);
    if (hnew_ok) {
      hg = hnew;
      break;
    }
    const double hrat = hnew / hg;
    if (hrat > 0.5 && hrat < 2.0) hnew_ok = 1;
    if (count > 1 && hrat > 2.0) {
      hnew = hg;
      hnew_ok = 1;
    }
    hg = hnew;
  }
  double h0 = 0.5 * hg;
  if (h0 < hlb) h0 = hlb;
  if (h0 > hub) h0 = hub;
  return h0;
}

/* ---------------- dominant eigenvalue ---------------- */
/* difference-quotient J v around (tn, yn, fn), sigma = 1/||v||_wrms */
static void dq_jv(orc_integrator* I, double t, const double* y, const double* fy,
                  const double* v, double* jv, double* work, long* counter)
{
  const double sig = 1.0 / orc_vwrmsnorm(v, I->ewt, I->n);
  vlinsum(sig, v, 1.0, y, work, I->n);
  rhs(I, t, work, jv);
  (*counter)++;
  const double siginv = 1.0 / sig;
  vlinsum(siginv, jv, -siginv, fy, jv, I->n);
}

static void dee_estimate(orc_integrator* I, double* lr, double* li)
{
  double* V = I->dee_V;
  double* q = I->dee_q;
  int npre = I->dee_pre;
  if (!I->dee_started) {
    orc_seed_vector(I->n, DEE_SEED, V);
    const double s = 1.0 / sqrt(orc_vdot(V, V, I->n));
    for (int64_t i = 0; i < I->n; i++) V[i] = s * V[i];
    I->dee_started = 1;
    npre = I->dee_init_pre;
  }
  for (int k = 0; k < npre; k++) {
    dq_jv(I, I->tn, I->yn, I->fn, V, q, I->dee_work, &I->nfe_dee);
    I->ndee_iters++;
    const double s = 1.0 / sqrt(orc_vdot(q, q, I->n));
    for (int64_t i = 0; i < I->n; i++) V[i] = s * q[i];
  }
  double lambda = 0.0, lold = 0.0;
  for (int k = 0; k < I->dee_maxiters; k++) {
    dq_jv(I, I->tn, I->yn, I->fn, V, q, I->dee_work, &I->nfe_dee);
    I->ndee_iters++;
    lambda = orc_vdot(q, V, I->n);
    const double s = 1.0 / sqrt(orc_vdot(q, q, I->n));
    for (int64_t i = 0; i < I->n; i++) V[i] = s * q[i];
    const double res = fabs(lambda - lold) / fabs(lambda);
    if (res < I->dee_rtol) break;
    lold = lambda;
  }
  *lr = lambda;
  *li = 0.0;
}

static void compute_dom_eig(orc_integrator* I)
{
  double lr, li;
  if (!I->fn_is_current) {
    rhs(I, I->tn, I->yn, I->fn);
    I->fn_is_current = 1;
  }
  if (I->internal_eig) dee_estimate(I, &lr, &li);
  else orc_dom_eig(I->P, &lr, &li);
  I->rho = sqrt(lr * lr + li * li) * I->dom_eig_safety;
  if (I->rho > I->rho_max) I->rho_max = I->rho;
  I->ndomeig++;
  I->dom_eig_update = 0;
  I->dom_eig_nst = I->nst;
}

/* ---------------- LSRK STS step ---------------- */
static int lsrk_takestep(orc_integrator* I, double* dsm)
{
  const double h = I->h, tn = I->tn;
  const int64_t n = I->n;
  if (I->dom_eig_update) compute_dom_eig(I);

  double ss;
  if (I->method == ORC_RKC) ss = ceil(sqrt(1.54 * fabs(h) * I->rho));
  else ss = ceil((sqrt(9.0 + 8.0 * fabs(h) * I->rho) - 1.0) / 2.0);
  ss = fmax(ss, 2.0);
  if (ss >= (double)I->smax) {
    if (I->fixedstep) return ORC_MAX_STAGE_LIMIT_FAIL;
    /* largest step whose stage count is smax-1 (a step sized for smax
       stages can round back up to smax and never make progress) */
    const double sm = (double)I->smax - 1.0;
    const double hmax = (I->method == ORC_RKC)
                          ? I->safety * (sm * sm) / (1.54 * I->rho)
                          : I->safety * (sm * sm + sm - 2.0) / (2.0 * I->rho);
    I->eta = hmax / h;
    return ORC_RETRY_STEP;
  }
  const int s = (int)ss;
  I->last_stages = s;
  if (s > I->stage_max) I->stage_max = s;
  if (!I->fn_is_current) {
    rhs(I, tn, I->yn, I->fn);
    I->fn_is_current = 1;
  }

  double* Yjm2 = I->tv2; /* Y_{j-2} */
  double* Yjm1 = I->tv1; /* Y_{j-1} */
  double* Yj = I->ycur;
  double c[5];
  const double* X[5];
  vcopy(Yjm2, I->yn, n);

  if (I->method == ORC_RKC) {
    const double w0 = 1.0 + 2.0 / (13.0 * ((double)s * (double)s));
    const double temp1 = w0 * w0 - 1.0;
    const double temp2 = sqrt(temp1);
    const double arg = (double)s * log(w0 + temp2);
    const double w1 =
      sinh(arg) * temp1 / (cosh(arg) * (double)s * temp2 - w0 * sinh(arg));
    double bjm1 = 1.0 / ((2.0 * w0) * (2.0 * w0)), bjm2 = bjm1;
    double mus = w1 * bjm1;
    vlinsum(1.0, I->yn, h * mus, I->fn, Yjm1, n);
    double thjm2 = 0.0, thjm1 = mus;
    double zjm1 = w0, zjm2 = 1.0, dzjm1 = 1.0, dzjm2 = 0.0, d2zjm1 = 0.0, d2zjm2 = 0.0;
    for (int j = 2; j <= s; j++) {
      const double zj = 2.0 * w0 * zjm1 - zjm2;
      const double dzj = 2.0 * w0 * dzjm1 - dzjm2 + 2.0 * zjm1;
      const double d2zj = 2.0 * w0 * d2zjm1 - d2zjm2 + 4.0 * dzjm1;
      const double bj = d2zj / (dzj * dzj);
      const double ajm1 = 1.0 - zjm1 * bjm1;
      const double mu = 2.0 * w0 * bj / bjm1;
      const double nu = -bj / bjm2;
      mus = mu * w1 / w0;
      rhs(I, tn + h * thjm1, Yjm1, Yj);
      c[0] = mus * h; X[0] = Yj;
      c[1] = nu; X[1] = Yjm2;
      c[2] = 1.0 - mu - nu; X[2] = I->yn;
      c[3] = mu; X[3] = Yjm1;
      c[4] = -mus * ajm1 * h; X[4] = I->fn;
      orc_vlinearcombination(5, c, X, Yj, n);
      if (j < s) {
        double* t = Yjm2;
        Yjm2 = Yjm1; Yjm1 = Yj; Yj = t;
        const double thj = mu * thjm1 + nu * thjm2 + mus * (1.0 - ajm1);
        thjm2 = thjm1; thjm1 = thj;
        bjm2 = bjm1; bjm1 = bj;
        zjm2 = zjm1; zjm1 = zj;
        dzjm2 = dzjm1; dzjm1 = dzj;
        d2zjm2 = d2zjm1; d2zjm1 = d2zj;
      }
    }
  } else {
    const double w1 = 4.0 / (((double)s + 2.0) * ((double)s - 1.0));
    double bjm2 = 1.0 / 3.0, bjm1 = bjm2;
    double mus = w1 * bjm1;
    double cjm1 = mus;
    vlinsum(1.0, I->yn, h * mus, I->fn, Yjm1, n);
    for (int j = 2; j <= s; j++) {
      const double dj = (double)j;
      const double temj = (dj + 2.0) * (dj - 1.0);
      const double bj = temj / (2.0 * dj * (dj + 1.0));
      const double ajm1 = 1.0 - bjm1;
      const double mu = (2.0 * dj - 1.0) / dj * (bj / bjm1);
      const double nu = -(dj - 1.0) / dj * (bj / bjm2);
      mus = w1 * mu;
      const double cj = temj * w1 / 4.0;
      rhs(I, tn + h * cjm1, Yjm1, Yj);
      c[0] = mus * h; X[0] = Yj;
      c[1] = nu; X[1] = Yjm2;
      c[2] = 1.0 - mu - nu; X[2] = I->yn;
      c[3] = mu; X[3] = Yjm1;
      c[4] = -mus * ajm1 * h; X[4] = I->fn;
      orc_vlinearcombination(5, c, X, Yj, n);
      if (j < s) {
        double* t = Yjm2;
        Yjm2 = Yjm1; Yjm1 = Yj; Yj = t;
        cjm1 = cj;
        bjm2 = bjm1; bjm1 = bj;
      }
    }
  }
  /* the final stage may live in any of the three buffers: normalise so that
     ycur holds Y_s and tv1/tv2 are scratch */
  if (Yj != I->ycur) {
    double* t = I->ycur;
    if (I->tv1 == Yj) I->tv1 = t;
    else I->tv2 = t;
    I->ycur = Yj;
  }

  if (!I->fixedstep) {
    rhs(I, tn + h, I->ycur, I->tv2);
    c[0] = 0.8; X[0] = I->yn;
    c[1] = -0.8; X[1] = I->ycur;
    c[2] = 0.4 * h; X[2] = I->fn;
    c[3] = 0.4 * h; X[3] = I->tv2;
    orc_vlinearcombination(4, c, X, I->tv1, n);
    *dsm = orc_vwrmsnorm(I->tv1, I->ewt, n);
  } else *dsm = 0.0;

  if (I->fixedstep || *dsm <= 1.0) {
    if (!I->fixedstep) {
      double* t = I->fn;
      I->fn = I->tv2; I->tv2 = t;
      I->fn_is_current = 1;
    } else I->fn_is_current = 0;
    if (I->dom_eig_freq > 0 && (I->nst + 1 - I->dom_eig_nst) >= I->dom_eig_freq)
      I->dom_eig_update = 1;
  } else if (I->dom_eig_nst != I->nst) I->dom_eig_update = 1;
  return ORC_SUCCESS;
}

/* ---------------- DIRK: Newton + PCG + Jacobi ---------------- */
/* (I - gamma J) v; J v by the difference quotient of arkLs (the reference sets
   no Jtimes, main.cpp:302-329), or exactly as the 5-point stencil (the forcing
   does not depend on u) when the analytic option is on */
static void atimes(orc_integrator* I, double t, double gamma, const double* v, double* Av)
{
  if (I->jtimes_analytic) orc_stencil(I->P, v, I->jv);
  else dq_jv(I, t, I->zcur, I->fz, v, I->jv, I->jwork, &I->nfe_ls);
  I->njtimes++;
  vlinsum(1.0, v, -gamma, I->jv, Av, I->n);
}

/* precond 1: Jacobi (preconditioner_jacobi.cpp); 2: one multigrid V-cycle
   (orc_mg.c) for the gamma of the last setup */
static void psolve(orc_integrator* I, const double* r, double* z)
{
  if (I->precond == 2) {
    orc_mg_apply(I->P, I->gammap, r, z);
    I->npsolve++;
  } else if (I->precond) {
    orc_psolve(I->P, r, z);
    I->npsolve++;
  } else vcopy(z, r, I->n);
}

/* a non-finite scalar in a Krylov iteration: the solve cannot recover (no
   residual test can pass once a NaN is in r), so it stops with a message
   instead of iterating to the cap */
static int krylov_nonfinite(const char* what, double v)
{
  if (isfinite(v)) return 0;
  fprintf(stderr, "[oracle PCG] non-finite %s (%g): linear solve failed\n", what, v);
  return 1;
}

/* solves (I - gamma J) x = b, zero initial guess; 0 converged, 1 residual
   reduced, 2 failed, -1 a non-finite value (unrecoverable) */
static int pcg(orc_integrator* I, double t, double gamma, const double* b, double* x)
{
  const int64_t n = I->n;
  memset(x, 0, sizeof(double) * (size_t)n);
  vcopy(I->r, b, n);
  /* relative-tolerance mode measures the unweighted 2-norm */
  const int rel = (I->ls_reltol > 0.0);
  const double bnorm = rel ? sqrt(orc_vdot(I->r, I->r, n)) : vwl2(I->r, I->ewt, n);
  const double eplifac = (I->epslin > 0.0) ? I->epslin : 0.05;
  const double delta = (I->ls_reltol > 0.0) ? I->ls_reltol * bnorm
                                            : eplifac * I->nlscoef * sqrt((double)n);
  if (krylov_nonfinite("|b|", bnorm)) return -1;
  if (bnorm <= delta) return 0;
  psolve(I, I->r, I->zv);
  vcopy(I->pv, I->zv, n);
  double rz = orc_vdot(I->r, I->zv, n);
  if (krylov_nonfinite("r.z", rz)) return -1;
  double rnorm = bnorm;
  for (int l = 1; l <= I->maxl; l++) {
    atimes(I, t, gamma, I->pv, I->Ap);
    const double pap = orc_vdot(I->pv, I->Ap, n);
    const double alpha = rz / pap;
    if (krylov_nonfinite("alpha = r.z / p.Ap", alpha)) return -1;
    vlinsum(1.0, x, alpha, I->pv, x, n);
    vlinsum(1.0, I->r, -alpha, I->Ap, I->r, n);
    I->nli++;
    rnorm = rel ? sqrt(orc_vdot(I->r, I->r, n)) : vwl2(I->r, I->ewt, n);
    if (krylov_nonfinite("|r|", rnorm)) return -1;
    if (rnorm <= delta) return 0;
    psolve(I, I->r, I->zv);
    const double rznew = orc_vdot(I->r, I->zv, n);
    if (krylov_nonfinite("r.z", rznew)) return -1;
    const double beta = rznew / rz;
    rz = rznew;
    vlinsum(1.0, I->zv, beta, I->pv, I->pv, n);
  }
  return (rnorm < bnorm) ? 1 : 2;
}

static int newton(orc_integrator* I, double t, double gamma, int force_setup)
{
  const int64_t n = I->n;
  vcopy(I->zcur, I->yn, n);
  if (I->precond) {
    const int need = !I->psetup_done || force_setup ||
                     fabs(gamma / I->gammap - 1.0) > 0.2 ||
                     (I->msbp > 0 && I->nst >= I->nstlp + I->msbp);
    if (need) {
      if (I->precond != 2) orc_psetup(I->P, gamma);
      I->npsetup++;
      I->gammap = gamma;
      I->nstlp = I->nst;
      I->psetup_done = 1;
    }
  }
  double crate = 1.0, delp = 0.0;
  for (int m = 0; m < I->maxcor; m++) {
    rhs(I, t, I->zcur, I->fz);
    const double c3[3] = {1.0, -1.0, gamma};
    const double* X3[3] = {I->sdata, I->zcur, I->fz};
    orc_vlinearcombination(3, c3, X3, I->rhs, n);
    const int ls = pcg(I, t, gamma, I->rhs, I->dz);
    if (ls < 0) return ORC_LSOLVE_FAIL;
    if (ls == 2) {
      I->nlsfail++;
      return ORC_CONV_RECOVER;
    }
    vlinsum(1.0, I->zcur, 1.0, I->dz, I->zcur, n);
    I->nni++;
    if (I->linear) return 0;
    const double del = orc_vwrmsnorm(I->dz, I->ewt, n);
    if (m > 0) crate = fmax(0.3 * crate, del / delp);
    const double dcon = del * fmin(1.0, crate) / I->nlscoef;
    if (dcon <= 1.0) return 0;
    if (m > 0 && del > 2.3 * delp) return ORC_CONV_RECOVER;
    delp = del;
  }
  return ORC_CONV_RECOVER;
}

static int dirk_takestep(orc_integrator* I, double* dsm, int force_setup)
{
  const double h = I->h, tn = I->tn;
  const int64_t n = I->n;
  const int s = I->stages;
  double c[5];
  const double* X[5];
  for (int i = 0; i < s; i++) {
    const double ti = tn + I->c[i] * h;
    if (i == 0 && I->A[0] == 0.0) {
      if (!I->fn_is_current) {
        rhs(I, tn, I->yn, I->fn);
        I->fn_is_current = 1;
      }
      vcopy(I->F[0], I->fn, n);
      continue;
    }
    c[0] = 1.0; X[0] = I->yn;
    for (int j = 0; j < i; j++) {
      c[j + 1] = h * I->A[4 * i + j];
      X[j + 1] = I->F[j];
    }
    orc_vlinearcombination(i + 1, c, X, I->sdata, n);
    const double gamma = h * I->A[4 * i + i];
    const int ret = newton(I, ti, gamma, force_setup);
    force_setup = 0;
    if (ret != 0) return ret;
    rhs(I, ti, I->zcur, I->F[i]);
  }
  c[0] = 1.0; X[0] = I->yn;
  for (int j = 0; j < s; j++) {
    c[j + 1] = h * I->b[j];
    X[j + 1] = I->F[j];
  }
  orc_vlinearcombination(s + 1, c, X, I->ycur, n);
  if (!I->fixedstep) {
    for (int j = 0; j < s; j++) {
      c[j] = h * (I->b[j] - I->bh[j]);
      X[j] = I->F[j];
    }
    orc_vlinearcombination(s, c, X, I->tv1, n);
    *dsm = orc_vwrmsnorm(I->tv1, I->ewt, n);
  } else *dsm = 0.0;
  I->fn_is_current = 0;
  return ORC_SUCCESS;
}

/* ---------------- time stepping ---------------- */
static int take_one_step(orc_integrator* I, double tout)
{
  int nef = 0, ncf = 0, force_setup = 0;
  double dsm = 0.0;
  int clipped;
  for (;;) {
    clipped = 0;
    if (I->tn + I->h >= tout) {
      I->h = tout - I->tn;
      clipped = 1;
    }
    I->attempts++;
    int ret;
    if (I->method == ORC_DIRK) ret = dirk_takestep(I, &dsm, force_setup);
    else ret = lsrk_takestep(I, &dsm);
    force_setup = 0;
    if (ret == ORC_RETRY_STEP) {
      I->h *= I->eta;
      I->nstlim++;
      continue;
    }
    if (ret == ORC_CONV_RECOVER) {
      I->ncfn++;
      ncf++;
      if (ncf >= I->maxncf) return ORC_CONV_FAILURE;
      I->h *= I->etacf;
      force_setup = 1;
      continue;
    }
    if (ret < 0) return ret;
    if (!I->fixedstep && dsm > 1.0) {
      I->netf++;
      nef++;
      if (nef >= I->maxnef) return ORC_ERR_FAILURE;
      on_error_test_fail(I, dsm, nef);
      continue;
    }
    break;
  }
  I->hold = I->h;
  I->tn = clipped ? tout : I->tn + I->h;
  double* t = I->yn;
  I->yn = I->ycur;
  I->ycur = t;
  I->nst++;
  ewt_set(I, I->yn);
  if (!I->fixedstep) on_accept(I, dsm);
  else I->h = I->hfixed;
  return ORC_SUCCESS;
}

int orc_evolve(orc_integrator* I, double tout, double* yout, double* tret)
{
  int ret = ORC_SUCCESS;
  if (!I->initialized) {
    if (ewt_set(I, I->yn)) return ORC_ILL_INPUT;
    if (I->fixedstep) I->h = I->hfixed;
    else if (I->h0u != 0.0) I->h = I->h0u;
    else {
      if (!I->fn_is_current) {
        rhs(I, I->tn, I->yn, I->fn);
        I->fn_is_current = 1;
      }
      I->h = initial_step(I, tout);
    }
    I->h0_used = I->h;
    I->initialized = 1;
  }
  long nstloc = 0;
  while (I->tn < tout) {
    if (I->mxstep > 0 && nstloc >= I->mxstep) {
      ret = ORC_TOO_MUCH_WORK;
      break;
    }
    ret = take_one_step(I, tout);
    if (ret < 0) break;
    nstloc++;
  }
  if (yout) vcopy(yout, I->yn, I->n);
  if (tret) *tret = I->tn;
  return ret;
}

int orc_get_stats(const orc_integrator* I, long* out, int nout)
{
  long v[ORC_ST_COUNT];
  v[ORC_ST_NST] = I->nst; v[ORC_ST_ATTEMPTS] = I->attempts; v[ORC_ST_NFE] = I->nfe;
  v[ORC_ST_NETF] = I->netf; v[ORC_ST_NSTLIM] = I->nstlim;
  v[ORC_ST_NDOMEIG] = I->ndomeig; v[ORC_ST_NDEE_ITERS] = I->ndee_iters;
  v[ORC_ST_STAGE_MAX] = I->stage_max; v[ORC_ST_LAST_STAGES] = I->last_stages;
  v[ORC_ST_NFE_DEE] = I->nfe_dee; v[ORC_ST_NLI] = I->nli;
  v[ORC_ST_NPSETUP] = I->npsetup; v[ORC_ST_NPSOLVE] = I->npsolve;
  v[ORC_ST_NNI] = I->nni; v[ORC_ST_NCFN] = I->ncfn; v[ORC_ST_NLSFAIL] = I->nlsfail;
  v[ORC_ST_NJTIMES] = I->njtimes; v[ORC_ST_NFE_LS] = I->nfe_ls;
  int m = nout < ORC_ST_COUNT ? nout : ORC_ST_COUNT;
  for (int k = 0; k < m; k++) out[k] = v[k];
  return m;
}

double orc_get_current_h(const orc_integrator* I) { return I->h; }

/* threads of the OpenMP timing build (the checker build: always 1) */
int orc_set_threads(int n)
{
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}
double orc_get_h0_used(const orc_integrator* I) { return I->h0_used; }
double orc_get_last_dsm(const orc_integrator* I) { return I->dsm_last; }
double orc_get_spectral_radius(const orc_integrator* I) { return I->rho; }

/* ARKodePrintAllStats(SUN_OUTPUTFORMAT_TABLE)-style lines; the tokens are the
   ones diffusion_2D/runtests-diffusion2d.py:36-54 parses */
void orc_print_stats(const orc_integrator* I, FILE* fp)
{
  fprintf(fp, "Current time                 = %.16g\n", I->tn);
  fprintf(fp, "Steps                        = %ld\n", I->nst);
  fprintf(fp, "Step attempts                = %ld\n", I->attempts);
  fprintf(fp, "Stability limited steps      = %ld\n", I->nstlim);
  fprintf(fp, "Error test fails             = %ld\n", I->netf);
  fprintf(fp, "NLS step fails               = %ld\n", I->ncfn);
  fprintf(fp, "Initial step size            = %.16g\n", I->h0_used);
  fprintf(fp, "Last step size               = %.16g\n", I->hold);
  fprintf(fp, "Current step size            = %.16g\n", I->h);
  if (I->method == ORC_DIRK) {
    fprintf(fp, "Explicit RHS fn evals        = %d\n", 0);
    fprintf(fp, "Implicit RHS fn evals        = %ld\n", I->nfe - I->nfe_ls);
    fprintf(fp, "NLS iters                    = %ld\n", I->nni);
    fprintf(fp, "NLS fails                    = %ld\n", I->ncfn);
    fprintf(fp, "LS setups                    = %ld\n", I->npsetup);
    fprintf(fp, "LS RHS fn evals              = %ld\n", I->nfe_ls);
    fprintf(fp, "Prec setup evals             = %ld\n", I->npsetup);
    fprintf(fp, "Prec solves                  = %ld\n", I->npsolve);
    fprintf(fp, "LS iters                     = %ld\n", I->nli);
    fprintf(fp, "LS fails                     = %ld\n", I->nlsfail);
    fprintf(fp, "Jac-times evals              = %ld\n", I->njtimes);
  } else {
    fprintf(fp, "RHS fn evals                 = %ld\n", I->nfe);
    fprintf(fp, "Number of dom_eig updates    = %ld\n", I->ndomeig);
    fprintf(fp, "Max. num. of stages used     = %ld\n", I->stage_max);
    fprintf(fp, "Max. spectral radius         = %.16g\n", I->rho_max);
    fprintf(fp, "DEE power iterations         = %ld\n", I->ndee_iters);
  }
}

/* test hook: solve (I - gamma J) x = b with the DIRK path's PCG, DQ Jv around
   (t, yn) and the Jacobi preconditioner, relative tolerance reltol */
int orc_test_linear_solve(orc_integrator* I, double t, double gamma, const double* b,
                          double* x, double reltol, int maxl, long* iters)
{
  if (I->method != ORC_DIRK) return -1;
  if (ewt_set(I, I->yn)) return -1;
  vcopy(I->zcur, I->yn, I->n);
  rhs(I, t, I->zcur, I->fz);
  if (I->precond) {
    orc_psetup(I->P, gamma);
    I->npsetup++;
  }
  const double rt = I->ls_reltol;
  const int ml = I->maxl;
  I->ls_reltol = reltol;
  I->maxl = maxl;
  const long nli0 = I->nli;
  const int ret = pcg(I, t, gamma, b, x);
  I->ls_reltol = rt;
  I->maxl = ml;
  if (iters) *iters = I->nli - nli0;
  return ret;
}
