/* ---------------------------------------------------------------------------
 * CPU oracle, problem layer — TEST INFRASTRUCTURE ONLY (see orc.h).
 *
 * ceda-periodic follows the reference code:
 *   grid        diffusion_2D.cpp:155-160 (dx=(xu-xl)/(nx-1)), defaults
 *               diffusion_2D.hpp:66-93
 *   periodicity diffusion_2D.cpp:242-327 (every HaveNbr* true, one rank
 *               exchanges with itself: Wrecv <- column nx-1, Erecv <- column
 *               0, Srecv <- row ny-1, Nrecv <- row 0; tags :423-497,
 *               packing buffers.cpp:20-43)
 *   stencil     diffusion.cpp:9-209 (f zeroed at :31, formula :48-53 and the
 *               same c,w,e,s,n association order in every face block)
 *   coeffs      diffusion_2D.cpp:887-897
 *   IC          initial.cpp:20-48
 *   Jacobi      preconditioner_jacobi.cpp:9-62 (faces at (is+i)dx, no xl
 *               offset, kept as written)
 *   dom_eig     main.cpp:536-550
 * mms-dirichlet is the manufactured problem of diffusion_2D/README.md:11-32
 * on [xl,xu]x[yl,yu] (default [0,1]^2): the state holds the nx*ny interior
 * nodes x_i = xl+(i+1)dx, dx=(xu-xl)/(nx+1), with zero Dirichlet ghosts.
 * ------------------------------------------------------------------------- */
#include "orc.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846

void orc_problem_defaults(orc_problem* p, int kind)
{
  memset(p, 0, sizeof(*p));
  p->kind = kind;
  p->kx = 1.0;
  p->ky = 1.0;
  p->tf = 1.0;
  if (kind == ORC_CEDA_PERIODIC) {
    p->xl = -ORC_PI; p->xu = ORC_PI; p->yl = -6.0; p->yu = 6.0;
    p->nx = 64; p->ny = 64;
  } else {
    p->xl = 0.0; p->xu = 1.0; p->yl = 0.0; p->yu = 1.0;
    p->nx = 128; p->ny = 128;
  }
}

/* diffusion_2D.cpp:887-897 */
static double coeff(double k, int inhomogeneous, double s)
{
  if (inhomogeneous) return k * (1.0 + 0.99 * sin(s));
  return k;
}

int orc_problem_setup(orc_problem* p)
{
  int64_t i;
  if (p->nx < 2 || p->ny < 2) return -1;
  if (p->kind == ORC_MMS_DIRICHLET && p->inhomogeneous) return -1;
  p->n = p->nx * p->ny;
  if (p->kind == ORC_CEDA_PERIODIC) {
    p->dx = (p->xu - p->xl) / (double)(p->nx - 1);
    p->dy = (p->yu - p->yl) / (double)(p->ny - 1);
  } else {
    p->dx = (p->xu - p->xl) / (double)(p->nx + 1);
    p->dy = (p->yu - p->yl) / (double)(p->ny + 1);
  }
  p->fx = (double*)malloc(sizeof(double) * (size_t)(p->nx + 1));
  p->fy = (double*)malloc(sizeof(double) * (size_t)(p->ny + 1));
  p->px = (double*)malloc(sizeof(double) * (size_t)(p->nx + 1));
  p->py = (double*)malloc(sizeof(double) * (size_t)(p->ny + 1));
  p->sx = (double*)calloc((size_t)p->nx, sizeof(double));
  p->sy = (double*)calloc((size_t)p->ny, sizeof(double));
  p->cx = (double*)calloc((size_t)p->nx, sizeof(double));
  p->cy = (double*)calloc((size_t)p->ny, sizeof(double));
  p->diag = (double*)malloc(sizeof(double) * (size_t)p->n);
  if (!p->fx || !p->fy || !p->px || !p->py || !p->sx || !p->sy || !p->cx ||
      !p->cy || !p->diag)
    return -1;

  /* face k sits at xl + (k - 0.5) dx: the west face of node k and the east
     face of node k-1 (diffusion.cpp:43-46 evaluates both to the same double) */
  const double dx2 = p->dx * p->dx, dy2 = p->dy * p->dy;
  const double xo = (p->kind == ORC_CEDA_PERIODIC) ? 0.0 : 1.0;
  for (i = 0; i <= p->nx; i++) {
    const double xf = p->xl + ((double)i + xo - 0.5) * p->dx;
    p->fx[i] = coeff(p->kx, p->inhomogeneous, xf) / dx2;
    p->px[i] = coeff(p->kx, p->inhomogeneous, (double)i * p->dx) / dx2;
  }
  for (i = 0; i <= p->ny; i++) {
    const double yf = p->yl + ((double)i + xo - 0.5) * p->dy;
    p->fy[i] = coeff(p->ky, p->inhomogeneous, yf) / dy2;
    p->py[i] = coeff(p->ky, p->inhomogeneous, (double)i * p->dy) / dy2;
  }
  if (p->kind == ORC_MMS_DIRICHLET) {
    /* cos^2 - sin^2 evaluated as cos(2 pi x): no sin/cos pair of one
       argument, which a compiler may fuse into sincos() (it rounds
       differently in rare cases, breaking bitwise reproducibility) */
    const double tpp = 2.0 * ORC_PI * ORC_PI;
    for (i = 0; i < p->nx; i++) {
      const double x = p->xl + (double)(i + 1) * p->dx;
      const double s = sin(ORC_PI * x);
      p->sx[i] = s * s;
      p->cx[i] = p->kx * (-tpp * cos(2.0 * ORC_PI * x));
    }
    for (i = 0; i < p->ny; i++) {
      const double y = p->yl + (double)(i + 1) * p->dy;
      const double s = sin(ORC_PI * y);
      p->sy[i] = s * s;
      p->cy[i] = p->ky * (-tpp * cos(2.0 * ORC_PI * y));
    }
  }
  for (i = 0; i < p->n; i++) p->diag[i] = 1.0;
  p->nrhs = 0;
  return 0;
}

void orc_problem_free(orc_problem* p)
{
  free(p->fx); free(p->fy); free(p->px); free(p->py);
  free(p->sx); free(p->sy); free(p->cx); free(p->cy); free(p->diag);
  p->fx = p->fy = p->px = p->py = p->sx = p->sy = p->cx = p->cy = p->diag = NULL;
}

/* initial.cpp:20-48 (periodic: every node, since all HaveNbr* are true);
   MMS: u(0) = sin^2(pi x) sin^2(pi y) (diffusion_2D/README.md:20) */
void orc_initial(const orc_problem* p, double* u)
{
  int64_t i, j;
  if (p->kind == ORC_CEDA_PERIODIC) {
    for (j = 0; j < p->ny; j++)
      for (i = 0; i < p->nx; i++) {
        const double x = p->xl + (double)i * p->dx;
        const double y = p->yl + (double)j * p->dy;
        u[j * p->nx + i] =
          (1.0 + 0.3 * sin(2.0 * x)) / sqrt(5.5 * ORC_PI) * exp(-(y * y) / 5.5);
      }
  } else {
    for (j = 0; j < p->ny; j++)
      for (i = 0; i < p->nx; i++) u[j * p->nx + i] = p->sx[i] * p->sy[j];
  }
}

/* README.md:32: u = sin^2(pi x) sin^2(pi y) cos^2(pi t) */
void orc_exact(const orc_problem* p, double t, double* u)
{
  int64_t i, j;
  double T1, T2;
  orc_mms_time(t, &T1, &T2);
  for (j = 0; j < p->ny; j++)
    for (i = 0; i < p->nx; i++) u[j * p->nx + i] = (p->sx[i] * p->sy[j]) * T2;
}

/* time factors of the README.md:28 forcing:
   b = T1 sx sy + T2 (kx Cx sy + ky Cy sx), T1 = -2 pi sin(pi t) cos(pi t)
   = -pi sin(2 pi t), T2 = cos^2(pi t) (no sin/cos pair of one argument) */
void orc_mms_time(double t, double* T1, double* T2)
{
  const double c = cos(ORC_PI * t);
  *T1 = -ORC_PI * sin(2.0 * ORC_PI * t);
  *T2 = c * c;
}

/* 5-point stencil, one association order for every cell (diffusion.cpp:48-53):
   t = -((Dw+De)+(Ds+Dn))*uc; t += Dw*uw; t += De*ue; t += Ds*us; t += Dn*un;
   f = 0 + t (f was zeroed, :31).  Periodic: wrap; Dirichlet: ghost = 0. */
static inline double cell(double Dw, double De, double Ds, double Dn, double uc,
                          double uw, double ue, double us, double un)
{
  double t = -((Dw + De) + (Ds + Dn)) * uc;
  t = t + Dw * uw;
  t = t + De * ue;
  t = t + Ds * us;
  t = t + Dn * un;
  return 0.0 + t;
}

void orc_stencil(const orc_problem* p, const double* u, double* f)
{
  const int64_t nx = p->nx, ny = p->ny;
  const int per = (p->kind == ORC_CEDA_PERIODIC);
  int64_t j;
#pragma omp parallel for schedule(static)
  for (j = 0; j < ny; j++) {
    int64_t i;
    const double Ds = p->fy[j], Dn = p->fy[j + 1];
    const double* uc = u + j * nx;
    const double* us = (j > 0) ? u + (j - 1) * nx : (per ? u + (ny - 1) * nx : NULL);
    const double* un = (j < ny - 1) ? u + (j + 1) * nx : (per ? u : NULL);
    for (i = 0; i < nx; i++) {
      const double Dw = p->fx[i], De = p->fx[i + 1];
      const double uw = (i > 0) ? uc[i - 1] : (per ? uc[nx - 1] : 0.0);
      const double ue = (i < nx - 1) ? uc[i + 1] : (per ? uc[0] : 0.0);
      const double vs = us ? us[i] : 0.0;
      const double vn = un ? un[i] : 0.0;
      f[j * nx + i] = cell(Dw, De, Ds, Dn, uc[i], uw, ue, vs, vn);
    }
  }
}

/* ARKRhsFn: diffusion() (diffusion_2D.cpp:23-35) = laplacian(), plus the MMS
   forcing b(t,x,y) for the Dirichlet problem:
   f = stencil + ((T1*sx)*sy + T2*(cx*sy + cy*sx)) */
int orc_rhs(orc_problem* p, double t, const double* u, double* f)
{
  orc_stencil(p, u, f);
  if (p->kind == ORC_MMS_DIRICHLET) {
    double T1, T2;
    int64_t i, j;
    orc_mms_time(t, &T1, &T2);
#pragma omp parallel for schedule(static) private(i)
    for (j = 0; j < p->ny; j++)
      for (i = 0; i < p->nx; i++) {
        const double b =
          (T1 * p->sx[i]) * p->sy[j] + T2 * (p->cx[i] * p->sy[j] + p->cy[j] * p->sx[i]);
        f[j * p->nx + i] = f[j * p->nx + i] + b;
      }
  }
  p->nrhs++;
  return 0;
}

/* preconditioner_jacobi.cpp:9-44: diag = 1/(1 - gamma*(-((Dw+De)+(Ds+Dn))))
   with the faces evaluated at (is+i)dx, (is+i+1)dx (no xl offset).  The MMS
   problem uses its stencil faces (constant coefficients, identical values). */
void orc_psetup(orc_problem* p, double gamma)
{
  const double* ax = (p->kind == ORC_CEDA_PERIODIC) ? p->px : p->fx;
  const double* ay = (p->kind == ORC_CEDA_PERIODIC) ? p->py : p->fy;
  int64_t i, j;
  for (j = 0; j < p->ny; j++)
    for (i = 0; i < p->nx; i++) {
      const double d = -((ax[i] + ax[i + 1]) + (ay[j] + ay[j + 1]));
      p->diag[j * p->nx + i] = 1.0 / (1.0 - gamma * d);
    }
}

/* preconditioner_jacobi.cpp:58: z = diag .* r (N_VProd) */
void orc_psolve(const orc_problem* p, const double* r, double* z)
{
  int64_t i;
  for (i = 0; i < p->n; i++) z[i] = p->diag[i] * r[i];
}

/* main.cpp:536-550 */
void orc_dom_eig(const orc_problem* p, double* lr, double* li)
{
  const double a = p->kx / p->dx / p->dx, b = p->ky / p->dy / p->dy;
  *lr = -8.0 * (a > b ? a : b);
  *li = 0.0;
}

/* SplitMix64 (Steele, Lea & Flood 2014): the deterministic start vector of the
   power iteration, reproduced bit-for-bit by the device generator. */
uint64_t orc_splitmix64(uint64_t x)
{
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

void orc_seed_vector(int64_t n, uint64_t seed, double* v)
{
  int64_t i;
  for (i = 0; i < n; i++)
    v[i] = (double)(orc_splitmix64(seed + (uint64_t)i) >> 11) * 0x1.0p-53 - 0.5;
}
