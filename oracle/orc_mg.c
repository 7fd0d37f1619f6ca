/* CPU oracle of the geometric-multigrid preconditioner (test infrastructure
 * only: used by tests/ to check the GPU V-cycle; never by the product).
 *
 * The reference's alternative preconditioner is one hypre PFMG V-cycle with
 * zero initial guess, symmetric red/black Gauss-Seidel, 2 pre and 2 post
 * sweeps (preconditioner_hypre.cpp:51-102, 142-160; settings
 * diffusion_2D.hpp:187-193).  hypre is absent, so this is the build's own
 * V-cycle with those settings, restated serially:
 *   level 0: the operator of the implicit system, A = I - gamma J with J the
 *            5-point stencil of the RHS face tables fx, fy (diffusion.cpp:48-53;
 *            f[k] is the face west of node k);
 *   Dirichlet (manufactured problem): coarsening n -> n/2 vertex-centred
 *            (coarse node I = fine node 2I+1, so for even n the last coarse
 *            node is the last fine node), until a side would drop below 3;
 *            coarse face I sits on fine node 2I: coefficient
 *            (f[2I] + f[min(2I+1, n)]) / 8 (mean of the fine faces beside it,
 *            times h^2/H^2 = 1/4); transfers see zero ghosts;
 *   periodic (the reference's problem, every side even): coarse node I =
 *            fine node 2I, n -> n/2 while the coarse sides stay even and at
 *            least 4; coarse face I spans fine faces 2I-1 and 2I:
 *            (f[2I-1] + f[2I]) / 8 with f[-1] = f[n-1]; transfers and
 *            smoothing wrap around;
 *   restriction: full weighting (1/16)[1 2 1; 2 4 2; 1 2 1] = P^T / 4;
 *   prolongation: bilinear;
 *   smoothing: nu pairs of red (i+j even) then black sweeps before; after,
 *            nu pairs of black then red (the adjoint: relax type 2, the
 *            V-cycle is symmetric) or red then black again (relax type 3,
 *            diffusion_2D.cpp:137-151); the coarsest level gets 4 pairs in
 *            each order.  Default relax 2, nu 2.
 * Expression orders are fixed here and in the GPU kernels (no FMA), so the
 * GPU V-cycle equals this one bit for bit. */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

#define MG_MAXL 24

typedef struct {
  int nl, nu1, nu2, nuc, relax, per;
  int64_t nx[MG_MAXL], ny[MG_MAXL];
  double *fx[MG_MAXL], *fy[MG_MAXL];
  double *x[MG_MAXL], *b[MG_MAXL], *r[MG_MAXL];
  double gamma;
} orc_mg;

static void mg_free(orc_mg* M)
{
  for (int l = 0; l < M->nl; l++) {
    free(M->fx[l]);
    free(M->fy[l]);
    free(M->x[l]);
    free(M->b[l]);
    free(M->r[l]);
  }
  free(M);
}

static orc_mg* mg_setup(const double* fx0, const double* fy0, int64_t nx, int64_t ny, double gamma,
                        int relax, int nu, int per)
{
  orc_mg* M = calloc(1, sizeof(orc_mg));
  M->per = per;
  M->nu1 = nu;
  M->nu2 = nu;
  M->nuc = 4;
  M->relax = relax;
  M->gamma = gamma;
  int l = 0;
  for (;;) {
    M->nx[l] = nx;
    M->ny[l] = ny;
    M->fx[l] = malloc(sizeof(double) * (nx + 1));
    M->fy[l] = malloc(sizeof(double) * (ny + 1));
    if (l == 0) {
      memcpy(M->fx[0], fx0, sizeof(double) * (nx + 1));
      memcpy(M->fy[0], fy0, sizeof(double) * (ny + 1));
    } else {
      const double* px = M->fx[l - 1];
      const double* py = M->fy[l - 1];
      const int64_t fxn = M->nx[l - 1], fyn = M->ny[l - 1];
      if (per) {
        for (int64_t i = 0; i <= nx; i++) M->fx[l][i] = (px[i > 0 ? 2 * i - 1 : fxn - 1] + px[2 * i]) / 8.0;
        for (int64_t j = 0; j <= ny; j++) M->fy[l][j] = (py[j > 0 ? 2 * j - 1 : fyn - 1] + py[2 * j]) / 8.0;
      } else {
        for (int64_t i = 0; i <= nx; i++)
          M->fx[l][i] = (px[2 * i < fxn ? 2 * i : fxn] + px[2 * i + 1 < fxn ? 2 * i + 1 : fxn]) / 8.0;
        for (int64_t j = 0; j <= ny; j++)
          M->fy[l][j] = (py[2 * j < fyn ? 2 * j : fyn] + py[2 * j + 1 < fyn ? 2 * j + 1 : fyn]) / 8.0;
      }
    }
    M->x[l] = calloc(nx * ny, sizeof(double));
    M->b[l] = calloc(nx * ny, sizeof(double));
    M->r[l] = calloc(nx * ny, sizeof(double));
    l++;
    const int64_t cx = nx / 2, cy = ny / 2;
    if (per ? (cx % 2 || cy % 2 || cx < 4 || cy < 4) : (cx < 3 || cy < 3)) break;
    if (l == MG_MAXL) break;
    nx = cx;
    ny = cy;
  }
  M->nl = l;
  return M;
}

/* off-diagonal sum and diagonal of A at (i, j) of level l */
static void mg_row(const orc_mg* M, int l, const double* x, int64_t i, int64_t j, double* off, double* d)
{
  const int64_t nx = M->nx[l], ny = M->ny[l];
  const double* fx = M->fx[l];
  const double* fy = M->fy[l];
  const int64_t k = j * nx + i;
  double xw, xe, xs, xn;
  if (M->per) {
    xw = x[j * nx + (i > 0 ? i - 1 : nx - 1)];
    xe = x[j * nx + (i < nx - 1 ? i + 1 : 0)];
    xs = x[(j > 0 ? j - 1 : ny - 1) * nx + i];
    xn = x[(j < ny - 1 ? j + 1 : 0) * nx + i];
  } else {
    xw = i > 0 ? x[k - 1] : 0.0;
    xe = i < nx - 1 ? x[k + 1] : 0.0;
    xs = j > 0 ? x[k - nx] : 0.0;
    xn = j < ny - 1 ? x[k + nx] : 0.0;
  }
  *off = ((fx[i] * xw + fx[i + 1] * xe) + fy[j] * xs) + fy[j + 1] * xn;
  *d = 1.0 + M->gamma * ((fx[i] + fx[i + 1]) + (fy[j] + fy[j + 1]));
}

static void mg_sweep(orc_mg* M, int l, int color)
{
  const int64_t nx = M->nx[l], ny = M->ny[l];
  double* x = M->x[l];
  const double* b = M->b[l];
  for (int64_t j = 0; j < ny; j++)
    for (int64_t i = (j + color) & 1; i < nx; i += 2) {
      double off, d;
      mg_row(M, l, x, i, j, &off, &d);
      x[j * nx + i] = (b[j * nx + i] + M->gamma * off) / d;
    }
}

static void mg_vcycle(orc_mg* M, int l)
{
  const int64_t nx = M->nx[l], ny = M->ny[l];
  memset(M->x[l], 0, sizeof(double) * nx * ny);
  if (l == M->nl - 1) {
    for (int s = 0; s < M->nuc; s++) {
      mg_sweep(M, l, 0);
      mg_sweep(M, l, 1);
    }
    for (int s = 0; s < M->nuc; s++) {
      mg_sweep(M, l, M->relax == 2 ? 1 : 0);
      mg_sweep(M, l, M->relax == 2 ? 0 : 1);
    }
    return;
  }
  for (int s = 0; s < M->nu1; s++) {
    mg_sweep(M, l, 0);
    mg_sweep(M, l, 1);
  }
  /* residual */
  double* x = M->x[l];
  double* r = M->r[l];
  for (int64_t j = 0; j < ny; j++)
    for (int64_t i = 0; i < nx; i++) {
      double off, d;
      mg_row(M, l, x, i, j, &off, &d);
      double t = d * x[j * nx + i];
      t = t - M->gamma * off;
      r[j * nx + i] = M->b[l][j * nx + i] - t;
    }
  /* full-weighting restriction */
  const int64_t cx = M->nx[l + 1], cy = M->ny[l + 1];
  double* bc = M->b[l + 1];
  const int per = M->per;
  for (int64_t J = 0; J < cy; J++)
    for (int64_t I = 0; I < cx; I++) {
      const int64_t i = per ? 2 * I : 2 * I + 1, j = per ? 2 * J : 2 * J + 1;
#define R_(a, c)                                                                              \
  (per ? r[(((c) + ny) % ny) * nx + (((a) + nx) % nx)] : (((a) < nx && (c) < ny) ? r[(c) * nx + (a)] : 0.0))
      double t = 4.0 * R_(i, j);
      t = t + 2.0 * ((R_(i - 1, j) + R_(i + 1, j)) + (R_(i, j - 1) + R_(i, j + 1)));
      t = t + ((R_(i - 1, j - 1) + R_(i + 1, j - 1)) + (R_(i - 1, j + 1) + R_(i + 1, j + 1)));
#undef R_
      bc[J * cx + I] = t * 0.0625;
    }
  mg_vcycle(M, l + 1);
  /* bilinear prolongation, added */
  const double* xc = M->x[l + 1];
  for (int64_t j = 0; j < ny; j++)
    for (int64_t i = 0; i < nx; i++) {
#define XC(a, c)                                                                              \
  (per ? xc[(((c) + cy) % cy) * cx + (((a) + cx) % cx)]                                      \
       : (((a) >= 0 && (a) < cx && (c) >= 0 && (c) < cy) ? xc[(c) * cx + (a)] : 0.0))
      double v;
      /* a fine index coincides with coarse index I (Dirichlet: odd i,
         I = i/2; periodic: even i, I = i/2) or lies between I0 and I0+1
         (Dirichlet: I0 = i/2 - 1; periodic: I0 = i/2, wrapping) */
      const int cxo = per ? !(i & 1) : (int)(i & 1), cyo = per ? !(j & 1) : (int)(j & 1);
      const int64_t I = per ? i / 2 : (cxo ? i / 2 : i / 2 - 1);
      const int64_t J = per ? j / 2 : (cyo ? j / 2 : j / 2 - 1);
      if (cxo && cyo) v = XC(I, J);
      else if (cyo) v = 0.5 * (XC(I, J) + XC(I + 1, J));
      else if (cxo) v = 0.5 * (XC(I, J) + XC(I, J + 1));
      else v = 0.25 * ((XC(I, J) + XC(I + 1, J)) + (XC(I, J + 1) + XC(I + 1, J + 1)));
#undef XC
      x[j * nx + i] = x[j * nx + i] + v;
    }
  for (int s = 0; s < M->nu2; s++) {
    mg_sweep(M, l, M->relax == 2 ? 1 : 0);
    mg_sweep(M, l, M->relax == 2 ? 0 : 1);
  }
}

/* z = one V-cycle (smoother relax type 2|3, nu sweep pairs) applied to r
   for A = I - gamma J of the problem's grid; returns the number of levels,
   or -1 for a periodic grid with an odd side */
int orc_mg_apply_opts(const orc_problem* p, double gamma, int relax, int nu, const double* r, double* z)
{
  const int per = (p->kind == ORC_CEDA_PERIODIC);
  if (per && (p->nx % 2 || p->ny % 2)) return -1;
  orc_mg* M = mg_setup(p->fx, p->fy, p->nx, p->ny, gamma, relax, nu, per);
  memcpy(M->b[0], r, sizeof(double) * p->nx * p->ny);
  mg_vcycle(M, 0);
  memcpy(z, M->x[0], sizeof(double) * p->nx * p->ny);
  const int nl = M->nl;
  mg_free(M);
  return nl;
}

int orc_mg_apply(const orc_problem* p, double gamma, const double* r, double* z)
{
  return orc_mg_apply_opts(p, gamma, 2, 2, r, z);
}
