/* CPU oracle of the learned preconditioner (test infrastructure only: used by
 * tests/ to check the GPU network kernels; never by the product).
 *
 * Restates, serially in FP64, the network of include/learned_prec.h and its
 * training objective:
 *   s = RMS(r); x_0 = bf16(r/s); x_{l+1} = bf16(relu(W_l * x_l + b_l));
 *   u = out_scale * (W_L * x_L + b_L);  d = u - e/s;
 *   loss = mean_b d^T A d / ((e.r)/s^2),  A = I - rx Dxx - ry Dyy (unit-spacing second
 *   differences, zero BC; rx == ry evaluated as I - rx*L with the 5-point L)
 * and the gradient of the loss with respect to every parameter by reverse
 * accumulation.  With emu = 1 the BF16 roundings of the GPU path are emulated
 * at the same points (weights; network input; every stored activation; the
 * output gradient and every stored activation gradient), accumulation is
 * FP64 (the GPU accumulates in FP32, so agreement is to rounding, not bitwise).
 * With emu = 0 every rounding is skipped: the exact FP64 function, whose
 * gradient the tests pin by finite differences.
 *
 * The symmetric network (orc_lp_sym_lossgrad) is K followed by its adjoint,
 * with linear layers and no biases:
 *   h_0 = bf16(r/s); h_{l+1} = bf16(W_l * h_l); u_nl = h_nl;
 *   u_l = bf16(W_l^T * u_{l+1}); u = out_scale * (W_0^T * u_1)
 * and the same loss; each layer's gradient sums its adjoint-layer and its
 * layer contributions.
 *
 * Parity pinning: forward convolution against a brute-force numpy
 * convolution; gradient against central differences of the emu = 0 loss; the
 * symmetric network's symmetry (x.N y = y.N x) and positivity at emu = 0 and
 * its adjoint convolution against the transpose of the dense matrix of the
 * forward one; no reference counterpart exists for this component. */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

#define LP_C 16
#define LP_MAXL 8

static double bfr(double x, int emu)
{
  if (!emu) return x;
  float f = (float)x;
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7FFFFFFFu) > 0x7F800000u) return x; /* NaN stays NaN */
  u += 0x7FFFu + ((u >> 16) & 1u);
  u &= 0xFFFF0000u;
  memcpy(&f, &u, 4);
  return (double)f;
}

typedef struct {
  int nl, np;
  int dil[LP_MAXL], cin[LP_MAXL], cout[LP_MAXL], woff[LP_MAXL], boff[LP_MAXL];
} lpnet;

static void lp_layout(lpnet* N, int nl, const int* dil, int sym)
{
  int off = 0;
  N->nl = nl;
  for (int l = 0; l < nl; l++) {
    N->dil[l] = dil[l];
    N->cin[l] = (l == 0) ? 1 : LP_C;
    N->cout[l] = (l == nl - 1 && !sym) ? 1 : LP_C;
    N->woff[l] = off;
    off += N->cout[l] * N->cin[l] * 9;
    N->boff[l] = off;
    off += N->cout[l];
  }
  N->np = off;
}

int orc_lp_num_params(int nl, const int* dil)
{
  lpnet N;
  lp_layout(&N, nl, dil, 0);
  return N.np;
}

int orc_lp_sym_num_params(int nl, const int* dil)
{
  lpnet N;
  lp_layout(&N, nl, dil, 1);
  return N.np;
}

/* activation of image b, pixel (x, y), channel c in an [nb][ny][nx][C] array */
#define IDX(b, y, x, c, C) ((((int64_t)(b) * ny + (y)) * nx + (x)) * (C) + (c))

/* pre-activation of one layer: out[p][co] = b[co] + sum_t sum_ci W[co][ci][t] in[p+off_t][ci] */
static void conv_fwd(int nb, int nx, int ny, int cin, int cout, int d, const double* W, const double* bias,
                     const double* in, double* out)
{
  for (int b = 0; b < nb; b++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++)
        for (int co = 0; co < cout; co++) {
          double acc = 0.0;
          for (int t = 0; t < 9; t++) {
            const int xs = x + (t % 3 - 1) * d, ys = y + (t / 3 - 1) * d;
            if (xs < 0 || xs >= nx || ys < 0 || ys >= ny) continue;
            for (int ci = 0; ci < cin; ci++) acc += W[(co * cin + ci) * 9 + t] * in[IDX(b, ys, xs, ci, cin)];
          }
          out[IDX(b, y, x, co, cout)] = acc + (bias ? bias[co] : 0.0);
        }
}

/* adjoint convolution: out[p][ci] = sum_t sum_co W[co][ci][t] in[p - off_t][co] */
static void conv_adj(int nb, int nx, int ny, int cin, int cout, int d, const double* W, const double* in,
                     double* out)
{
  for (int b = 0; b < nb; b++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++)
        for (int ci = 0; ci < cin; ci++) {
          double acc = 0.0;
          for (int t = 0; t < 9; t++) {
            const int xs = x - (t % 3 - 1) * d, ys = y - (t / 3 - 1) * d;
            if (xs < 0 || xs >= nx || ys < 0 || ys >= ny) continue;
            for (int co = 0; co < cout; co++) acc += W[(co * cin + ci) * 9 + t] * in[IDX(b, ys, xs, co, cout)];
          }
          out[IDX(b, y, x, ci, cin)] = acc;
        }
}

/* gw[co][ci][t] += sum_p dy[p][co] x[p + off_t][ci] */
static void wgrad_acc(int nb, int nx, int ny, int cin, int cout, int d, const double* dy, const double* x,
                      double* gw)
{
  for (int b = 0; b < nb; b++)
    for (int y = 0; y < ny; y++)
      for (int xx = 0; xx < nx; xx++)
        for (int co = 0; co < cout; co++) {
          const double g = dy[IDX(b, y, xx, co, cout)];
          for (int t = 0; t < 9; t++) {
            const int xs = xx + (t % 3 - 1) * d, ys = y + (t / 3 - 1) * d;
            if (xs < 0 || xs >= nx || ys < 0 || ys >= ny) continue;
            for (int ci = 0; ci < cin; ci++) gw[(co * cin + ci) * 9 + t] += g * x[IDX(b, ys, xs, ci, cin)];
          }
        }
}

static void apply_A(int nb, int nx, int ny, double rx, double ry, const double* x, double* y)
{
  for (int b = 0; b < nb; b++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        const int64_t k = ((int64_t)b * ny + j) * nx + i;
        const double c = x[k];
        const double w = i > 0 ? x[k - 1] : 0.0, e = i < nx - 1 ? x[k + 1] : 0.0;
        const double s = j > 0 ? x[k - nx] : 0.0, no = j < ny - 1 ? x[k + nx] : 0.0;
        y[k] = rx == ry ? c - rx * ((((w + e) + s) + no) - 4.0 * c)
                        : c - (rx * ((w + e) - 2.0 * c) + ry * ((s + no) - 2.0 * c));
      }
}

/* loss of a batch and (grad != NULL) its parameter gradient; u (optional)
   receives the network output in normalised units */
double orc_lp_lossgrad(int nl, const int* dil, const float* params, double out_scale, int nb, int nx,
                       int ny, double ratio_x, double ratio_y, const double* r, const double* e, int emu,
                       double* grad, double* u_out)
{
  lpnet N;
  lp_layout(&N, nl, dil, 0);
  const int64_t npix = (int64_t)nx * ny, tot = npix * nb;
  double* W = malloc(sizeof(double) * N.np);  /* rounded weights, FP32 biases */
  for (int l = 0; l < nl; l++) {
    for (int k = N.woff[l]; k < N.boff[l]; k++) W[k] = bfr((double)params[k], emu);
    for (int k = N.boff[l]; k < N.boff[l] + N.cout[l]; k++) W[k] = (double)params[k];
  }
  double* inv = malloc(sizeof(double) * nb);
  for (int b = 0; b < nb; b++) {
    double s2 = 0.0;
    for (int64_t k = 0; k < npix; k++) s2 += r[b * npix + k] * r[b * npix + k];
    double s = sqrt(s2 / (double)npix);
    if (!(s > 0.0)) s = 1.0;
    inv[b] = 1.0 / s;
  }
  /* x[l] = input of layer l ([nb][ny][nx][cin]) */
  double* x[LP_MAXL + 1];
  x[0] = malloc(sizeof(double) * tot);
  for (int64_t k = 0; k < tot; k++) x[0][k] = bfr(r[k] * inv[k / npix], emu);
  double* u = malloc(sizeof(double) * tot);
  for (int l = 0; l < nl; l++) {
    const int last = (l == nl - 1);
    double* pre = malloc(sizeof(double) * tot * N.cout[l]);
    conv_fwd(nb, nx, ny, N.cin[l], N.cout[l], N.dil[l], W + N.woff[l], W + N.boff[l], x[l], pre);
    if (last) {
      for (int64_t k = 0; k < tot; k++) u[k] = pre[k] * out_scale;
      free(pre);
    } else {
      for (int64_t k = 0; k < tot * LP_C; k++) pre[k] = bfr(pre[k] > 0.0 ? pre[k] : 0.0, emu);
      x[l + 1] = pre;
    }
  }
  if (u_out) memcpy(u_out, u, sizeof(double) * tot);
  double* dd = malloc(sizeof(double) * tot);
  double* ad = malloc(sizeof(double) * tot);
  for (int64_t k = 0; k < tot; k++) dd[k] = u[k] - e[k] * inv[k / npix];
  apply_A(nb, nx, ny, ratio_x, ratio_y, dd, ad);
  double loss = 0.0;
  double* coef = malloc(sizeof(double) * nb);
  for (int b = 0; b < nb; b++) {
    double s1 = 0.0, s2 = 0.0;
    for (int64_t k = 0; k < npix; k++) {
      s1 += dd[b * npix + k] * ad[b * npix + k];
      s2 += e[b * npix + k] * r[b * npix + k];
    }
    const double er = s2 * inv[b] * inv[b];
    loss += s1 / er;
    coef[b] = 2.0 * out_scale / ((double)nb * er);
  }
  loss /= (double)nb;
  if (grad) {
    /* gradient of the last pre-activation, rounded as the GPU stores it */
    double* dy = malloc(sizeof(double) * tot);
    for (int64_t k = 0; k < tot; k++) dy[k] = bfr(coef[k / npix] * ad[k], emu);
    for (int l = nl - 1; l >= 0; l--) {
      const int ci_n = N.cin[l], co_n = N.cout[l], d = N.dil[l];
      double* gw = grad + N.woff[l];
      double* gb = grad + N.boff[l];
      for (int k = 0; k < co_n * ci_n * 9; k++) gw[k] = 0.0;
      for (int co = 0; co < co_n; co++) gb[co] = 0.0;
      for (int b = 0; b < nb; b++)
        for (int y = 0; y < ny; y++)
          for (int xx = 0; xx < nx; xx++)
            for (int co = 0; co < co_n; co++) {
              const double g = dy[IDX(b, y, xx, co, co_n)];
              gb[co] += g;
              for (int t = 0; t < 9; t++) {
                const int xs = xx + (t % 3 - 1) * d, ys = y + (t / 3 - 1) * d;
                if (xs < 0 || xs >= nx || ys < 0 || ys >= ny) continue;
                for (int ci = 0; ci < ci_n; ci++) gw[(co * ci_n + ci) * 9 + t] += g * x[l][IDX(b, ys, xs, ci, ci_n)];
              }
            }
      if (l == 0) break;
      /* gradient of the layer input, times relu' of the layer below */
      double* dx = malloc(sizeof(double) * tot * ci_n);
      const double* Wl = W + N.woff[l];
      for (int b = 0; b < nb; b++)
        for (int y = 0; y < ny; y++)
          for (int xx = 0; xx < nx; xx++)
            for (int ci = 0; ci < ci_n; ci++) {
              double acc = 0.0;
              for (int t = 0; t < 9; t++) {
                const int xs = xx - (t % 3 - 1) * d, ys = y - (t / 3 - 1) * d;
                if (xs < 0 || xs >= nx || ys < 0 || ys >= ny) continue;
                for (int co = 0; co < co_n; co++) acc += Wl[(co * ci_n + ci) * 9 + t] * dy[IDX(b, ys, xs, co, co_n)];
              }
              const double m = x[l][IDX(b, y, xx, ci, ci_n)] > 0.0 ? acc : 0.0;
              dx[IDX(b, y, xx, ci, ci_n)] = bfr(m, emu);
            }
      free(dy);
      dy = dx;
    }
    free(dy);
  }
  for (int l = 0; l < nl; l++) free(x[l]);
  free(W);
  free(inv);
  free(u);
  free(dd);
  free(ad);
  free(coef);
  return loss;
}

/* the symmetric network: loss, gradient (grad != NULL) and normalised output */
double orc_lp_sym_lossgrad(int nl, const int* dil, const float* params, double out_scale, int nb, int nx,
                           int ny, double ratio_x, double ratio_y, const double* r, const double* e, int emu,
                           double* grad, double* u_out)
{
  lpnet N;
  lp_layout(&N, nl, dil, 1);
  const int64_t npix = (int64_t)nx * ny, tot = npix * nb;
  double* W = malloc(sizeof(double) * N.np);
  for (int k = 0; k < N.np; k++) W[k] = bfr((double)params[k], emu);
  double* inv = malloc(sizeof(double) * nb);
  for (int b = 0; b < nb; b++) {
    double s2 = 0.0;
    for (int64_t k = 0; k < npix; k++) s2 += r[b * npix + k] * r[b * npix + k];
    double s = sqrt(s2 / (double)npix);
    if (!(s > 0.0)) s = 1.0;
    inv[b] = 1.0 / s;
  }
  /* h[l] = input of layer l, u[l] = input of the adjoint of layer l - 1 */
  double* h[LP_MAXL + 1];
  double* u[LP_MAXL + 1];
  h[0] = malloc(sizeof(double) * tot);
  for (int64_t k = 0; k < tot; k++) h[0][k] = bfr(r[k] * inv[k / npix], emu);
  for (int l = 0; l < nl; l++) {
    h[l + 1] = malloc(sizeof(double) * tot * LP_C);
    conv_fwd(nb, nx, ny, N.cin[l], LP_C, N.dil[l], W + N.woff[l], NULL, h[l], h[l + 1]);
    for (int64_t k = 0; k < tot * LP_C; k++) h[l + 1][k] = bfr(h[l + 1][k], emu);
  }
  u[nl] = h[nl];
  for (int l = nl - 1; l >= 1; l--) {
    u[l] = malloc(sizeof(double) * tot * LP_C);
    conv_adj(nb, nx, ny, LP_C, LP_C, N.dil[l], W + N.woff[l], u[l + 1], u[l]);
    for (int64_t k = 0; k < tot * LP_C; k++) u[l][k] = bfr(u[l][k], emu);
  }
  double* z = malloc(sizeof(double) * tot);
  conv_adj(nb, nx, ny, 1, LP_C, N.dil[0], W + N.woff[0], u[1], z);
  for (int64_t k = 0; k < tot; k++) z[k] *= out_scale;
  if (u_out) memcpy(u_out, z, sizeof(double) * tot);
  double* dd = malloc(sizeof(double) * tot);
  double* ad = malloc(sizeof(double) * tot);
  for (int64_t k = 0; k < tot; k++) dd[k] = z[k] - e[k] * inv[k / npix];
  apply_A(nb, nx, ny, ratio_x, ratio_y, dd, ad);
  double loss = 0.0;
  double* coef = malloc(sizeof(double) * nb);
  for (int b = 0; b < nb; b++) {
    double s1 = 0.0, s2 = 0.0;
    for (int64_t k = 0; k < npix; k++) {
      s1 += dd[b * npix + k] * ad[b * npix + k];
      s2 += e[b * npix + k] * r[b * npix + k];
    }
    const double er = s2 * inv[b] * inv[b];
    loss += s1 / er;
    coef[b] = 2.0 * out_scale / ((double)nb * er);
  }
  loss /= (double)nb;
  if (grad) {
    for (int k = 0; k < N.np; k++) grad[k] = 0.0;
    /* adjoint layers 0 .. nl-1: output gradient g, input u[l+1] */
    double* g = malloc(sizeof(double) * tot);
    for (int64_t k = 0; k < tot; k++) g[k] = bfr(coef[k / npix] * ad[k], emu);
    for (int l = 0; l < nl; l++) {
      const int ci_n = N.cin[l];
      wgrad_acc(nb, nx, ny, ci_n, LP_C, N.dil[l], u[l + 1], g, grad + N.woff[l]);
      double* gn = malloc(sizeof(double) * tot * LP_C);
      conv_fwd(nb, nx, ny, ci_n, LP_C, N.dil[l], W + N.woff[l], NULL, g, gn);
      for (int64_t k = 0; k < tot * LP_C; k++) gn[k] = bfr(gn[k], emu);
      free(g);
      g = gn;
    }
    /* layers nl-1 .. 0: output gradient g, input h[l] */
    for (int l = nl - 1; l >= 0; l--) {
      const int ci_n = N.cin[l];
      wgrad_acc(nb, nx, ny, ci_n, LP_C, N.dil[l], g, h[l], grad + N.woff[l]);
      if (l == 0) break;
      double* gn = malloc(sizeof(double) * tot * LP_C);
      conv_adj(nb, nx, ny, LP_C, LP_C, N.dil[l], W + N.woff[l], g, gn);
      for (int64_t k = 0; k < tot * LP_C; k++) gn[k] = bfr(gn[k], emu);
      free(g);
      g = gn;
    }
    free(g);
  }
  for (int l = 0; l <= nl; l++) free(h[l]);
  for (int l = 1; l < nl; l++) free(u[l]);
  free(W);
  free(inv);
  free(z);
  free(dd);
  free(ad);
  free(coef);
  return loss;
}
