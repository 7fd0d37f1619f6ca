// diffusion_2D problem layer on gfx950: stencil RHS, Jacobi preconditioner,
// initial/exact states and the fused STS stage kernel.
//
// Reference: diffusion_2D/diffusion.cpp:9-209 (laplacian), diffusion_2D.cpp:
// 23-35 (diffusion), :155-160 and :242-327 (grid, one-rank periodicity),
// :887-897 (coefficients), initial.cpp:20-48, preconditioner_jacobi.cpp:9-62,
// main.cpp:536-550 (dom_eig); manufactured problem diffusion_2D/README.md:11-32.
//
// Stencil design.  The 5-point update is HBM-bound (8 B read + 8 B written per
// cell at FP64).  Each lane owns two adjacent columns (16-byte loads of the
// centre, south and north rows, 8-byte loads of the west/east neighbours that
// hit L1), a wavefront covers 128 columns of one row and the four wavefronts
// of a workgroup take four adjacent rows, so north/south rows are shared in L1.
// Tiles are enumerated x-fastest and the workgroup->tile map is XCD-aware:
// the hardware deals workgroups round-robin to the 8 XCDs, so workgroup b is
// given tile (b mod 8)*ceil(T/8)+(b div 8) and every XCD sweeps a contiguous
// band of rows, keeping the row above/below in its own L2.  The periodic
// self-exchange of the reference (Wrecv <- column nx-1 etc.) is the index wrap;
// the Dirichlet ghosts are zeros.  Face coefficients Kx(x_{i-1/2})/dx^2 are a
// 1-D table (the coefficient is separable), so no transcendental is evaluated
// per cell.  The expression order is exactly the reference's:
//   t = -((Dw+De)+(Ds+Dn))*uc + Dw*uw + De*ue + Ds*us + Dn*un ; f = 0 + t
// and the library is built without FMA contraction, which makes the result
// bitwise identical to a serial CPU evaluation.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>

#include "diffusion_internal.hpp"
#include "stencil_ops.hpp"

using namespace ceda;

namespace {

constexpr double kPi = 3.14159265358979323846;

struct KArgs {
  const double* u;
  double* f;
  const double *fx, *fy, *sx, *sy, *cx, *cy;
  const double* coef;  // device {c0..c4, T1, T2}; nullptr: T1/T2 by value
  double T1, T2;
  const double *yjm2, *yn, *fn;  // fused stage operands
  const StageEntry* ent;          // non-null: operands come from the table
  int64_t nx, ny;
  int tiles_x, rows;
};

// FUSED = false: f = F(u).  FUSED = true: f = stage combination with F(u).
template <int CPL, bool PER, bool MMS, bool FUSED>
__global__ __launch_bounds__(kBlock) void k_stencil(const KArgs a)
{
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const unsigned B = gridDim.x, b = blockIdx.x;
  const unsigned q = B >> 3, r = B & 7u, xcd = b & 7u, loc = b >> 3;
  const unsigned tile = xcd * q + (xcd < r ? xcd : r) + loc;
  const int tx = (int)(tile % (unsigned)a.tiles_x), ty = (int)(tile / (unsigned)a.tiles_x);
  const int64_t nx = a.nx, ny = a.ny;
  const int64_t i = (int64_t)tx * (kWave * CPL) + (int64_t)lane * CPL;
  if (i >= nx) return;

  double T1 = a.T1, T2 = a.T2;
  double c[5] = {0, 0, 0, 0, 0};
  const double* __restrict__ u = a.u;
  double* __restrict__ fout = a.f;
  const double* yjm2 = a.yjm2;
  const double* yn = a.yn;
  const double* fn = a.fn;
  const double* coef = a.coef;
  if (a.ent) {
    u = a.ent->yjm1;
    fout = a.ent->out;
    yjm2 = a.ent->yjm2;
    yn = a.ent->yn;
    fn = a.ent->fn;
    T1 = a.ent->T1;
    T2 = a.ent->T2;
    if (FUSED) {
#pragma unroll
      for (int k = 0; k < 5; k++) c[k] = a.ent->c[k];
    }
  } else if (coef) {
    T1 = coef[5];
    T2 = coef[6];
    if (FUSED) {
#pragma unroll
      for (int k = 0; k < 5; k++) c[k] = coef[k];
    }
  }
  const double Dw0 = a.fx[i], De0 = a.fx[i + 1];
  const double De1 = (CPL == 2) ? a.fx[i + 2] : 0.0;
  const int64_t iw = (i > 0) ? i - 1 : (PER ? nx - 1 : -1);
  const int64_t ie = (i + CPL < nx) ? i + CPL : (PER ? 0 : -1);
  double sx0 = 0, sx1 = 0, cx0 = 0, cx1 = 0;
  if (MMS) {
    sx0 = a.sx[i];
    cx0 = a.cx[i];
    if (CPL == 2) {
      sx1 = a.sx[i + 1];
      cx1 = a.cx[i + 1];
    }
  }
  for (int rr = 0; rr < a.rows; rr++) {
    const int64_t j = (int64_t)ty * (4 * a.rows) + rr * 4 + wave;
    if (j >= ny) break;
    const int64_t js = (j > 0) ? j - 1 : (PER ? ny - 1 : -1);
    const int64_t jn = (j < ny - 1) ? j + 1 : (PER ? 0 : -1);
    const double Ds = a.fy[j], Dn = a.fy[j + 1];
    const int64_t o = j * nx + i;
    const double uw = (iw >= 0) ? u[j * nx + iw] : 0.0;
    const double ue = (ie >= 0) ? u[j * nx + ie] : 0.0;
    if (CPL == 2) {
      const double2 C = *reinterpret_cast<const double2*>(u + o);
      const double2 S = (js >= 0) ? *reinterpret_cast<const double2*>(u + js * nx + i) : make_double2(0.0, 0.0);
      const double2 N = (jn >= 0) ? *reinterpret_cast<const double2*>(u + jn * nx + i) : make_double2(0.0, 0.0);
      double f0 = cell(Dw0, De0, Ds, Dn, C.x, uw, C.y, S.x, N.x);
      double f1 = cell(De0, De1, Ds, Dn, C.y, C.x, ue, S.y, N.y);
      if (MMS) {
        const double syj = a.sy[j], cyj = a.cy[j];
        f0 = f0 + forcing(T1, T2, sx0, cx0, syj, cyj);
        f1 = f1 + forcing(T1, T2, sx1, cx1, syj, cyj);
      }
      if (FUSED) {
        const double2 A2 = *reinterpret_cast<const double2*>(yjm2 + o);
        const double2 Y0 = *reinterpret_cast<const double2*>(yn + o);
        const double2 G = *reinterpret_cast<const double2*>(fn + o);
        f0 = lc5(c, f0, A2.x, Y0.x, C.x, G.x);
        f1 = lc5(c, f1, A2.y, Y0.y, C.y, G.y);
      }
      *reinterpret_cast<double2*>(fout + o) = make_double2(f0, f1);
    } else {
      const double C = u[o];
      const double S = (js >= 0) ? u[js * nx + i] : 0.0;
      const double N = (jn >= 0) ? u[jn * nx + i] : 0.0;
      double f0 = cell(Dw0, De0, Ds, Dn, C, uw, ue, S, N);
      if (MMS) f0 = f0 + forcing(T1, T2, sx0, cx0, a.sy[j], a.cy[j]);
      if (FUSED) f0 = lc5(c, f0, yjm2[o], yn[o], C, fn[o]);
      fout[o] = f0;
    }
  }
}


// y-marching variant: a workgroup owns 256*CPL columns x `rows` consecutive
// rows; each lane walks down its columns keeping the south/centre/north rows
// in registers (one new row load per output row), gets west/east neighbours
// from the adjacent lanes by wavefront shuffles (only lanes 0 and 63 load
// them), and prefetches the next row's operands before computing the current
// one.  Same arithmetic, same association order as k_stencil.
template <int CPL, bool PER, bool MMS, bool FUSED, bool NT>
__global__ __launch_bounds__(kBlock) void k_march(const KArgs a)
{
  const int lane = threadIdx.x & (kWave - 1);
  const unsigned B = gridDim.x, b = blockIdx.x;
  const unsigned q = B >> 3, r = B & 7u, xcd = b & 7u, loc = b >> 3;
  const unsigned tile = xcd * q + (xcd < r ? xcd : r) + loc;
  const int tx = (int)(tile % (unsigned)a.tiles_x), ty = (int)(tile / (unsigned)a.tiles_x);
  const int64_t nx = a.nx, ny = a.ny;
  const int64_t i = (int64_t)tx * (kBlock * CPL) + (int64_t)threadIdx.x * CPL;
  const bool act = i < nx;

  double T1 = a.T1, T2 = a.T2;
  double c[5] = {0, 0, 0, 0, 0};
  const double* __restrict__ u = a.u;
  double* __restrict__ fout = a.f;
  const double* yjm2 = a.yjm2;
  const double* yn = a.yn;
  const double* fn = a.fn;
  if (a.ent) {
    u = a.ent->yjm1;
    fout = a.ent->out;
    yjm2 = a.ent->yjm2;
    yn = a.ent->yn;
    fn = a.ent->fn;
    T1 = a.ent->T1;
    T2 = a.ent->T2;
    if (FUSED) {
#pragma unroll
      for (int k = 0; k < 5; k++) c[k] = a.ent->c[k];
    }
  } else if (a.coef) {
    T1 = a.coef[5];
    T2 = a.coef[6];
    if (FUSED) {
#pragma unroll
      for (int k = 0; k < 5; k++) c[k] = a.coef[k];
    }
  }
  const int64_t ic = act ? i : 0;  // clamp for the table reads of idle lanes
  const double Dw0 = a.fx[ic], De0 = a.fx[ic + 1];
  const double De1 = (CPL == 2) ? a.fx[ic + 2] : 0.0;
  const int64_t iw = (i > 0) ? i - 1 : (PER ? nx - 1 : -1);
  const int64_t ie = (i + CPL < nx) ? i + CPL : (PER ? 0 : -1);
  const bool need_w = (lane == 0);
  const bool need_e = (lane == kWave - 1) || (i + CPL >= nx);
  double sx0 = 0, sx1 = 0, cx0 = 0, cx1 = 0;
  if (MMS) {
    sx0 = a.sx[ic];
    cx0 = a.cx[ic];
    if (CPL == 2) {
      sx1 = a.sx[ic + 1];
      cx1 = a.cx[ic + 1];
    }
  }
  const int64_t j0 = (int64_t)ty * a.rows;
  const int64_t j1 = (j0 + a.rows < ny) ? j0 + a.rows : ny;

  // row loader: row index wrapped (periodic) or zero outside (Dirichlet)
  auto rowld = [&](int64_t rr, double& v0, double& v1) {
    if (rr < 0) rr = PER ? ny - 1 : -1;
    else if (rr >= ny) rr = PER ? 0 : -1;
    v0 = 0.0;
    v1 = 0.0;
    if (act && rr >= 0) {
      if (CPL == 2) {
        const double2 t = *reinterpret_cast<const double2*>(u + rr * nx + i);
        v0 = t.x;
        v1 = t.y;
      } else v0 = u[rr * nx + i];
    }
  };
  auto edges = [&](int64_t rr, double& w, double& e) {
    w = (act && need_w && iw >= 0) ? u[rr * nx + iw] : 0.0;
    e = (act && need_e && ie >= 0) ? u[rr * nx + ie] : 0.0;
  };
  auto streams = [&](int64_t rr, double* A, double* Y, double* G) {
    if (!FUSED || !act) return;
    const int64_t o = rr * nx + i;
    if (CPL == 2) {
      const double2 x = ldv2<NT>(yjm2 + o), y = ldv2<NT>(yn + o), g = ldv2<NT>(fn + o);
      A[0] = x.x; A[1] = x.y; Y[0] = y.x; Y[1] = y.y; G[0] = g.x; G[1] = g.y;
    } else {
      A[0] = lds1<NT>(yjm2 + o); Y[0] = lds1<NT>(yn + o); G[0] = lds1<NT>(fn + o);
    }
  };

  double S0, S1, C0, C1, N0, N1, we, ee;
  double A[2] = {0, 0}, Y[2] = {0, 0}, G[2] = {0, 0};
  rowld(j0 - 1, S0, S1);
  rowld(j0, C0, C1);
  rowld(j0 + 1, N0, N1);
  edges(j0, we, ee);
  streams(j0, A, Y, G);
  for (int64_t j = j0; j < j1; j++) {
    // prefetch the next output row's operands
    double M0 = 0, M1 = 0, we2 = 0, ee2 = 0;
    double A2[2] = {0, 0}, Y2[2] = {0, 0}, G2[2] = {0, 0};
    const bool more = (j + 1 < j1);
    if (more) {
      rowld(j + 2, M0, M1);
      edges(j + 1, we2, ee2);
      streams(j + 1, A2, Y2, G2);
    }
    double wv = __shfl_up(CPL == 2 ? C1 : C0, 1, kWave);
    double ev = __shfl_down(C0, 1, kWave);
    if (need_w) wv = we;
    if (need_e) ev = ee;
    if (act) {
      const double Ds = a.fy[j], Dn = a.fy[j + 1];
      const int64_t o = j * nx + i;
      if (CPL == 2) {
        double f0 = cell(Dw0, De0, Ds, Dn, C0, wv, C1, S0, N0);
        double f1 = cell(De0, De1, Ds, Dn, C1, C0, ev, S1, N1);
        if (MMS) {
          const double syj = a.sy[j], cyj = a.cy[j];
          f0 = f0 + forcing(T1, T2, sx0, cx0, syj, cyj);
          f1 = f1 + forcing(T1, T2, sx1, cx1, syj, cyj);
        }
        if (FUSED) {
          f0 = lc5(c, f0, A[0], Y[0], C0, G[0]);
          f1 = lc5(c, f1, A[1], Y[1], C1, G[1]);
        }
        stv2<NT>(fout + o, f0, f1);
      } else {
        double f0 = cell(Dw0, De0, Ds, Dn, C0, wv, ev, S0, N0);
        if (MMS) f0 = f0 + forcing(T1, T2, sx0, cx0, a.sy[j], a.cy[j]);
        if (FUSED) f0 = lc5(c, f0, A[0], Y[0], C0, G[0]);
        sts1<NT>(fout + o, f0);
      }
    }
    S0 = C0; S1 = C1;
    C0 = N0; C1 = N1;
    N0 = M0; N1 = M1;
    we = we2; ee = ee2;
    A[0] = A2[0]; A[1] = A2[1]; Y[0] = Y2[0]; Y[1] = Y2[1]; G[0] = G2[0]; G[1] = G2[1];
  }
}

template <int CPL, bool PER, bool MMS, bool FUSED>
void launch_t(const KArgs& a, const StencilCfg& cfg, unsigned grid, hipStream_t s)
{
  if (cfg.variant == 0) hipLaunchKernelGGL((k_stencil<CPL, PER, MMS, FUSED>), dim3(grid), dim3(kBlock), 0, s, a);
  else if (cfg.nt) hipLaunchKernelGGL((k_march<CPL, PER, MMS, FUSED, true>), dim3(grid), dim3(kBlock), 0, s, a);
  else hipLaunchKernelGGL((k_march<CPL, PER, MMS, FUSED, false>), dim3(grid), dim3(kBlock), 0, s, a);
}

bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

void launch_stencil(DiffusionProblem* P, KArgs a, bool fused, hipStream_t s)
{
  // table mode: the integrator's own hipMalloc'd vectors (256-byte aligned)
  bool al = (P->p.nx % 2 == 0) &&
            (a.ent || (aligned16(a.u) && aligned16(a.f) &&
                       (!fused || (aligned16(a.yjm2) && aligned16(a.yn) && aligned16(a.fn)))));
  const int cpl = al ? 2 : 1;
  const StencilCfg& cfg = P->cfg[fused ? 1 : 0];
  const int rows = cfg.rows > 0 ? cfg.rows : 1;
  a.fx = P->d_fx;
  a.fy = P->d_fy;
  a.sx = P->d_sx;
  a.sy = P->d_sy;
  a.cx = P->d_cx;
  a.cy = P->d_cy;
  a.nx = P->p.nx;
  a.ny = P->p.ny;
  a.rows = rows;
  int64_t tiles_y;
  if (cfg.variant == 0) {
    a.tiles_x = (int)((P->p.nx + kWave * cpl - 1) / (kWave * cpl));
    tiles_y = (P->p.ny + 4 * rows - 1) / (4 * rows);
  } else {
    a.tiles_x = (int)((P->p.nx + kBlock * cpl - 1) / (kBlock * cpl));
    tiles_y = (P->p.ny + rows - 1) / rows;
  }
  const unsigned grid = (unsigned)(a.tiles_x * tiles_y);
  const bool per = (P->p.kind == DIFFUSION_CEDA_PERIODIC);
#define CEDA_DISPATCH(C, PE, M)                                     \
  if (fused) launch_t<C, PE, M, true>(a, cfg, grid, s);             \
  else launch_t<C, PE, M, false>(a, cfg, grid, s);
  if (cpl == 2) {
    if (per) { CEDA_DISPATCH(2, true, false) }
    else { CEDA_DISPATCH(2, false, true) }
  } else {
    if (per) { CEDA_DISPATCH(1, true, false) }
    else { CEDA_DISPATCH(1, false, true) }
  }
#undef CEDA_DISPATCH
}

// first STS stage from the table: out = c0*yn + c1*fn (N_VLinearSum order)
__global__ __launch_bounds__(kBlock) void k_stage1(int64_t n, const StageEntry* e)
{
  const double a0 = e->c[0], a1 = e->c[1];
  const double* __restrict__ x = e->yn;
  const double* __restrict__ y = e->fn;
  double* __restrict__ z = e->out;
  const int64_t np = n >> 1;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < np; k += stride) {
    const double2 X = *reinterpret_cast<const double2*>(x + 2 * k);
    const double2 Y = *reinterpret_cast<const double2*>(y + 2 * k);
    *reinterpret_cast<double2*>(z + 2 * k) = make_double2(a0 * X.x + a1 * Y.x, a0 * X.y + a1 * Y.y);
  }
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) z[n - 1] = a0 * x[n - 1] + a1 * y[n - 1];
}


__global__ __launch_bounds__(kBlock) void k_psetup(int64_t nx, int64_t ny, const double* ax,
                                                   const double* ay, double gamma, double* diag)
{
  const int64_t n = nx * ny;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += stride) {
    const int64_t j = k / nx, i = k - j * nx;
    const double d = -((ax[i] + ax[i + 1]) + (ay[j] + ay[j + 1]));
    diag[k] = 1.0 / (1.0 - gamma * d);
  }
}

// u = A[i]*B[j] (periodic IC factors) or sx[i]*sy[j]*T2 (MMS exact, T2 = 1 at t=0)
__global__ __launch_bounds__(kBlock) void k_outer(int64_t nx, int64_t ny, const double* A,
                                                  const double* B, int scale, double T2,
                                                  double* u)
{
  const int64_t n = nx * ny;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += stride) {
    const int64_t j = k / nx, i = k - j * nx;
    double v = A[i] * B[j];
    if (scale) v = v * T2;
    u[k] = v;
  }
}

double coeff(double k, int inhomogeneous, double s)
{
  if (inhomogeneous) return k * (1.0 + 0.99 * sin(s));
  return k;
}

// the problem's 1-D tables (face coefficients, forcing and initial-condition
// factors) in one device allocation, each table starting on a 256-byte
// boundary: one allocation and one copy per problem, not one of each per
// table (problems created from several host threads contend for both)
void upload_tables(DiffusionProblem* P)
{
  const std::vector<double>* hs[] = {&P->fx, &P->fy, &P->px, &P->py, &P->sx,
                                     &P->sy, &P->cx, &P->cy, &P->ax, &P->by};
  double** ds[] = {&P->d_fx, &P->d_fy, &P->d_px, &P->d_py, &P->d_sx, &P->d_sy, &P->d_cx, &P->d_cy, &P->d_ax, &P->d_by};
  constexpr int nt = sizeof(hs) / sizeof(hs[0]);
  size_t off[nt], tot = 0;
  for (int i = 0; i < nt; i++) {
    off[i] = tot;
    tot += (std::max<size_t>(hs[i]->size(), 1) + 31) / 32 * 32;
  }
  std::vector<double> h(tot, 0.0);
  for (int i = 0; i < nt; i++) std::copy(hs[i]->begin(), hs[i]->end(), h.begin() + (ptrdiff_t)off[i]);
  CEDA_HIP_CHECK(ceda::dev_malloc(&P->d_tables, sizeof(double) * tot));
  CEDA_HIP_CHECK(ceda::copy_sync(P->d_tables, h.data(), sizeof(double) * tot, hipMemcpyHostToDevice));
  for (int i = 0; i < nt; i++) *ds[i] = P->d_tables + off[i];
}

// autotuner results, per (nx, ny, kind) for the life of the process
struct TuneKey {
  int64_t nx, ny;
  int kind;
  bool operator<(const TuneKey& o) const
  {
    return nx != o.nx ? nx < o.nx : (ny != o.ny ? ny < o.ny : kind < o.kind);
  }
};
std::map<TuneKey, std::array<StencilCfg, 3>> g_tune;
std::mutex g_tune_mu;

bool vec_ok(const DiffusionProblem* P, N_Vector v)
{
  return v && v->content && N_VGetLength(v) == P->n && N_VGetDeviceArrayPointer(v);
}

}  // namespace

namespace ceda {

// T1 = -2 pi sin(pi t) cos(pi t) = -pi sin(2 pi t), T2 = cos^2(pi t)
void mms_time_factors(double t, double* T1, double* T2)
{
  const double c = cos(kPi * t);
  *T1 = -kPi * sin(2.0 * kPi * t);
  *T2 = c * c;
}

void launch_fused_stage(DiffusionProblem* P, const double* coef, const double* yjm1,
                        const double* yjm2, const double* yn, const double* fn, double* out,
                        hipStream_t s)
{
  KArgs a{};
  a.u = yjm1;
  a.f = out;
  a.coef = coef;
  a.yjm2 = yjm2;
  a.yn = yn;
  a.fn = fn;
  launch_stencil(P, a, true, s);
}

void launch_stage1_entry(DiffusionProblem* P, const StageEntry* e, hipStream_t s)
{
  hipLaunchKernelGGL(k_stage1, dim3(stream_grid(P->n)), dim3(kBlock), 0, s, P->n, e);
}

void launch_fused_entry(DiffusionProblem* P, const StageEntry* e, hipStream_t s)
{
  KArgs a{};
  a.ent = e;
  launch_stencil(P, a, true, s);
}

void launch_rhs_entry(DiffusionProblem* P, const StageEntry* e, hipStream_t s)
{
  KArgs a{};
  a.ent = e;
  launch_stencil(P, a, false, s);
}

void launch_rhs_devtime(DiffusionProblem* P, const double* coef, const double* u, double* f,
                        hipStream_t s)
{
  KArgs a{};
  a.u = u;
  a.f = f;
  a.coef = coef;
  launch_stencil(P, a, false, s);
}

}  // namespace ceda

extern "C" {

void DiffusionParams_Default(DiffusionParams* p, int kind)
{
  memset(p, 0, sizeof(*p));
  p->kind = kind;
  p->kx = 1.0;
  p->ky = 1.0;
  p->tf = 1.0;
  if (kind == DIFFUSION_CEDA_PERIODIC) {
    p->xl = -kPi; p->xu = kPi; p->yl = -6.0; p->yu = 6.0;
    p->nx = 64; p->ny = 64;
  } else {
    p->xl = 0.0; p->xu = 1.0; p->yl = 0.0; p->yu = 1.0;
    p->nx = 128; p->ny = 128;
  }
}

DiffusionProblem* DiffusionProblem_Create(const DiffusionParams* prm, SUNContext ctx)
{
  if (!prm || !ctx) return nullptr;
  if (prm->nx < 2 || prm->ny < 2) return nullptr;
  if (prm->kind != DIFFUSION_CEDA_PERIODIC && prm->kind != DIFFUSION_MMS_DIRICHLET) return nullptr;
  if (prm->kind == DIFFUSION_MMS_DIRICHLET && prm->inhomogeneous) return nullptr;
  DiffusionProblem* P = new DiffusionProblem();
  P->p = *prm;
  P->ctx = ctx;
  P->exec = context_exec(ctx);
  const int64_t nx = prm->nx, ny = prm->ny;
  P->n = nx * ny;
  const bool per = (prm->kind == DIFFUSION_CEDA_PERIODIC);
  // diffusion_2D.cpp:159-160 (periodic: nodes at both ends are distinct);
  // interior nodes of [xl,xu] for the Dirichlet problem
  P->dx = per ? (prm->xu - prm->xl) / (double)(nx - 1) : (prm->xu - prm->xl) / (double)(nx + 1);
  P->dy = per ? (prm->yu - prm->yl) / (double)(ny - 1) : (prm->yu - prm->yl) / (double)(ny + 1);
  const double dx = P->dx, dy = P->dy, dx2 = dx * dx, dy2 = dy * dy;
  const double xo = per ? 0.0 : 1.0;
  P->fx.resize(nx + 1); P->px.resize(nx + 1);
  P->fy.resize(ny + 1); P->py.resize(ny + 1);
  // face k at xl+(k-1/2)dx: west face of node k, east face of node k-1
  // (diffusion.cpp:36-46); Jacobi faces at k*dx (preconditioner_jacobi.cpp:23-37)
  for (int64_t i = 0; i <= nx; i++) {
    const double xf = prm->xl + ((double)i + xo - 0.5) * dx;
    P->fx[i] = coeff(prm->kx, prm->inhomogeneous, xf) / dx2;
    P->px[i] = coeff(prm->kx, prm->inhomogeneous, (double)i * dx) / dx2;
  }
  for (int64_t j = 0; j <= ny; j++) {
    const double yf = prm->yl + ((double)j + xo - 0.5) * dy;
    P->fy[j] = coeff(prm->ky, prm->inhomogeneous, yf) / dy2;
    P->py[j] = coeff(prm->ky, prm->inhomogeneous, (double)j * dy) / dy2;
  }
  P->sx.assign(nx, 0.0); P->cx.assign(nx, 0.0); P->ax.assign(nx, 0.0);
  P->sy.assign(ny, 0.0); P->cy.assign(ny, 0.0); P->by.assign(ny, 0.0);
  if (per) {
    // initial.cpp:42-43 factorised: ((1+0.3 sin 2x)/sqrt(5.5 pi)) * exp(-y^2/5.5)
    for (int64_t i = 0; i < nx; i++) {
      const double x = prm->xl + (double)i * dx;
      P->ax[i] = (1.0 + 0.3 * sin(2.0 * x)) / sqrt(5.5 * kPi);
    }
    for (int64_t j = 0; j < ny; j++) {
      const double y = prm->yl + (double)j * dy;
      P->by[j] = exp(-(y * y) / 5.5);
    }
  } else {
    // README.md:28 forcing factors; cos^2 - sin^2 is evaluated as cos(2 pi x)
    // so no sin/cos pair of one argument exists for a compiler to fuse into
    // sincos(), which rounds differently in rare cases
    const double tpp = 2.0 * kPi * kPi;
    for (int64_t i = 0; i < nx; i++) {
      const double x = prm->xl + (double)(i + 1) * dx;
      const double s = sin(kPi * x);
      P->sx[i] = s * s;
      P->cx[i] = prm->kx * (-tpp * cos(2.0 * kPi * x));
    }
    for (int64_t j = 0; j < ny; j++) {
      const double y = prm->yl + (double)(j + 1) * dy;
      const double s = sin(kPi * y);
      P->sy[j] = s * s;
      P->cy[j] = prm->ky * (-tpp * cos(2.0 * kPi * y));
    }
  }
  upload_tables(P);
  CEDA_HIP_CHECK(ceda::dev_malloc(&P->d_tf, sizeof(double) * 8));
  CEDA_HIP_CHECK(ceda::host_malloc(&P->h_tf, sizeof(double) * 8, hipHostMallocDefault));
  P->diag = N_VNew_HipNative(P->n, ctx);
  N_VConst(1.0, P->diag);
  P->cfg[0] = StencilCfg{0, 8, 0};
  P->cfg[1] = StencilCfg{0, 2, 0};
  P->cfg[2] = StencilCfg{0, 32, 0};
  return P;
}

void DiffusionProblem_Destroy(DiffusionProblem* P)
{
  if (!P) return;
  CEDA_HIP_CHECK(hipStreamSynchronize(P->exec->stream));
  N_VDestroy(P->diag);
  ceda::mg_destroy(P->mg);
  double* ds[] = {P->d_tables, P->d_tf};
  for (double* d : ds)
    if (d) CEDA_HIP_CHECK(ceda::dev_free(d));
  if (P->h_tf) CEDA_HIP_CHECK(ceda::host_free(P->h_tf));
  delete P;
}

void DiffusionProblem_GetGrid(const DiffusionProblem* P, int64_t* nx, int64_t* ny, double* dx,
                              double* dy)
{
  if (nx) *nx = P->p.nx;
  if (ny) *ny = P->p.ny;
  if (dx) *dx = P->dx;
  if (dy) *dy = P->dy;
}

long DiffusionProblem_GetNumRhsEvals(const DiffusionProblem* P) { return P->nrhs; }

int DiffusionProblem_SetStencilRows(DiffusionProblem* P, int rows)
{
  if (rows != 0 && rows != 1 && rows != 2 && rows != 4 && rows != 8) return -1;
  for (int k = 0; k < 2; k++) {
    P->cfg[k].variant = 0;
    P->cfg[k].rows = rows ? rows : 1;
    P->cfg[k].nt = 0;
  }
  return 0;
}

int DiffusionProblem_SetStencilConfig(DiffusionProblem* P, int which, int variant, int rows, int nt)
{
  if (!P || which < 0 || which > 2 || rows < 1 || rows > 1024 || nt < 0 || nt > (which == 2 ? 3 : 1)) return -1;
  if (which < 2 && (variant < 0 || variant > 1 || rows > 256 || (variant == 0 && nt))) return -1;
  if (which == 2 && (variant < 0 || variant > kMaxPassStages)) return -1;
  P->cfg[which].variant = variant;
  P->cfg[which].rows = rows;
  P->cfg[which].nt = nt;
  return 0;
}

int DiffusionProblem_GetStencilConfig(const DiffusionProblem* P, int which, int* variant, int* rows,
                                      int* nt)
{
  if (!P || which < 0 || which > 2) return -1;
  if (variant) *variant = P->cfg[which].variant;
  if (rows) *rows = P->cfg[which].rows;
  if (nt) *nt = P->cfg[which].nt;
  return 0;
}

}  // extern "C"

namespace {
// device stage table of a K-stage timing pass: the timing coefficients,
// inputs ops[0..3] = Y_{j-1}, Y_{j-2}, Y_n, F_n; stage K-1 -> ops[4],
// stage K -> ops[5]
StageEntry* pass_table(double* const* ops, int K, hipStream_t s)
{
  // the kernel reads entries 0..K-1: the table holds the deepest pass
  if (K < 1 || K > kMaxPassStages) return nullptr;
  StageEntry h[kMaxPassStages];
  memset(h, 0, sizeof(h));
  const double hc[5] = {1.0e-9, -0.5, 0.25, 1.25, -1.0e-10};
  for (int m = 0; m < kMaxPassStages; m++) {
    for (int k = 0; k < 5; k++) h[m].c[k] = hc[k];
    h[m].T2 = 1.0;
    h[m].yjm1 = ops[0];
    h[m].yjm2 = ops[1];
    h[m].yn = ops[2];
    h[m].fn = ops[3];
    h[m].out = ops[5];
  }
  if (K >= 2) h[K - 2].out = ops[4];
  StageEntry* d = nullptr;
  CEDA_HIP_CHECK(ceda::dev_malloc(&d, sizeof(h)));
  CEDA_HIP_CHECK(hipMemcpyAsync(d, h, sizeof(h), hipMemcpyHostToDevice, s));
  CEDA_HIP_CHECK(hipStreamSynchronize(s));
  return d;
}
}  // namespace

extern "C" {

int DiffusionProblem_Autotune(DiffusionProblem* P, int reps)
{
  if (!P || reps < 1) return -1;
  const TuneKey key{P->p.nx, P->p.ny, P->p.kind};
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      for (int k = 0; k < 3; k++) P->cfg[k] = it->second[k];
      P->tuned = true;
      return 0;
    }
  }
  hipStream_t s = P->exec->stream;
  double* buf[6];
  for (int k = 0; k < 6; k++) {
    CEDA_HIP_CHECK(ceda::dev_malloc(&buf[k], sizeof(double) * P->n));
    CEDA_HIP_CHECK(hipMemsetAsync(buf[k], 0, sizeof(double) * P->n, s));
  }
  double* coef = nullptr;
  const double hc[8] = {1.0e-9, -0.5, 0.25, 1.25, -1.0e-10, 0.0, 1.0, 0.0};
  CEDA_HIP_CHECK(ceda::dev_malloc(&coef, sizeof(hc)));
  CEDA_HIP_CHECK(hipMemcpyAsync(coef, hc, sizeof(hc), hipMemcpyHostToDevice, s));
  hipEvent_t e0, e1;
  CEDA_HIP_CHECK(hipEventCreate(&e0));
  CEDA_HIP_CHECK(hipEventCreate(&e1));
  std::vector<StencilCfg> cands;
  for (int rows : {1, 2, 4, 8}) cands.push_back(StencilCfg{0, rows, 0});
  for (int rows : {4, 8, 16, 32, 64})
    for (int nt : {0, 1}) cands.push_back(StencilCfg{1, rows, nt});
  float best_stage = 1e30f;
  for (int which = 0; which < 2; which++) {
    float best = 1e30f;
    StencilCfg bestc = P->cfg[which];
    for (const StencilCfg& c : cands) {
      if (c.variant == 1 && c.rows > P->p.ny) continue;
      P->cfg[which] = c;
      auto launch = [&]() {
        if (which == 0) launch_rhs_devtime(P, coef, buf[0], buf[4], s);
        else launch_fused_stage(P, coef, buf[0], buf[1], buf[2], buf[3], buf[4], s);
      };
      launch();
      CEDA_HIP_CHECK(hipEventRecord(e0, s));
      for (int r = 0; r < reps; r++) launch();
      CEDA_HIP_CHECK(hipEventRecord(e1, s));
      CEDA_HIP_CHECK(hipEventSynchronize(e1));
      float ms = 0.0f;
      CEDA_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) {
        best = ms;
        bestc = c;
      }
    }
    P->cfg[which] = bestc;
    if (which == 1) best_stage = best;
  }
  // multi-stage passes: kept only when a stage costs less than the best
  // single-stage kernel (per stage: pass time / K)
  StencilCfg bestp{0, 32, 0};
  float bestp_ms = best_stage * 0.97f;
  // heights between 16 and 64 matter: the K-1 halo rows above and below a
  // tile are re-read (mostly from HBM), while tall tiles leave too few
  // wavefronts or a partial last round of them
  for (int K : {2, 3, 4, 5, 6}) {
    StageEntry* dtab = pass_table(buf, K, s);
    for (int rows : {8, 16, 24, 32, 40, 48, 64, 96, 128})
      for (int nt : {0, 1, 2, 3}) {
        if (rows > 2 * P->p.ny) continue;
        launch_pass(P, K, rows, nt, dtab, s);
        CEDA_HIP_CHECK(hipEventRecord(e0, s));
        for (int r = 0; r < reps; r++) launch_pass(P, K, rows, nt, dtab, s);
        CEDA_HIP_CHECK(hipEventRecord(e1, s));
        CEDA_HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.0f;
        CEDA_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (ms / K < bestp_ms) {
          bestp_ms = ms / K;
          bestp = StencilCfg{K, rows, nt};
        }
      }
    CEDA_HIP_CHECK(hipStreamSynchronize(s));
    CEDA_HIP_CHECK(ceda::dev_free(dtab));
  }
  P->cfg[2] = bestp;
  CEDA_HIP_CHECK(hipEventDestroy(e0));
  CEDA_HIP_CHECK(hipEventDestroy(e1));
  CEDA_HIP_CHECK(hipStreamSynchronize(s));
  for (int k = 0; k < 6; k++) CEDA_HIP_CHECK(ceda::dev_free(buf[k]));
  CEDA_HIP_CHECK(ceda::dev_free(coef));
  P->tuned = true;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune[key] = {P->cfg[0], P->cfg[1], P->cfg[2]};
  return 0;
}

int DiffusionProblem_TimeKernel(DiffusionProblem* P, int which, N_Vector u, N_Vector a,
                                N_Vector b, N_Vector c, N_Vector out, int reps, double* avg_ms)
{
  if (!P || reps < 1 || !avg_ms || !vec_ok(P, u) || !vec_ok(P, out)) return -1;
  if (which >= 1 && (!vec_ok(P, a) || !vec_ok(P, b) || !vec_ok(P, c))) return -1;
  if (which < 0 || which > 2 || (which == 2 && P->cfg[2].variant < 1)) return -1;
  hipStream_t s = P->exec->stream;
  double* coef = nullptr;
  // stage coefficients of a mid-loop RKL2 stage; the values do not change
  // the traffic, only keep the result finite
  const double hc[8] = {1.0e-9, -0.5, 0.25, 1.25, -1.0e-10, 0.0, 1.0, 0.0};
  CEDA_HIP_CHECK(ceda::dev_malloc(&coef, sizeof(hc)));
  CEDA_HIP_CHECK(ceda::copy_sync(coef, hc, sizeof(hc), hipMemcpyHostToDevice));
  hipEvent_t e0, e1;
  CEDA_HIP_CHECK(hipEventCreate(&e0));
  CEDA_HIP_CHECK(hipEventCreate(&e1));
  // a pass writes two stages: the next-to-last goes to a scratch vector
  double* mid = nullptr;
  StageEntry* dtab = nullptr;
  if (which == 2) {
    CEDA_HIP_CHECK(ceda::dev_malloc(&mid, sizeof(double) * P->n));
    CEDA_HIP_CHECK(ceda::set_sync(mid, 0, sizeof(double) * P->n));
    double* ops[6] = {dptr(u), dptr(a), dptr(b), dptr(c), mid, dptr(out)};
    dtab = pass_table(ops, P->cfg[2].variant, s);
  }
  auto launch = [&]() {
    if (which == 0) launch_rhs_devtime(P, coef, dptr(u), dptr(out), s);
    else if (which == 1) launch_fused_stage(P, coef, dptr(u), dptr(a), dptr(b), dptr(c), dptr(out), s);
    else launch_pass(P, P->cfg[2].variant, P->cfg[2].rows, P->cfg[2].nt, dtab, s);
  };
  launch();  // warm
  CEDA_HIP_CHECK(hipEventRecord(e0, s));
  for (int r = 0; r < reps; r++) launch();
  CEDA_HIP_CHECK(hipEventRecord(e1, s));
  CEDA_HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0.0f;
  CEDA_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  *avg_ms = (double)ms / reps;
  if (mid) {
    CEDA_HIP_CHECK(hipStreamSynchronize(s));
    CEDA_HIP_CHECK(ceda::dev_free(mid));
    CEDA_HIP_CHECK(ceda::dev_free(dtab));
  }
  CEDA_HIP_CHECK(hipEventDestroy(e0));
  CEDA_HIP_CHECK(hipEventDestroy(e1));
  CEDA_HIP_CHECK(ceda::dev_free(coef));
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int DiffusionProblem_StagePass(DiffusionProblem* P, int K, const double* coef, N_Vector yjm1,
                               N_Vector yjm2, N_Vector yn, N_Vector fn, N_Vector outkm1,
                               N_Vector outk)
{
  if (!P || K < 1 || K > kMaxPassStages || !coef || !vec_ok(P, yjm1) || !vec_ok(P, yjm2) || !vec_ok(P, yn) ||
      !vec_ok(P, fn) || !vec_ok(P, outk) || (K >= 2 && !vec_ok(P, outkm1)))
    return -1;
  const double* ins[4] = {dptr(yjm1), dptr(yjm2), dptr(yn), dptr(fn)};
  double* outs[2] = {K >= 2 ? dptr(outkm1) : nullptr, dptr(outk)};
  if (outs[0] == outs[1]) return -1;
  for (const double* x : ins)
    for (double* o : outs)
      if (o && o == x) return -1;
  hipStream_t s = P->exec->stream;
  StageEntry h[kMaxPassStages];  // K <= kMaxPassStages (checked above)
  memset(h, 0, sizeof(h));
  for (int m = 0; m < K; m++) {
    for (int k = 0; k < 5; k++) h[m].c[k] = coef[8 * m + k];
    h[m].T1 = coef[8 * m + 5];
    h[m].T2 = coef[8 * m + 6];
    h[m].yjm1 = ins[0];
    h[m].yjm2 = ins[1];
    h[m].yn = ins[2];
    h[m].fn = ins[3];
    h[m].out = outs[1];
  }
  if (K >= 2) h[K - 2].out = outs[0];
  StageEntry* d = nullptr;
  CEDA_HIP_CHECK(ceda::dev_malloc(&d, sizeof(h)));
  CEDA_HIP_CHECK(hipMemcpyAsync(d, h, sizeof(h), hipMemcpyHostToDevice, s));
  launch_pass(P, K, P->cfg[2].rows, P->cfg[2].nt, d, s);
  CEDA_HIP_CHECK(hipGetLastError());
  CEDA_HIP_CHECK(hipStreamSynchronize(s));
  CEDA_HIP_CHECK(ceda::dev_free(d));
  return 0;
}

int Initial(sunrealtype, N_Vector u, void* user_data)
{
  DiffusionProblem* P = static_cast<DiffusionProblem*>(user_data);
  if (!P || !vec_ok(P, u)) return -1;
  const bool per = (P->p.kind == DIFFUSION_CEDA_PERIODIC);
  hipLaunchKernelGGL(k_outer, dim3(stream_grid(2 * P->n)), dim3(kBlock), 0, vstream(u), P->p.nx,
                     P->p.ny, per ? P->d_ax : P->d_sx, per ? P->d_by : P->d_sy, 0, 1.0, dptr(u));
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int Exact(sunrealtype t, N_Vector u, void* user_data)
{
  DiffusionProblem* P = static_cast<DiffusionProblem*>(user_data);
  if (!P || !vec_ok(P, u) || P->p.kind != DIFFUSION_MMS_DIRICHLET) return -1;
  double T1, T2;
  mms_time_factors(t, &T1, &T2);
  hipLaunchKernelGGL(k_outer, dim3(stream_grid(2 * P->n)), dim3(kBlock), 0, vstream(u), P->p.nx,
                     P->p.ny, P->d_sx, P->d_sy, 1, T2, dptr(u));
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// diffusion.cpp:9-209 on one GPU (no forcing term)
int laplacian(sunrealtype, N_Vector u, N_Vector f, void* user_data)
{
  DiffusionProblem* P = static_cast<DiffusionProblem*>(user_data);
  if (!P || !vec_ok(P, u) || !vec_ok(P, f)) return -1;
  KArgs a{};
  a.u = dptr(u);
  a.f = dptr(f);
  a.T1 = 0.0;
  a.T2 = 0.0;
  // the Dirichlet kernel adds the forcing; with T1 = T2 = 0 it adds +0
  launch_stencil(P, a, false, vstream(f));
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int JTimes(N_Vector v, N_Vector Jv, sunrealtype t, N_Vector, N_Vector, void* user_data, N_Vector)
{
  return laplacian(t, v, Jv, user_data);
}

// diffusion_2D.cpp:23-35 (+ README.md:28 forcing for the manufactured problem)
int diffusion(sunrealtype t, N_Vector u, N_Vector f, void* user_data)
{
  DiffusionProblem* P = static_cast<DiffusionProblem*>(user_data);
  if (!P || !vec_ok(P, u) || !vec_ok(P, f)) return -1;
  KArgs a{};
  a.u = dptr(u);
  a.f = dptr(f);
  if (P->p.kind == DIFFUSION_MMS_DIRICHLET) mms_time_factors(t, &a.T1, &a.T2);
  launch_stencil(P, a, false, vstream(f));
  P->nrhs++;
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// preconditioner_jacobi.cpp:9-44
int PSetup(sunrealtype, N_Vector, N_Vector, sunbooleantype, sunbooleantype* jcurPtr,
           sunrealtype gamma, void* user_data)
{
  DiffusionProblem* P = static_cast<DiffusionProblem*>(user_data);
  if (!P) return -1;
  const bool per = (P->p.kind == DIFFUSION_CEDA_PERIODIC);
  hipLaunchKernelGGL(k_psetup, dim3(stream_grid(2 * P->n)), dim3(kBlock), 0, vstream(P->diag),
                     P->p.nx, P->p.ny, per ? P->d_px : P->d_fx, per ? P->d_py : P->d_fy, gamma,
                     dptr(P->diag));
  P->diag_gamma = gamma;
  P->diag_gamma_set = true;
  if (jcurPtr) *jcurPtr = SUNTRUE;
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// preconditioner_jacobi.cpp:47-62: z = diag .* r
int PSolve(sunrealtype, N_Vector, N_Vector, N_Vector r, N_Vector z, sunrealtype, sunrealtype,
           int, void* user_data)
{
  DiffusionProblem* P = static_cast<DiffusionProblem*>(user_data);
  if (!P || !vec_ok(P, r) || !vec_ok(P, z)) return -1;
  N_VProd(P->diag, r, z);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// main.cpp:536-550
int dom_eig(sunrealtype, N_Vector, N_Vector, sunrealtype* lambdaR, sunrealtype* lambdaI,
            void* user_data, N_Vector, N_Vector, N_Vector)
{
  DiffusionProblem* P = static_cast<DiffusionProblem*>(user_data);
  if (!P) return -1;
  const double a = P->p.kx / P->dx / P->dx, b = P->p.ky / P->dy / P->dy;
  *lambdaR = -8.0 * (a > b ? a : b);
  *lambdaI = 0.0;
  return 0;
}

}  // extern "C"
