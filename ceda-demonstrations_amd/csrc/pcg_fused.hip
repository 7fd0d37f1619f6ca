// Fused preconditioned CG for the implicit path: (I - gamma J) x = b with J
// the 5-point stencil of the RHS (the analytic Jacobian-vector product,
// exactly laplacian()'s arithmetic, diffusion.cpp:48-53) and the Jacobi or
// multigrid preconditioner.  The algorithm is the integrator's PCG
// (SUNLinSol_PCG, main.cpp:305; Hestenes-Stiefel with z = M r).
//
// With Jacobi (or no preconditioner) an iteration l is two kernels:
//   k_pcg_dirap:  the residual test of iteration l-1 and
//                 beta = r.z(l-1) / r.z(l-2); the direction
//                 p_l = z + beta p_{l-1}, formed at the five stencil points
//                 and stored into the other of two direction buffers;
//                 Ap = p_l - gamma * J p_l and the partial sums of p_l.Ap
//   k_pcg_update: alpha = r.z / p.Ap; x += alpha p, r -= alpha Ap,
//                 z = Dinv r (Dinv formed from the face coefficients, not
//                 read), and the partial sums of |r|^2 and r.z
// streaming 32 + 56 = 88 bytes per cell (176 for the separate vector
// operations).  The scalars never go to the host: every workgroup first
// sums the previous kernel's per-workgroup partials itself, in one fixed
// order (every workgroup gets bitwise the same alpha or beta, the value a
// separate one-workgroup combine would give), and the residual test is
// recorded in a device status word (running, converged at iteration l,
// non-finite) that every later kernel reads first, returning at once when
// the solve has ended.  The host enqueues iterations in batches and reads the
// status once per batch (after a one-workgroup kernel has tested the batch's
// last iteration): the iteration count, x and the statistics are those of a
// test after every iteration, and a NaN or infinity stops the solve with an
// error instead of running to the cap.  The arithmetic of every value is the
// one of the separate vector operations; only the reductions' summation
// order differs from them (deterministic, not bitwise equal), which is why
// this path is an option of the integrator (ARKodeSetDiffusionFastPCG) and
// the N_Vector PCG stays the reference.
//
// With a preconditioner given as a callback (multigrid, the learned
// network), the update leaves z alone, the callback's z = M r is enqueued on
// the same stream after it, and k_pcg_dot sums r.z for the next iteration's
// test and beta: the scalars stay on the device there too.
#include <algorithm>
#include <cmath>

#include "arkode_internal.hpp"
#include "diffusion_internal.hpp"
#include "stencil_ops.hpp"

using namespace ceda;

namespace ceda_ark {

namespace {

constexpr int kPcgBlocks = 1024;   // workgroups of the reduction kernels
constexpr int kPcgBatchMax = 32;   // iterations enqueued between host checks
constexpr int kPcgChunk = 4;       // callback preconditioner: iterations between host waits in the bulk
// workspace layout (doubles): partials of p.Ap, of (|r|^2, r.z), of r.z
// (callback preconditioner), then the solve state
constexpr int kOffA = 0, kOffB = kPcgBlocks, kOffC = 3 * kPcgBlocks, kOffState = 4 * kPcgBlocks;
constexpr int kWsDoubles = kOffState + 8;

// solve state in device memory (doubles): r.z of the current iteration in
// slot l & 1, the status word, the iteration it was set at, the last |r|^2
enum { kRz0 = 0, kRz1 = 1, kStatus = 2, kIter = 3, kRr = 4, kDelta = 5 };
enum { kRunning = 0, kConverged = 1, kNonFinite = 2 };

// workgroup sum of NACC accumulators into partials[blockIdx.x * NACC + a]
template <int NACC>
__device__ inline void block_sums(double (&acc)[NACC], double* __restrict__ partials)
{
  __shared__ double s[kBlock / kWave][NACC];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
  for (int a = 0; a < NACC; a++) {
    double v = acc[a];
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    if (lane == 0) s[wave][a] = v;
  }
  __syncthreads();
  if (threadIdx.x < NACC) {
    double v = s[0][threadIdx.x];
#pragma unroll
    for (int w = 1; w < kBlock / kWave; w++) v += s[w][threadIdx.x];
    partials[(int64_t)blockIdx.x * NACC + threadIdx.x] = v;
  }
}

// the sum of accumulator a over nblocks per-workgroup partials, returned to
// every thread of the workgroup: thread t adds partials t, t + 256, ... in
// order, then the wavefront butterfly, then the four wavefront sums in order
// (every workgroup gets bitwise the same value)
__device__ inline double sum_partials(const double* __restrict__ partials, int nblocks, int nacc, int a)
{
  __shared__ double s[kBlock / kWave];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  double v = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += kBlock) v += partials[(int64_t)b * nacc + a];
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  if (lane == 0) s[wave] = v;
  __syncthreads();
  double t = s[0];
#pragma unroll
  for (int w = 1; w < kBlock / kWave; w++) t += s[w];
  __syncthreads();  // s is reused by the next call
  return t;
}

// a grid-stride walk over the cells that keeps (i, j) without a division
// per cell: one division per thread for the start and the stride
// XCD-aware (as xcd_pairs): when the grid is a multiple of 8, workgroup b
// walks the (b % 8)-th of 8 contiguous ranges of cells, so that a cell's
// rows above and below are read by the same XCD at about the same time
struct CellWalk {
  int64_t k, i, j, n, nx, step, di, dj;
  __device__ CellWalk(int64_t n_, int64_t nx_) : n(n_), nx(nx_)
  {
    const unsigned G = gridDim.x, b = blockIdx.x;
    const int parts = (G % 8 == 0) ? 8 : 1;
    const int64_t per = (n_ + parts - 1) / parts;
    const int part = (int)(b % parts);
    k = part * per + (int64_t)(b / parts) * kBlock + threadIdx.x;
    n = min(n_, (part + 1) * per);
    j = k / nx;
    i = k - j * nx;
    step = (int64_t)(G / parts) * kBlock;
    dj = step / nx;
    di = step - dj * nx;
  }
  __device__ bool valid() const { return k < n; }
  __device__ void next()
  {
    k += step;
    i += di;
    j += dj;
    if (i >= nx) {
      i -= nx;
      j++;
    }
  }
};

__device__ inline bool solve_ended(const double* __restrict__ st) { return st[kStatus] != kRunning; }

// alpha = rz / p.Ap with p.Ap from the stencil kernel's partials (parts_a), then
// x += alpha p, r -= alpha Ap and
// mode 0: no preconditioner (z = r), 1: Jacobi (z = dinv r), 2: z computed
// elsewhere (only x and r are updated; acc[1] = 0).  wl2: the residual norm
// is the weighted one, sum (r w)^2, else sum r^2.  rz comes from the state
// (slot l & 1) when st is given, else from the argument.
// With JAC the update kernel forms the Jacobi inverse diagonal itself
// instead of reading it: dinv = 1 / (1 - gp d), d = -((ax[i] + ax[i+1]) +
// (ay[j] + ay[j+1])), exactly k_psetup's value for the gamma gp of the last
// preconditioner setup: 88 bytes per cell and iteration instead of 96.  (Not
// storing z at all and forming it at the stencil's five points instead
// saves 16 more bytes, but the eight divisions per pair of cells make the
// stencil kernel slower than the traffic saved.)
struct JacobiDiag {
  const double *ax, *ay;
  double gp;
};
__device__ inline double jac_dinv(const JacobiDiag& J, double sxi, double syj)
{
  const double d = -(sxi + syj);
  return 1.0 / (1.0 - J.gp * d);
}

// (the cell-wise update; JAC: mode 1 with the inverse diagonal formed from
// the face coefficients as above)
template <bool JAC>
__global__ __launch_bounds__(kBlock) void k_pcg_update(int64_t n, int64_t nx, int nparts,
                                                       const double* __restrict__ parts_a, double rz_arg,
                                                       const double* __restrict__ st, int l,
                                                       const double* __restrict__ p,
                                                       const double* __restrict__ ap, double* __restrict__ x,
                                                       double* __restrict__ r, double* __restrict__ z,
                                                       const double* __restrict__ dinv, JacobiDiag J,
                                                       const double* __restrict__ w, int mode,
                                                       double* __restrict__ partials)
{
  if (st && solve_ended(st)) return;
  const double pap = sum_partials(parts_a, nparts, 1, 0);
  const double rz = st ? st[l & 1] : rz_arg;
  const double alpha = rz / pap;
  double acc[2] = {0.0, 0.0};
  for (CellWalk c(n, nx); c.valid(); c.next()) {
    const int64_t k = c.k;
    x[k] = x[k] + alpha * p[k];
    const double rk = r[k] + (-alpha) * ap[k];
    r[k] = rk;
    const double rw = w ? rk * w[k] : rk;
    acc[0] += rw * rw;
    if (mode != 2) {
      double zk = rk;
      if (JAC) zk = jac_dinv(J, J.ax[c.i] + J.ax[c.i + 1], J.ay[c.j] + J.ay[c.j + 1]) * rk;
      else if (mode == 1) zk = dinv[k] * rk;
      z[k] = zk;
      acc[1] += rk * zk;
    }
  }
  block_sums<2>(acc, partials);
}

// the residual test of iteration l-1 from k_pcg_update's partials (|r|^2,
// r.z): converged, non-finite (|r| or beta), else running with
// beta = r.z(l-1) / r.z(l-2).  Workgroup 0 records the outcome; the caller
// returns when the solve has ended.
// r.z comes from parts_b's second accumulator, or, when parts_z is given
// (a preconditioner callback computed z after the update), from k_pcg_dot's
// partials parts_z.
__device__ inline bool test_iteration(int nparts, const double* __restrict__ parts_b,
                                      const double* __restrict__ parts_z, double* __restrict__ st, int l,
                                      double* beta)
{
  const double rr = sum_partials(parts_b, nparts, 2, 0);
  const double rznew = parts_z ? sum_partials(parts_z, nparts, 1, 0) : sum_partials(parts_b, nparts, 2, 1);
  const double rnorm = sqrt(rr);
  *beta = rznew / st[(l - 1) & 1];
  int status = kRunning;
  if (!isfinite(rnorm)) status = kNonFinite;
  else if (rnorm <= st[kDelta]) status = kConverged;
  else if (!isfinite(*beta)) status = kNonFinite;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st[kRr] = rr;
    st[kIter] = l - 1;
    if (status == kRunning) st[l & 1] = rznew;  // the r.z iteration l uses
    else st[kStatus] = status;
  }
  return status != kRunning;
}

__global__ __launch_bounds__(kBlock) void k_pcg_dirap(int64_t nx, int64_t ny, const double* __restrict__ fx,
                                                      const double* __restrict__ fy, int per, double gamma,
                                                      const double* __restrict__ z,
                                                      const double* __restrict__ pold, double* __restrict__ pnew,
                                                      double* __restrict__ ap, int nparts,
                                                      const double* __restrict__ parts_b,
                                                      const double* __restrict__ parts_z,
                                                      double* __restrict__ parts_a, double* __restrict__ st, int l)
{
  if (solve_ended(st)) return;
  double beta = 0.0;
  if (l > 1 && test_iteration(nparts, parts_b, parts_z, st, l, &beta)) return;
  // p_l at a cell: the first direction is z itself, already in pold
  auto pv = [&](int64_t k) { return (l > 1) ? z[k] + beta * pold[k] : pold[k]; };
  double acc[1] = {0.0};
  for (CellWalk c(nx * ny, nx); c.valid(); c.next()) {
    const int64_t k = c.k, i = c.i, j = c.j;
    double vw, ve, vs, vn;
    if (per) {
      vw = pv(j * nx + (i > 0 ? i - 1 : nx - 1));
      ve = pv(j * nx + (i < nx - 1 ? i + 1 : 0));
      vs = pv((j > 0 ? j - 1 : ny - 1) * nx + i);
      vn = pv((j < ny - 1 ? j + 1 : 0) * nx + i);
    } else {
      vw = i > 0 ? pv(k - 1) : 0.0;
      ve = i < nx - 1 ? pv(k + 1) : 0.0;
      vs = j > 0 ? pv(k - nx) : 0.0;
      vn = j < ny - 1 ? pv(k + nx) : 0.0;
    }
    const double pk = pv(k);
    if (l > 1) pnew[k] = pk;
    const double jp = cell(fx[i], fx[i + 1], fy[j], fy[j + 1], pk, vw, ve, vs, vn);
    const double a = pk + (-gamma) * jp;  // N_VLinearSum(1, p, -gamma, Jp) order
    ap[k] = a;
    acc[0] += pk * a;
  }
  block_sums<1>(acc, parts_a);
}

// The same two kernels on pairs of cells (k, k + 1), k even, for an even
// nx (a pair never straddles a row): 16-byte loads and stores, half the
// load instructions of the stencil kernel, which reads z and p_{l-1} at the
// five points.  Every cell value is the one the cell-wise kernels compute;
// the dot products group their terms by pairs (a thread adds cell 2q, then
// 2q + 1), so they differ from the cell-wise kernels' sums in the last bits
// but are as deterministic.
// XCD-aware walk over [0, npairs): workgroup b runs on XCD b % 8 and, when
// the grid is a multiple of 8, walks the (b % 8)-th of 8 contiguous ranges,
// so that the rows above and below a pair are read by the same XCD at about
// the same time and come from its L2
struct PairWalk {
  int64_t q, end, step;
};
__device__ inline PairWalk xcd_pairs(int64_t npairs)
{
  const unsigned G = gridDim.x, b = blockIdx.x;
  const int parts = (G % 8 == 0) ? 8 : 1;
  const int64_t per = (npairs + parts - 1) / parts;
  const int part = (int)(b % parts);
  PairWalk w;
  w.q = part * per + (int64_t)(b / parts) * kBlock + threadIdx.x;
  w.end = min(npairs, (part + 1) * per);
  w.step = (int64_t)(G / parts) * kBlock;
  return w;
}

__device__ inline double2 ld2(const double* __restrict__ a, int64_t k)
{
  return *reinterpret_cast<const double2*>(a + k);
}
__device__ inline void st2(double* __restrict__ a, int64_t k, double2 v) { *reinterpret_cast<double2*>(a + k) = v; }

template <bool PER>
__global__ __launch_bounds__(kBlock) void k_pcg_dirap_pairs(int64_t nx, int64_t ny, const double* __restrict__ fx,
                                                            const double* __restrict__ fy, double gamma,
                                                            const double* __restrict__ z,
                                                            const double* __restrict__ pold,
                                                            double* __restrict__ pnew, double* __restrict__ ap,
                                                            int nparts, const double* __restrict__ parts_b,
                                                            const double* __restrict__ parts_z,
                                                            double* __restrict__ parts_a, double* __restrict__ st,
                                                            int l)
{
  if (solve_ended(st)) return;
  double beta = 0.0;
  if (l > 1 && test_iteration(nparts, parts_b, parts_z, st, l, &beta)) return;
  const bool first = (l == 1);
  auto pv1 = [&](int64_t k) { return first ? pold[k] : z[k] + beta * pold[k]; };
  auto pv2 = [&](int64_t k) {
    const double2 pp = ld2(pold, k);
    if (first) return pp;
    const double2 zz = ld2(z, k);
    return make_double2(zz.x + beta * pp.x, zz.y + beta * pp.y);
  };
  const double2 zero2 = make_double2(0.0, 0.0);
  double acc[1] = {0.0};
  // pair q is cells 2q, 2q + 1 of row j = 2q / nx, columns i, i + 1
  const int64_t hnx = nx >> 1, npairs = (nx * ny) >> 1;
  const PairWalk wk = xcd_pairs(npairs);
  const int64_t step = wk.step;
  int64_t q = wk.q;
  int64_t j = q / hnx, i2 = q - j * hnx;
  const int64_t dj = step / hnx, di = step - dj * hnx;
  for (; q < wk.end; q += step) {
    const int64_t k = 2 * q, i = 2 * i2;
    const double2 c = pv2(k);
    const double w = i > 0 ? pv1(k - 1) : (PER ? pv1(k + nx - 1) : 0.0);
    const double e = i + 2 < nx ? pv1(k + 2) : (PER ? pv1(k + 2 - nx) : 0.0);
    const double2 sv = j > 0 ? pv2(k - nx) : (PER ? pv2(k + (ny - 1) * nx) : zero2);
    const double2 nv = j < ny - 1 ? pv2(k + nx) : (PER ? pv2(k - (ny - 1) * nx) : zero2);
    if (!first) st2(pnew, k, c);
    const double fyj = fy[j], fyj1 = fy[j + 1];
    const double j0 = cell(fx[i], fx[i + 1], fyj, fyj1, c.x, w, c.y, sv.x, nv.x);
    const double j1 = cell(fx[i + 1], fx[i + 2], fyj, fyj1, c.y, c.x, e, sv.y, nv.y);
    const double a0 = c.x + (-gamma) * j0;  // N_VLinearSum(1, p, -gamma, Jp) order
    const double a1 = c.y + (-gamma) * j1;
    st2(ap, k, make_double2(a0, a1));
    acc[0] += c.x * a0;
    acc[0] += c.y * a1;
    i2 += di;
    j += dj;
    if (i2 >= hnx) {
      i2 -= hnx;
      j++;
    }
  }
  block_sums<1>(acc, parts_a);
}

template <bool JAC>
__global__ __launch_bounds__(kBlock) void k_pcg_update_pairs(int64_t nx, int64_t ny, int nparts,
                                                             const double* __restrict__ parts_a,
                                                             const double* __restrict__ st, int l,
                                                             const double* __restrict__ p,
                                                             const double* __restrict__ ap, double* __restrict__ x,
                                                             double* __restrict__ r, double* __restrict__ z,
                                                             const double* __restrict__ dinv, JacobiDiag J,
                                                             const double* __restrict__ w, int mode,
                                                             double* __restrict__ partials)
{
  if (solve_ended(st)) return;
  const double pap = sum_partials(parts_a, nparts, 1, 0);
  const double alpha = st[l & 1] / pap;
  double acc[2] = {0.0, 0.0};
  const int64_t hnx = nx >> 1, npairs = (nx * ny) >> 1;
  const PairWalk wk = xcd_pairs(npairs);
  const int64_t step = wk.step;
  int64_t q = wk.q;
  int64_t j = q / hnx, i2 = q - j * hnx;
  const int64_t dj = step / hnx, di = step - dj * hnx;
  for (; q < wk.end; q += step) {
    const int64_t k = 2 * q;
    const double2 xv = ld2(x, k), pv = ld2(p, k), rv = ld2(r, k), av = ld2(ap, k);
    st2(x, k, make_double2(xv.x + alpha * pv.x, xv.y + alpha * pv.y));
    const double r0 = rv.x + (-alpha) * av.x, r1 = rv.y + (-alpha) * av.y;
    st2(r, k, make_double2(r0, r1));
    double rw0 = r0, rw1 = r1;
    if (w) {
      const double2 wv = ld2(w, k);
      rw0 = r0 * wv.x;
      rw1 = r1 * wv.y;
    }
    acc[0] += rw0 * rw0;
    acc[0] += rw1 * rw1;
    if (mode != 2) {  // 2: z is computed elsewhere
      double z0 = r0, z1 = r1;
      if (JAC) {
        const int64_t i = 2 * i2;
        const double syj = J.ay[j] + J.ay[j + 1];
        z0 = jac_dinv(J, J.ax[i] + J.ax[i + 1], syj) * r0;
        z1 = jac_dinv(J, J.ax[i + 1] + J.ax[i + 2], syj) * r1;
      } else if (mode == 1) {
        const double2 dv = ld2(dinv, k);
        z0 = dv.x * r0;
        z1 = dv.y * r1;
      }
      st2(z, k, make_double2(z0, z1));
      acc[1] += r0 * z0;
      acc[1] += r1 * z1;
    }
    i2 += di;
    j += dj;
    if (i2 >= hnx) {
      i2 -= hnx;
      j++;
    }
  }
  block_sums<2>(acc, partials);
}

// the residual test of the last iteration of a batch, so that the host sees it
__global__ __launch_bounds__(kBlock) void k_pcg_check(int nparts, const double* __restrict__ parts_b,
                                                      const double* __restrict__ parts_z, double* __restrict__ st,
                                                      int l)
{
  if (solve_ended(st)) return;
  double beta;
  test_iteration(nparts, parts_b, parts_z, st, l, &beta);
}

// the residual test alone of iteration l (callback preconditioner: before
// z = M r is computed for it): converged or non-finite, recorded as
// test_iteration does; otherwise the state is left for the next iteration's
// full test (beta needs r.z)
__global__ __launch_bounds__(kBlock) void k_pcg_rtest(int nparts, const double* __restrict__ parts_b,
                                                      double* __restrict__ st, int l)
{
  if (solve_ended(st)) return;
  const double rr = sum_partials(parts_b, nparts, 2, 0);
  const double rnorm = sqrt(rr);
  int status = kRunning;
  if (!isfinite(rnorm)) status = kNonFinite;
  else if (rnorm <= st[kDelta]) status = kConverged;
  if (threadIdx.x == 0 && status != kRunning) {
    st[kRr] = rr;
    st[kIter] = l;
    st[kStatus] = status;
  }
}

__global__ __launch_bounds__(kBlock) void k_pcg_dot(int64_t n, const double* __restrict__ a,
                                                    const double* __restrict__ b, double* __restrict__ partials,
                                                    const double* __restrict__ st)
{
  if (st && solve_ended(st)) return;
  double acc[1] = {0.0};
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += (int64_t)gridDim.x * kBlock)
    acc[0] += a[k] * b[k];
  block_sums<1>(acc, partials);
}

__global__ void k_pcg_init(double* __restrict__ st, double rz, double delta)
{
  if (threadIdx.x == 0) {
    st[kRz0] = 0.0;
    st[kRz1] = rz;  // iteration 1 reads slot 1
    st[kStatus] = kRunning;
    st[kIter] = 0;
    st[kRr] = 0.0;
    st[kDelta] = delta;
  }
}

unsigned blocks_for(int64_t n)
{
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g > kPcgBlocks) g = kPcgBlocks;
  return (unsigned)(g < 1 ? 1 : g);
}

// the fused solver's workspace (partials and state), allocated on the
// solver's first fused solve and kept with it
void ensure_ws(SUNLinearSolver S)
{
  if (S->d_ws) return;
  CEDA_HIP_CHECK(dev_malloc(&S->d_ws, sizeof(double) * kWsDoubles));
  CEDA_HIP_CHECK(host_malloc(&S->h_ws, sizeof(double) * kWsDoubles, hipHostMallocDefault));
  CEDA_HIP_CHECK(dev_malloc(&S->d_p2, sizeof(double) * vlen(S->r)));
}

}  // namespace

int pcg_fused(ArkMem* A, SUNLinearSolver S, DiffusionProblem* P, int pmode, ARKLsPrecSolveFn ps, void* ud,
              double t, double gamma, N_Vector b, N_Vector x)
{
  // the solver's vectors' stream (the problem is the integrator's user data
  // and lives on the same context)
  hipStream_t s = vstream(S->r);
  const int64_t n = P->n, nx = P->p.nx, ny = P->p.ny;
  if (vlen(x) != n || vlen(b) != n || vlen(S->r) != n) {
    fprintf(stderr, "[ARKLS ERROR] fused PCG: vector length %lld does not match the problem's %lld cells\n",
            (long long)vlen(x), (long long)n);
    return -1;
  }
  const int per = (P->p.kind == DIFFUSION_CEDA_PERIODIC) ? 1 : 0;
  const unsigned g = blocks_for(n);
  ensure_ws(S);
  double* pa = S->d_ws + kOffA;
  double* pb = S->d_ws + kOffB;
  double* pc = S->d_ws + kOffC;
  double* st = S->d_ws + kOffState;
  double* dx = dptr(x);
  double* dr = dptr(S->r);
  double* dz = dptr(S->z);
  double* dp = dptr(S->p);
  double* dap = dptr(S->Ap);
  const double* dinv = (pmode == 1) ? dptr(P->diag) : nullptr;
  const bool wl2 = !(A->ls_reltol > 0.0);
  const double* w = wl2 ? dptr(A->ewt) : nullptr;

  N_VConst(0.0, x);
  N_VScale(1.0, b, S->r);
  const double bnorm = wl2 ? N_VWL2Norm(S->r, A->ewt) : sqrt(N_VDotProd(S->r, S->r));
  const double eplifac = (A->epslin > 0.0) ? A->epslin : 0.05;
  const double delta = (A->ls_reltol > 0.0) ? A->ls_reltol * bnorm
                                            : eplifac * A->nlscoef * sqrt((double)A->n);
  if (krylov_nonfinite("fused PCG", "|b|", bnorm)) return -1;
  if (bnorm <= delta) return 0;
  // z = M r
  if (pmode == 2) {
    A->npsolve++;
    if (ps(t, A->zcur, A->fz, S->r, S->z, gamma, delta, 1, ud)) return -1;
  } else if (pmode == 1) {
    A->npsolve++;
    N_VProd(P->diag, S->r, S->z);
  } else N_VScale(1.0, S->r, S->z);
  N_VScale(1.0, S->z, S->p);
  double rz = N_VDotProd(S->r, S->z);
  if (krylov_nonfinite("fused PCG", "r.z", rz)) return -1;

  // Scalars on the device; the host reads the state once per batch.  The
  // direction of iteration l lives in q[l & 1] (q[1] = S->p holds p_1 = z).
  // With a preconditioner callback (multigrid, the learned network) an
  // iteration is the stencil/direction kernel, the update, the callback's
  // z = M r (enqueued on the same stream) and r.z (k_pcg_dot).  Every kernel
  // returns at once after convergence, but the callback's own work would
  // still be spent, so the host enqueues a fraction (A->pcg_bulk, 3/4 by
  // default) of the previous solve's iteration count without waiting and
  // then tests each residual (k_pcg_rtest, one wait) before it calls the
  // preconditioner.
  double* q[2] = {S->d_p2, dp};
  // pairs of cells with 16-byte accesses: an even row length (every pair in
  // one row, every pair's first cell 16-byte aligned)
  const bool pairs = (nx % 2 == 0);
  // the Jacobi inverse diagonal formed in the update kernel (from the gamma
  // of its last setup, which is what P->diag holds)
  const bool jac = (pmode == 1 && P->diag_gamma_set);
  const JacobiDiag J{per ? P->d_px : P->d_fx, per ? P->d_py : P->d_fy, P->diag_gamma};
  hipLaunchKernelGGL(k_pcg_init, dim3(1), dim3(64), 0, s, st, rz, delta);
  if (pmode == 2) {
    const long bulk = (S->last_iters > 4) ? (long)(A->pcg_bulk * (double)S->last_iters) : 0;
    for (int it = 1; it <= S->maxl; it++) {
      const double* pold = it == 1 ? q[1] : q[(it - 1) & 1];
      if (pairs) {
        if (per)
          hipLaunchKernelGGL(k_pcg_dirap_pairs<true>, dim3(g), dim3(kBlock), 0, s, nx, ny, P->d_fx, P->d_fy,
                             gamma, dz, pold, q[it & 1], dap, (int)g, pb, pc, pa, st, it);
        else
          hipLaunchKernelGGL(k_pcg_dirap_pairs<false>, dim3(g), dim3(kBlock), 0, s, nx, ny, P->d_fx, P->d_fy,
                             gamma, dz, pold, q[it & 1], dap, (int)g, pb, pc, pa, st, it);
        hipLaunchKernelGGL(k_pcg_update_pairs<false>, dim3(g), dim3(kBlock), 0, s, nx, ny, (int)g, pa, st, it,
                           q[it & 1], dap, dx, dr, dz, dinv, J, w, pmode, pb);
      } else {
        hipLaunchKernelGGL(k_pcg_dirap, dim3(g), dim3(kBlock), 0, s, nx, ny, P->d_fx, P->d_fy, per, gamma, dz,
                           pold, q[it & 1], dap, (int)g, pb, pc, pa, st, it);
        hipLaunchKernelGGL(k_pcg_update<false>, dim3(g), dim3(kBlock), 0, s, n, nx, (int)g, pa, 0.0, st, it,
                           q[it & 1], dap, dx, dr, dz, dinv, J, w, pmode, pb);
      }
      // after the bulk, and every kPcgChunk iterations within it, the host
      // waits for the residual test (chunks keep streams that share a
      // hardware queue interleaved)
      if (it > bulk || it % kPcgChunk == 0) {
        hipLaunchKernelGGL(k_pcg_rtest, dim3(1), dim3(kBlock), 0, s, (int)g, pb, st, it);
        CEDA_HIP_CHECK(hipMemcpyAsync(S->h_ws, st, sizeof(double) * 8, hipMemcpyDeviceToHost, s));
        CEDA_HIP_CHECK(hipStreamSynchronize(s));
        const int status = (int)S->h_ws[kStatus];
        if (status != kRunning) {
          const long iters = (long)S->h_ws[kIter];
          A->nli += iters;
          A->njtimes += iters;
          A->npsolve += iters - 1;  // z of the converged iterate is not computed
          S->last_iters = iters;
          if (status == kNonFinite) {
            fprintf(stderr, "[ARKLS ERROR] fused PCG: non-finite residual or direction at iteration %ld "
                            "(|r|^2 = %g); linear solve failed\n", iters, S->h_ws[kRr]);
            return -1;
          }
          return 0;
        }
      }
      if (ps(t, A->zcur, A->fz, S->r, S->z, gamma, delta, 1, ud)) {
        fprintf(stderr, "[ARKLS ERROR] fused PCG: the preconditioner solve failed at iteration %d\n", it);
        return -1;
      }
      hipLaunchKernelGGL(k_pcg_dot, dim3(g), dim3(kBlock), 0, s, n, dr, dz, pc, st);
    }
    hipLaunchKernelGGL(k_pcg_check, dim3(1), dim3(kBlock), 0, s, (int)g, pb, pc, st, S->maxl + 1);
    CEDA_HIP_CHECK(hipMemcpyAsync(S->h_ws, st, sizeof(double) * 8, hipMemcpyDeviceToHost, s));
    CEDA_HIP_CHECK(hipStreamSynchronize(s));
    A->nli += S->maxl;
    A->njtimes += S->maxl;
    A->npsolve += S->maxl;
    S->last_iters = S->maxl;
    return (sqrt(S->h_ws[kRr]) < bnorm) ? 1 : 2;
  }

  int l = 1, batch = 4;
  while (l <= S->maxl) {
    const int last = std::min(S->maxl, l + batch - 1);
    for (int it = l; it <= last; it++) {
      // iteration 1 reads p_1 = z from q[1] and writes nothing
      const double* pold = it == 1 ? q[1] : q[(it - 1) & 1];
      if (pairs) {
        if (per)
          hipLaunchKernelGGL(k_pcg_dirap_pairs<true>, dim3(g), dim3(kBlock), 0, s, nx, ny, P->d_fx, P->d_fy,
                             gamma, dz, pold, q[it & 1], dap, (int)g, pb, nullptr, pa, st, it);
        else
          hipLaunchKernelGGL(k_pcg_dirap_pairs<false>, dim3(g), dim3(kBlock), 0, s, nx, ny, P->d_fx, P->d_fy,
                             gamma, dz, pold, q[it & 1], dap, (int)g, pb, nullptr, pa, st, it);
        if (jac)
          hipLaunchKernelGGL(k_pcg_update_pairs<true>, dim3(g), dim3(kBlock), 0, s, nx, ny, (int)g, pa, st, it,
                             q[it & 1], dap, dx, dr, dz, dinv, J, w, pmode, pb);
        else
          hipLaunchKernelGGL(k_pcg_update_pairs<false>, dim3(g), dim3(kBlock), 0, s, nx, ny, (int)g, pa, st, it,
                             q[it & 1], dap, dx, dr, dz, dinv, J, w, pmode, pb);
      } else {
        hipLaunchKernelGGL(k_pcg_dirap, dim3(g), dim3(kBlock), 0, s, nx, ny, P->d_fx, P->d_fy, per, gamma, dz,
                           pold, q[it & 1], dap, (int)g, pb, nullptr, pa, st, it);
        if (jac)
          hipLaunchKernelGGL(k_pcg_update<true>, dim3(g), dim3(kBlock), 0, s, n, nx, (int)g, pa, 0.0, st, it,
                             q[it & 1], dap, dx, dr, dz, dinv, J, w, pmode, pb);
        else
          hipLaunchKernelGGL(k_pcg_update<false>, dim3(g), dim3(kBlock), 0, s, n, nx, (int)g, pa, 0.0, st, it,
                             q[it & 1], dap, dx, dr, dz, dinv, J, w, pmode, pb);
      }
    }
    hipLaunchKernelGGL(k_pcg_check, dim3(1), dim3(kBlock), 0, s, (int)g, pb, nullptr, st, last + 1);
    CEDA_HIP_CHECK(hipMemcpyAsync(S->h_ws, st, sizeof(double) * 8, hipMemcpyDeviceToHost, s));
    CEDA_HIP_CHECK(hipStreamSynchronize(s));
    const int status = (int)S->h_ws[kStatus];
    const long iters = (long)S->h_ws[kIter];  // iterations run in this solve
    if (status != kRunning) {
      A->nli += iters;
      A->njtimes += iters;
      if (pmode == 1) A->npsolve += iters - 1;  // z of the converged iterate is not used
      S->last_iters = iters;
      if (status == kNonFinite) {
        fprintf(stderr, "[ARKLS ERROR] fused PCG: non-finite residual or direction at iteration %ld "
                        "(|r|^2 = %g); linear solve failed\n", iters, S->h_ws[kRr]);
        return -1;
      }
      return 0;
    }
    l = last + 1;
    batch = std::min(2 * batch, kPcgBatchMax);
  }
  A->nli += S->maxl;
  A->njtimes += S->maxl;
  if (pmode == 1) A->npsolve += S->maxl;
  S->last_iters = S->maxl;
  return (sqrt(S->h_ws[kRr]) < bnorm) ? 1 : 2;
}

}  // namespace ceda_ark
