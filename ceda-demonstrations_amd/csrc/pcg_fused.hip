// Fused preconditioned CG for the implicit path: (I - gamma J) x = b with J
// the 5-point stencil of the RHS (the analytic Jacobian-vector product,
// exactly laplacian()'s arithmetic, diffusion.cpp:48-53) and the Jacobi or
// multigrid preconditioner.  The algorithm is the integrator's PCG
// (SUNLinSol_PCG, main.cpp:305; Hestenes-Stiefel with z = M r), with each
// iteration in three kernels instead of about nine vector operations:
//   k_pcg_ap:     Ap = p - gamma * J p, and the partial sums of p.Ap
//   k_pcg_update: x += alpha p, r -= alpha Ap, z = Dinv r (Jacobi), and the
//                 partial sums of the residual norm and of r.z
//   k_pcg_dir:    p = z + beta p
// so an iteration streams about 104 bytes per cell (176 unfused) and waits
// for the host twice.  Sums are formed per workgroup in a fixed order and
// combined by one workgroup, so a solve is deterministic; they are not
// bitwise equal to the separate N_Vector reductions (a different summation
// order), which is why this path is an option of the integrator
// (ARKodeSetDiffusionFastPCG) and the N_Vector path stays the reference.
#include "arkode_internal.hpp"
#include "diffusion_internal.hpp"
#include "stencil_ops.hpp"

using namespace ceda;

namespace ceda_ark {

namespace {

constexpr int kPcgBlocks = 1024;  // workgroups of the reduction kernels

// one cell of J v with the periodic wrap or zero Dirichlet ghosts
__device__ inline double jv_cell(const double* __restrict__ v, const double* __restrict__ fx,
                                 const double* __restrict__ fy, int64_t nx, int64_t ny, int64_t i,
                                 int64_t j, bool per)
{
  const int64_t k = j * nx + i;
  double vw, ve, vs, vn;
  if (per) {
    vw = v[j * nx + (i > 0 ? i - 1 : nx - 1)];
    ve = v[j * nx + (i < nx - 1 ? i + 1 : 0)];
    vs = v[(j > 0 ? j - 1 : ny - 1) * nx + i];
    vn = v[(j < ny - 1 ? j + 1 : 0) * nx + i];
  } else {
    vw = i > 0 ? v[k - 1] : 0.0;
    ve = i < nx - 1 ? v[k + 1] : 0.0;
    vs = j > 0 ? v[k - nx] : 0.0;
    vn = j < ny - 1 ? v[k + nx] : 0.0;
  }
  return cell(fx[i], fx[i + 1], fy[j], fy[j + 1], v[k], vw, ve, vs, vn);
}

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

__global__ __launch_bounds__(kBlock) void k_pcg_ap(int64_t nx, int64_t ny, const double* __restrict__ fx,
                                                   const double* __restrict__ fy, int per, double gamma,
                                                   const double* __restrict__ p, double* __restrict__ ap,
                                                   double* __restrict__ partials)
{
  double acc[1] = {0.0};
  const int64_t n = nx * ny;
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += (int64_t)gridDim.x * kBlock) {
    const int64_t j = k / nx, i = k - j * nx;
    const double jp = jv_cell(p, fx, fy, nx, ny, i, j, per != 0);
    const double a = p[k] + (-gamma) * jp;  // N_VLinearSum(1, p, -gamma, Jp) order
    ap[k] = a;
    acc[0] += p[k] * a;
  }
  block_sums<1>(acc, partials);
}

// mode 0: no preconditioner (z = r), 1: Jacobi (z = dinv r), 2: z computed
// elsewhere (only x and r are updated; acc[1] = 0).  wl2: the residual norm
// is the weighted one, sum (r w)^2, else sum r^2.
__global__ __launch_bounds__(kBlock) void k_pcg_update(int64_t n, double alpha, const double* __restrict__ p,
                                                       const double* __restrict__ ap, double* __restrict__ x,
                                                       double* __restrict__ r, double* __restrict__ z,
                                                       const double* __restrict__ dinv,
                                                       const double* __restrict__ w, int mode,
                                                       double* __restrict__ partials)
{
  double acc[2] = {0.0, 0.0};
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += (int64_t)gridDim.x * kBlock) {
    x[k] = x[k] + alpha * p[k];
    const double rk = r[k] + (-alpha) * ap[k];
    r[k] = rk;
    const double rw = w ? rk * w[k] : rk;
    acc[0] += rw * rw;
    if (mode != 2) {
      const double zk = (mode == 1) ? dinv[k] * rk : rk;
      z[k] = zk;
      acc[1] += rk * zk;
    }
  }
  block_sums<2>(acc, partials);
}

__global__ __launch_bounds__(kBlock) void k_pcg_dot(int64_t n, const double* __restrict__ a,
                                                    const double* __restrict__ b, double* __restrict__ partials)
{
  double acc[1] = {0.0};
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += (int64_t)gridDim.x * kBlock)
    acc[0] += a[k] * b[k];
  block_sums<1>(acc, partials);
}

__global__ __launch_bounds__(kBlock) void k_pcg_dir(int64_t n, double beta, const double* __restrict__ z,
                                                    double* __restrict__ p)
{
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += (int64_t)gridDim.x * kBlock)
    p[k] = z[k] + beta * p[k];
}

// sum of the per-workgroup partials (nacc accumulators), in a fixed order
__global__ __launch_bounds__(kBlock) void k_pcg_final(int nblocks, int nacc, const double* __restrict__ partials,
                                                      double* __restrict__ out)
{
  __shared__ double s[kBlock / kWave];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  for (int a = 0; a < nacc; a++) {
    double v = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += kBlock) v += partials[(int64_t)b * nacc + a];
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    if (lane == 0) s[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = s[0];
      for (int w = 1; w < kBlock / kWave; w++) t += s[w];
      out[a] = t;
    }
    __syncthreads();
  }
}

unsigned blocks_for(int64_t n)
{
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g > kPcgBlocks) g = kPcgBlocks;
  return (unsigned)(g < 1 ? 1 : g);
}

// the nacc sums of the last reduction kernel, on the host
const double* finish(Exec* ex, unsigned g, int nacc)
{
  hipLaunchKernelGGL(k_pcg_final, dim3(1), dim3(kBlock), 0, ex->stream, (int)g, nacc, ex->d_partials,
                     ex->d_result);
  CEDA_HIP_CHECK(hipMemcpyAsync(ex->h_result, ex->d_result, sizeof(double) * nacc, hipMemcpyDeviceToHost,
                                ex->stream));
  CEDA_HIP_CHECK(hipStreamSynchronize(ex->stream));
  return ex->h_result;
}

}  // namespace

int pcg_fused(ArkMem* A, SUNLinearSolver S, DiffusionProblem* P, int pmode, ARKLsPrecSolveFn ps, void* ud,
              double t, double gamma, N_Vector b, N_Vector x)
{
  // the solver's vectors' stream and reduction workspace (the problem is the
  // integrator's user data and lives on the same context)
  Exec* ex = content(S->r)->exec;
  hipStream_t s = ex->stream;
  const int64_t n = P->n, nx = P->p.nx, ny = P->p.ny;
  const int per = (P->p.kind == DIFFUSION_CEDA_PERIODIC) ? 1 : 0;
  const unsigned g = blocks_for(n);
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
  double rnorm = bnorm;
  for (int l = 1; l <= S->maxl; l++) {
    hipLaunchKernelGGL(k_pcg_ap, dim3(g), dim3(kBlock), 0, s, nx, ny, P->d_fx, P->d_fy, per, gamma, dp, dap,
                       ex->d_partials);
    A->njtimes++;
    const double pap = finish(ex, g, 1)[0];
    const double alpha = rz / pap;
    hipLaunchKernelGGL(k_pcg_update, dim3(g), dim3(kBlock), 0, s, n, alpha, dp, dap, dx, dr, dz, dinv, w,
                       pmode, ex->d_partials);
    const double* sums = finish(ex, g, 2);
    A->nli++;
    rnorm = sqrt(sums[0]);
    if (rnorm <= delta) return 0;
    double rznew;
    if (pmode == 2) {
      A->npsolve++;
      if (ps(t, A->zcur, A->fz, S->r, S->z, gamma, delta, 1, ud)) return -1;
      hipLaunchKernelGGL(k_pcg_dot, dim3(g), dim3(kBlock), 0, s, n, dr, dz, ex->d_partials);
      rznew = finish(ex, g, 1)[0];
    } else {
      if (pmode == 1) A->npsolve++;
      rznew = sums[1];
    }
    const double beta = rznew / rz;
    rz = rznew;
    hipLaunchKernelGGL(k_pcg_dir, dim3(g), dim3(kBlock), 0, s, n, beta, dz, dp);
  }
  return (rnorm < bnorm) ? 1 : 2;
}

}  // namespace ceda_ark
