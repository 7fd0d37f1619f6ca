// Geometric-multigrid preconditioner for the implicit path on gfx950: one
// V-cycle per application for A = I - gamma J (J the 5-point stencil of the
// RHS face tables), with the settings of the reference's hypre PFMG
// preconditioner (preconditioner_hypre.cpp:51-102, 142-160;
// diffusion_2D.hpp:187-193): zero initial guess, one cycle, red/black
// Gauss-Seidel smoothing, by default symmetric (PFMG relax type 2) with 2 pre
// and 2 post sweeps; relax type 3 (nonsymmetric: red then black both before
// and after the coarse correction) and the sweep count are the reference's
// --pfmg_relax / --pfmg_nrelax options (diffusion_2D.cpp:137-151).
//
// Hierarchy (Dirichlet, the manufactured problem): vertex-centred coarsening
// n -> n/2 (coarse node I on fine node 2I+1) until a side would drop below 3;
// coarse operators rediscretised from the fine face tables (coarse face I on
// fine node 2I: (f[2I] + f[2I+1]) / 8).  Periodic (the reference's problem,
// even sides only): coarse node I on fine node 2I, coarse face I spanning
// fine faces 2I-1 and 2I ((f[2I-1] + f[2I]) / 8, f[-1] = f[n-1]), coarsening
// while the coarse sides stay even and at least 4, every neighbour, transfer
// and sweep wrapping around (red/black colours stay consistent across the
// wrap because every side is even);
// full-weighting restriction = P^T/4 of the bilinear prolongation; red then
// black sweeps before the coarse correction, black then red after (the
// adjoint), 4 symmetric pairs on the coarsest grid.  The cycle is therefore
// symmetric and usable inside PCG.
//
// Every operation is element-wise or in-place red/black with a fixed
// expression order and no FMA contraction, so the V-cycle equals the CPU
// oracle's bit for bit.  gamma lives in device memory and every buffer is
// allocated once, so the whole cycle (about 11 launches per level) is
// captured as one HIP graph and replayed per application; it clears with a
// fill kernel, not a memset.
#include <cstring>

#include "diffusion_internal.hpp"

using namespace ceda;

struct MGLevel {
  int64_t nx = 0, ny = 0;
  double *fx = nullptr, *fy = nullptr;  // face tables, nx+1 / ny+1
  double *x = nullptr, *b = nullptr, *r = nullptr;
};

// a level as the kernels see it
struct MGLevelD {
  int64_t nx, ny;
  const double *fx, *fy;
  double *x, *b, *r;
};

struct MGHierarchy {
  int nl = 0;
  MGLevel L[24];
  int per = 0;              // periodic grid
  double* gamma = nullptr;  // device scalar
  double* h_gamma = nullptr;
  hipGraphExec_t graph = nullptr;
  int relax = 2, nrelax = 2;  // the options the graph was captured with
  long cycles = 0;            // V-cycles applied (MGPSolve calls)
  int lc = 0;                 // first level of the single-workgroup coarse cycle
  size_t lc_lds = 0;          // its LDS bytes
  MGLevelD* d_levels = nullptr;
};

namespace {

constexpr int kMGMaxL = 24;
constexpr int kCoarse = 4;
constexpr int64_t kMGCoarseCells = 4096;  // levels this small run in one workgroup

unsigned grid_of(int64_t n)
{
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (unsigned)g;
}

__device__ inline void mg_row(int64_t nx, int64_t ny, const double* __restrict__ fx,
                              const double* __restrict__ fy, double g, const double* x, int64_t i,
                              int64_t j, int per, double& off, double& d)
{
  const int64_t k = j * nx + i;
  double xw, xe, xs, xn;
  if (per) {
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
  off = ((fx[i] * xw + fx[i + 1] * xe) + fy[j] * xs) + fy[j + 1] * xn;
  d = 1.0 + g * ((fx[i] + fx[i + 1]) + (fy[j] + fy[j + 1]));
}

// ---- per-point operations, shared by the per-level kernels and the
// single-workgroup coarse cycle (same arithmetic, same order)

// q-th point of one colour of a red/black Gauss-Seidel sweep (red: i + j
// even), in place
__device__ inline void pt_sweep(const MGLevelD& L, double g, int color, int per, int64_t q)
{
  const int64_t half = (L.nx + 1) / 2;
  const int64_t j = q / half;
  const int64_t i = 2 * (q - j * half) + ((j + color) & 1);
  if (i >= L.nx) return;
  double off, d;
  mg_row(L.nx, L.ny, L.fx, L.fy, g, L.x, i, j, per, off, d);
  L.x[j * L.nx + i] = (L.b[j * L.nx + i] + g * off) / d;
}

__device__ inline void pt_resid(const MGLevelD& L, double g, int per, int64_t k)
{
  const int64_t j = k / L.nx, i = k - j * L.nx;
  double off, d;
  mg_row(L.nx, L.ny, L.fx, L.fy, g, L.x, i, j, per, off, d);
  double t = d * L.x[k];
  t = t - g * off;
  L.r[k] = L.b[k] - t;
}

// full weighting of F's residual onto coarse node q of C: (I, J) = fine node
// (2I+1, 2J+1), or (2I, 2J) with wrap-around on a periodic grid
__device__ inline void pt_restrict(const MGLevelD& F, const MGLevelD& C, int per, int64_t q)
{
  const int64_t nx = F.nx, ny = F.ny;
  const double* __restrict__ r = F.r;
  const int64_t J = q / C.nx, I = q - J * C.nx;
  const int64_t i = per ? 2 * I : 2 * I + 1, j = per ? 2 * J : 2 * J + 1;
  auto R = [&](int64_t a, int64_t c) -> double {
    if (per) return r[((c + ny) % ny) * nx + ((a + nx) % nx)];
    return (a < nx && c < ny) ? r[c * nx + a] : 0.0;
  };
  double t = 4.0 * R(i, j);
  t = t + 2.0 * ((R(i - 1, j) + R(i + 1, j)) + (R(i, j - 1) + R(i, j + 1)));
  t = t + ((R(i - 1, j - 1) + R(i + 1, j - 1)) + (R(i - 1, j + 1) + R(i + 1, j + 1)));
  C.b[q] = t * 0.0625;
}

// bilinear prolongation of C's correction onto fine point k of F, added: a
// fine index coincides with coarse index I (Dirichlet: odd i, I = i/2;
// periodic: even i, I = i/2) or lies between I and I+1 (Dirichlet:
// I = i/2 - 1, zero outside; periodic: I = i/2, wrapping)
__device__ inline void pt_prolong(const MGLevelD& F, const MGLevelD& C, int per, int64_t k)
{
  const int64_t cx = C.nx, cy = C.ny;
  const double* __restrict__ xc = C.x;
  const int64_t j = k / F.nx, i = k - j * F.nx;
  auto XC = [&](int64_t a, int64_t c) -> double {
    if (per) return xc[((c + cy) % cy) * cx + ((a + cx) % cx)];
    return (a >= 0 && a < cx && c >= 0 && c < cy) ? xc[c * cx + a] : 0.0;
  };
  const bool cxo = per ? !(i & 1) : (i & 1), cyo = per ? !(j & 1) : (j & 1);
  const int64_t I = (per || cxo) ? i / 2 : i / 2 - 1;
  const int64_t J = (per || cyo) ? j / 2 : j / 2 - 1;
  double v;
  if (cxo && cyo) v = XC(I, J);
  else if (cyo) v = 0.5 * (XC(I, J) + XC(I + 1, J));
  else if (cxo) v = 0.5 * (XC(I, J) + XC(I, J + 1));
  else v = 0.25 * ((XC(I, J) + XC(I + 1, J)) + (XC(I, J + 1) + XC(I + 1, J + 1)));
  F.x[k] = F.x[k] + v;
}

#define CEDA_MG_GRID_STRIDE(q, n) \
  for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < (n); q += (int64_t)gridDim.x * kBlock)

__global__ __launch_bounds__(kBlock) void k_mg_fill0(int64_t n, double* __restrict__ x)
{
  CEDA_MG_GRID_STRIDE(k, n) x[k] = 0.0;
}

__global__ __launch_bounds__(kBlock) void k_mg_sweep(const MGLevelD L, const double* __restrict__ gp, int color,
                                                     int per)
{
  const double g = *gp;
  CEDA_MG_GRID_STRIDE(q, (L.nx + 1) / 2 * L.ny) pt_sweep(L, g, color, per, q);
}

__global__ __launch_bounds__(kBlock) void k_mg_resid(const MGLevelD L, const double* __restrict__ gp, int per)
{
  const double g = *gp;
  CEDA_MG_GRID_STRIDE(k, L.nx * L.ny) pt_resid(L, g, per, k);
}

__global__ __launch_bounds__(kBlock) void k_mg_restrict(const MGLevelD F, const MGLevelD C, int per)
{
  CEDA_MG_GRID_STRIDE(q, C.nx * C.ny) pt_restrict(F, C, per, q);
}

__global__ __launch_bounds__(kBlock) void k_mg_prolong(const MGLevelD F, const MGLevelD C, int per)
{
  CEDA_MG_GRID_STRIDE(k, F.nx * F.ny) pt_prolong(F, C, per, k);
}

// The V-cycle from level l0 down to the coarsest and back, in one workgroup:
// the same operations in the same order as the per-level launches, separated
// by workgroup barriers.  Every array of these levels (x, b, r and the face
// tables) is staged in LDS, so each of the ~200 barrier-separated phases
// waits for LDS rather than for an L2 round trip; b of level l0 comes in
// from global memory (the restriction of the level above, or the
// preconditioner's input) and x of level l0 goes back (for the prolongation
// to the level above, or as the result).  Used for the trailing levels whose
// arrays fit in a workgroup's LDS (mg_coarse_lds).
constexpr int kCoarseThreads = 1024;
__global__ __launch_bounds__(kCoarseThreads) void k_mg_coarse(const MGLevelD* __restrict__ Ls, int l0, int nl,
                                                               const double* __restrict__ gp, int relax,
                                                               int nrelax, int per)
{
  extern __shared__ double lds[];
  const double g = *gp;
  const int p0 = (relax == 2) ? 1 : 0;
  auto each = [&](int64_t n, auto&& f) {
    for (int64_t q = threadIdx.x; q < n; q += kCoarseThreads) f(q);
    __syncthreads();
  };
  // descriptors of the LDS copies of levels l0..nl-1 (shared, not per lane)
  __shared__ MGLevelD S[kMGMaxL];
  if (threadIdx.x == 0) {
    double* at = lds;
    for (int l = l0; l < nl; l++) {
      const MGLevelD G = Ls[l];
      const int64_t n = G.nx * G.ny;
      double* fx = at;
      double* fy = fx + (G.nx + 1);
      double* x = fy + (G.ny + 1);
      double* b = x + n;
      double* r = b + n;
      at = r + n;
      S[l] = MGLevelD{G.nx, G.ny, fx, fy, x, b, r};
    }
  }
  __syncthreads();
  for (int l = l0; l < nl; l++) {
    double* fx = const_cast<double*>(S[l].fx);
    double* fy = const_cast<double*>(S[l].fy);
    for (int64_t q = threadIdx.x; q <= Ls[l].nx; q += kCoarseThreads) fx[q] = Ls[l].fx[q];
    for (int64_t q = threadIdx.x; q <= Ls[l].ny; q += kCoarseThreads) fy[q] = Ls[l].fy[q];
  }
  each(S[l0].nx * S[l0].ny, [&](int64_t k) { S[l0].b[k] = Ls[l0].b[k]; });
  auto smooth = [&](const MGLevelD& L, int c0, int pairs) {
    for (int k = 0; k < pairs; k++) {
      each((L.nx + 1) / 2 * L.ny, [&](int64_t q) { pt_sweep(L, g, c0, per, q); });
      each((L.nx + 1) / 2 * L.ny, [&](int64_t q) { pt_sweep(L, g, 1 - c0, per, q); });
    }
  };
  for (int l = l0; l < nl - 1; l++) {
    const MGLevelD L = S[l], C = S[l + 1];
    each(L.nx * L.ny, [&](int64_t k) { L.x[k] = 0.0; });
    smooth(L, 0, nrelax);
    each(L.nx * L.ny, [&](int64_t k) { pt_resid(L, g, per, k); });
    each(C.nx * C.ny, [&](int64_t q) { pt_restrict(L, C, per, q); });
  }
  {
    const MGLevelD L = S[nl - 1];
    each(L.nx * L.ny, [&](int64_t k) { L.x[k] = 0.0; });
    smooth(L, 0, kCoarse);
    smooth(L, p0, kCoarse);
  }
  for (int l = nl - 2; l >= l0; l--) {
    const MGLevelD F = S[l], C = S[l + 1];
    each(F.nx * F.ny, [&](int64_t k) { pt_prolong(F, C, per, k); });
    smooth(F, p0, nrelax);
  }
  each(S[l0].nx * S[l0].ny, [&](int64_t k) { Ls[l0].x[k] = S[l0].x[k]; });
}

// LDS bytes k_mg_coarse needs for levels l0..nl-1
size_t mg_coarse_lds(const MGLevel* L, int l0, int nl)
{
  size_t words = 0;
  for (int l = l0; l < nl; l++) words += (size_t)(L[l].nx + 1) + (size_t)(L[l].ny + 1) + 3 * (size_t)(L[l].nx * L[l].ny);
  return words * sizeof(double);
}

MGLevelD dev_level(const MGLevel& L) { return MGLevelD{L.nx, L.ny, L.fx, L.fy, L.x, L.b, L.r}; }

void sweep(MGHierarchy* H, int l, int color, hipStream_t s)
{
  const MGLevel& L = H->L[l];
  hipLaunchKernelGGL(k_mg_sweep, dim3(grid_of((L.nx + 1) / 2 * L.ny)), dim3(kBlock), 0, s, dev_level(L),
                     H->gamma, color, H->per);
}

void vcycle(MGHierarchy* H, int l, hipStream_t s)
{
  if (l >= H->lc) {  // the rest of the cycle in one workgroup
    hipLaunchKernelGGL(k_mg_coarse, dim3(1), dim3(kCoarseThreads), H->lc_lds, s, H->d_levels, l, H->nl,
                       H->gamma, H->relax, H->nrelax, H->per);
    return;
  }
  const MGLevel& L = H->L[l];
  hipLaunchKernelGGL(k_mg_fill0, dim3(grid_of(L.nx * L.ny)), dim3(kBlock), 0, s, L.nx * L.ny, L.x);
  // post-smoothing order: black then red (the adjoint of the pre-smoother,
  // relax type 2) or red then black again (relax type 3)
  const int p0 = (H->relax == 2) ? 1 : 0;
  if (l == H->nl - 1) {
    for (int k = 0; k < kCoarse; k++) {
      sweep(H, l, 0, s);
      sweep(H, l, 1, s);
    }
    for (int k = 0; k < kCoarse; k++) {
      sweep(H, l, p0, s);
      sweep(H, l, 1 - p0, s);
    }
    return;
  }
  for (int k = 0; k < H->nrelax; k++) {
    sweep(H, l, 0, s);
    sweep(H, l, 1, s);
  }
  const MGLevel& C = H->L[l + 1];
  hipLaunchKernelGGL(k_mg_resid, dim3(grid_of(L.nx * L.ny)), dim3(kBlock), 0, s, dev_level(L), H->gamma,
                     H->per);
  hipLaunchKernelGGL(k_mg_restrict, dim3(grid_of(C.nx * C.ny)), dim3(kBlock), 0, s, dev_level(L),
                     dev_level(C), H->per);
  vcycle(H, l + 1, s);
  hipLaunchKernelGGL(k_mg_prolong, dim3(grid_of(L.nx * L.ny)), dim3(kBlock), 0, s, dev_level(L),
                     dev_level(C), H->per);
  for (int k = 0; k < H->nrelax; k++) {
    sweep(H, l, p0, s);
    sweep(H, l, 1 - p0, s);
  }
}

MGHierarchy* build(DiffusionProblem* P)
{
  MGHierarchy* H = new MGHierarchy();
  const bool per = (P->p.kind == DIFFUSION_CEDA_PERIODIC);
  H->per = per ? 1 : 0;
  std::vector<double> fx = P->fx, fy = P->fy;
  int64_t nx = P->p.nx, ny = P->p.ny;
  int l = 0;
  for (;;) {
    MGLevel& L = H->L[l];
    L.nx = nx;
    L.ny = ny;
    CEDA_HIP_CHECK(ceda::dev_malloc(&L.fx, sizeof(double) * (nx + 1)));
    CEDA_HIP_CHECK(ceda::dev_malloc(&L.fy, sizeof(double) * (ny + 1)));
    CEDA_HIP_CHECK(ceda::copy_sync(L.fx, fx.data(), sizeof(double) * (nx + 1), hipMemcpyHostToDevice));
    CEDA_HIP_CHECK(ceda::copy_sync(L.fy, fy.data(), sizeof(double) * (ny + 1), hipMemcpyHostToDevice));
    CEDA_HIP_CHECK(ceda::dev_malloc(&L.x, sizeof(double) * nx * ny));
    CEDA_HIP_CHECK(ceda::dev_malloc(&L.b, sizeof(double) * nx * ny));
    CEDA_HIP_CHECK(ceda::dev_malloc(&L.r, sizeof(double) * nx * ny));
    l++;
    const int64_t cx = nx / 2, cy = ny / 2;
    if (per ? (cx % 2 || cy % 2 || cx < 4 || cy < 4) : (cx < 3 || cy < 3)) break;
    if (l == kMGMaxL) break;
    std::vector<double> cfx(cx + 1), cfy(cy + 1);
    if (per) {
      for (int64_t i = 0; i <= cx; i++) cfx[i] = (fx[i > 0 ? 2 * i - 1 : nx - 1] + fx[2 * i]) / 8.0;
      for (int64_t j = 0; j <= cy; j++) cfy[j] = (fy[j > 0 ? 2 * j - 1 : ny - 1] + fy[2 * j]) / 8.0;
    } else {
      for (int64_t i = 0; i <= cx; i++)
        cfx[i] = (fx[2 * i < nx ? 2 * i : nx] + fx[2 * i + 1 < nx ? 2 * i + 1 : nx]) / 8.0;
      for (int64_t j = 0; j <= cy; j++)
        cfy[j] = (fy[2 * j < ny ? 2 * j : ny] + fy[2 * j + 1 < ny ? 2 * j + 1 : ny]) / 8.0;
    }
    fx.swap(cfx);
    fy.swap(cfy);
    nx = cx;
    ny = cy;
  }
  H->nl = l;
  // the coarse cycle takes the trailing levels whose arrays fit in one
  // workgroup's LDS (none when even the coarsest does not, e.g. a grid that
  // cannot be coarsened)
  int dev = 0;
  CEDA_HIP_CHECK(hipGetDevice(&dev));
  int lds_max = 0;
  CEDA_HIP_CHECK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  H->lc = l;
  // (the kernel's static descriptor array comes on top of the dynamic LDS)
  const size_t lds_dyn_max = (size_t)lds_max - sizeof(MGLevelD) * kMGMaxL;
  while (H->lc > 0 && H->L[H->lc - 1].nx * H->L[H->lc - 1].ny <= kMGCoarseCells &&
         mg_coarse_lds(H->L, H->lc - 1, l) <= lds_dyn_max)
    H->lc--;
  H->lc_lds = (H->lc < l) ? mg_coarse_lds(H->L, H->lc, l) : 0;
  if (H->lc_lds > 65536 &&
      hipFuncSetAttribute((const void*)k_mg_coarse, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)H->lc_lds) != hipSuccess) {
    (void)hipGetLastError();
    while (H->lc < l && mg_coarse_lds(H->L, H->lc, l) > 65536 - sizeof(MGLevelD) * kMGMaxL) H->lc++;
    H->lc_lds = (H->lc < l) ? mg_coarse_lds(H->L, H->lc, l) : 0;
  }
  std::vector<MGLevelD> hl(l);
  for (int k = 0; k < l; k++) hl[k] = dev_level(H->L[k]);
  CEDA_HIP_CHECK(ceda::dev_malloc(&H->d_levels, sizeof(MGLevelD) * l));
  CEDA_HIP_CHECK(ceda::copy_sync(H->d_levels, hl.data(), sizeof(MGLevelD) * l, hipMemcpyHostToDevice));
  CEDA_HIP_CHECK(ceda::dev_malloc(&H->gamma, sizeof(double)));
  CEDA_HIP_CHECK(ceda::host_malloc(&H->h_gamma, sizeof(double), hipHostMallocDefault));
  return H;
}

}  // namespace

namespace ceda {
void mg_destroy(MGHierarchy* H)
{
  if (!H) return;
  for (int l = 0; l < H->nl; l++) {
    double* ps[] = {H->L[l].fx, H->L[l].fy, H->L[l].x, H->L[l].b, H->L[l].r};
    for (double* p : ps)
      if (p) CEDA_HIP_CHECK(ceda::dev_free(p));
  }
  if (H->gamma) CEDA_HIP_CHECK(ceda::dev_free(H->gamma));
  if (H->d_levels) CEDA_HIP_CHECK(ceda::dev_free(H->d_levels));
  if (H->h_gamma) CEDA_HIP_CHECK(ceda::host_free(H->h_gamma));
  if (H->graph) CEDA_HIP_CHECK(ceda::graph_exec_destroy(H->graph));
  delete H;
}
}  // namespace ceda

extern "C" {

int MGPSetup(sunrealtype, N_Vector, N_Vector, sunbooleantype, sunbooleantype* jcurPtr, sunrealtype gamma,
             void* user_data)
{
  DiffusionProblem* P = static_cast<DiffusionProblem*>(user_data);
  // periodic grids need even sides (red/black colours consistent across the wrap)
  if (!P || (P->p.kind == DIFFUSION_CEDA_PERIODIC && (P->p.nx % 2 || P->p.ny % 2))) return -1;
  hipStream_t s = P->exec->stream;
  if (!P->mg) P->mg = build(P);
  MGHierarchy* H = P->mg;
  CEDA_HIP_CHECK(hipStreamSynchronize(s));  // the previous gamma may still be read
  *H->h_gamma = gamma;
  CEDA_HIP_CHECK(hipMemcpyAsync(H->gamma, H->h_gamma, sizeof(double), hipMemcpyHostToDevice, s));
  if (H->graph && (H->relax != P->mg_relax || H->nrelax != P->mg_nrelax)) {
    CEDA_HIP_CHECK(ceda::graph_exec_destroy(H->graph));
    H->graph = nullptr;
  }
  if (!H->graph) {
    H->relax = P->mg_relax;
    H->nrelax = P->mg_nrelax;
    hipGraph_t g;
    // no device-wide synchronisation from other threads meanwhile
    std::lock_guard<std::recursive_mutex> lk(ceda::device_sync_mutex());
    CEDA_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    vcycle(H, 0, s);
    CEDA_HIP_CHECK(hipStreamEndCapture(s, &g));
    CEDA_HIP_CHECK(hipGraphInstantiate(&H->graph, g, nullptr, nullptr, 0));
    CEDA_HIP_CHECK(hipGraphDestroy(g));
  }
  CEDA_HIP_CHECK(hipStreamSynchronize(s));
  if (jcurPtr) *jcurPtr = SUNTRUE;
  return 0;
}

int MGPSolve(sunrealtype, N_Vector, N_Vector, N_Vector r, N_Vector z, sunrealtype, sunrealtype, int,
             void* user_data)
{
  DiffusionProblem* P = static_cast<DiffusionProblem*>(user_data);
  if (!P || !P->mg || !r || !z || N_VGetLength(r) != P->n || N_VGetLength(z) != P->n) return -1;
  MGHierarchy* H = P->mg;
  hipStream_t s = P->exec->stream;
  const size_t bytes = sizeof(double) * (size_t)P->n;
  CEDA_HIP_CHECK(hipMemcpyAsync(H->L[0].b, dptr(r), bytes, hipMemcpyDeviceToDevice, s));
  CEDA_HIP_CHECK(hipGraphLaunch(H->graph, s));
  CEDA_HIP_CHECK(hipMemcpyAsync(dptr(z), H->L[0].x, bytes, hipMemcpyDeviceToDevice, s));
  H->cycles++;
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int DiffusionProblem_MGLevels(const DiffusionProblem* P) { return (P && P->mg) ? P->mg->nl : 0; }

int DiffusionProblem_SetMGOptions(DiffusionProblem* P, int relax, int nrelax)
{
  if (!P || (relax != 2 && relax != 3) || nrelax < 0 || nrelax > 16) return -1;
  P->mg_relax = relax;
  P->mg_nrelax = nrelax;
  return 0;
}

long DiffusionProblem_MGNumCycles(const DiffusionProblem* P) { return (P && P->mg) ? P->mg->cycles : 0; }

}  // extern "C"
