// Native HIP N_Vector for gfx950.
//
// Replaces the reference's nvector_parallel (diffusion_2D/main.cpp:176) behind
// the SUNDIALS ops table; ownership follows nvector_gkylzero.c:111-203.
// Element-wise kernels stream two FP64 values per lane per pass (16-byte
// global loads/stores, 1 KiB per wavefront instruction) with a grid-stride
// loop; every formula is evaluated in the order written and the library is
// compiled without FMA contraction, so results are bitwise equal to the CPU
// oracle's serial loops.  Reductions are two-pass and deterministic: a fixed
// grid reduces each lane's strided elements in index order, then wave64
// butterflies (__shfl_xor), then LDS across the 4 wavefronts of a workgroup,
// then one workgroup combines the per-group partials in fixed order.
#include <cfloat>
#include <cmath>
#include <cstring>

#include "nvhip_internal.hpp"

using namespace ceda;

namespace {

// ------------------------------------------------------------------ maps
template <class F>
__global__ __launch_bounds__(kBlock) void k_map2(int64_t n, F f)
{
  const int64_t np = n >> 1;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < np; k += stride)
    f.pair(2 * k);
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) f.one(n - 1);
}

template <class F>
__global__ __launch_bounds__(kBlock) void k_map1(int64_t n, F f)
{
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) f.one(i);
}

__device__ inline double2 ld2(const double* p, int64_t i)
{
  return *reinterpret_cast<const double2*>(p + i);
}
__device__ inline void st2(double* p, int64_t i, double2 v)
{
  *reinterpret_cast<double2*>(p + i) = v;
}

template <int NI, class E>
struct MapOp {
  const double* x[NI > 0 ? NI : 1];
  double* z;
  E e;
  __device__ void one(int64_t i) const
  {
    double v[NI > 0 ? NI : 1];
#pragma unroll
    for (int k = 0; k < NI; k++) v[k] = x[k][i];
    z[i] = e(v);
  }
  __device__ void pair(int64_t i) const
  {
    double a[NI > 0 ? NI : 1], b[NI > 0 ? NI : 1];
#pragma unroll
    for (int k = 0; k < NI; k++) {
      const double2 t = ld2(x[k], i);
      a[k] = t.x;
      b[k] = t.y;
    }
    double2 r;
    r.x = e(a);
    r.y = e(b);
    st2(z, i, r);
  }
};

struct ELinSum { double a, b; __device__ double operator()(const double* v) const { return a * v[0] + b * v[1]; } };
struct EConst { double c; __device__ double operator()(const double*) const { return c; } };
struct EProd { __device__ double operator()(const double* v) const { return v[0] * v[1]; } };
struct EDiv { __device__ double operator()(const double* v) const { return v[0] / v[1]; } };
struct EScale { double c; __device__ double operator()(const double* v) const { return c * v[0]; } };
struct EAbs { __device__ double operator()(const double* v) const { return fabs(v[0]); } };
struct EInv { __device__ double operator()(const double* v) const { return 1.0 / v[0]; } };
struct EAddConst { double b; __device__ double operator()(const double* v) const { return v[0] + b; } };
struct ECompare { double c; __device__ double operator()(const double* v) const { return (fabs(v[0]) >= c) ? 1.0 : 0.0; } };

inline bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

template <int NI, class E>
void launch_map(N_Vector z, const double* const* xs, E e)
{
  MapOp<NI, E> op;
  bool al = aligned16(dptr(z));
  for (int k = 0; k < NI; k++) {
    op.x[k] = xs[k];
    al = al && aligned16(xs[k]);
  }
  op.z = dptr(z);
  op.e = e;
  const int64_t n = vlen(z);
  if (n == 0) return;
  if (al) hipLaunchKernelGGL((k_map2<MapOp<NI, E>>), dim3(stream_grid(n)), dim3(kBlock), 0, vstream(z), n, op);
  else hipLaunchKernelGGL((k_map1<MapOp<NI, E>>), dim3(stream_grid(2 * n)), dim3(kBlock), 0, vstream(z), n, op);
  CEDA_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------- fused operations
struct LCArgs {
  int nv;
  double c[NVHIP_MAX_FUSED];
  const double* x[NVHIP_MAX_FUSED];
  double* z;
};

template <int NV>  // NV > 0: compile-time count; NV == 0: runtime a.nv
struct LCOp {
  LCArgs a;
  __device__ int count() const { return NV > 0 ? NV : a.nv; }
  __device__ void one(int64_t i) const
  {
    double acc = a.c[0] * a.x[0][i];
#pragma unroll
    for (int k = 1; k < (NV > 0 ? NV : NVHIP_MAX_FUSED); k++) {
      if (NV == 0 && k >= a.nv) break;
      acc = acc + a.c[k] * a.x[k][i];
    }
    a.z[i] = acc;
  }
  __device__ void pair(int64_t i) const
  {
    double2 t = ld2(a.x[0], i);
    double ax = a.c[0] * t.x, ay = a.c[0] * t.y;
#pragma unroll
    for (int k = 1; k < (NV > 0 ? NV : NVHIP_MAX_FUSED); k++) {
      if (NV == 0 && k >= a.nv) break;
      t = ld2(a.x[k], i);
      ax = ax + a.c[k] * t.x;
      ay = ay + a.c[k] * t.y;
    }
    st2(a.z, i, make_double2(ax, ay));
  }
};

template <int NV>
void launch_lc(const LCArgs& a, int64_t n, hipStream_t s, bool al)
{
  LCOp<NV> op{a};
  if (al) hipLaunchKernelGGL(k_map2<LCOp<NV>>, dim3(stream_grid(n)), dim3(kBlock), 0, s, n, op);
  else hipLaunchKernelGGL(k_map1<LCOp<NV>>, dim3(stream_grid(2 * n)), dim3(kBlock), 0, s, n, op);
  CEDA_HIP_CHECK(hipGetLastError());
}

struct SAMArgs {
  int nv;
  double a[NVHIP_MAX_FUSED];
  const double* x;
  const double* y[NVHIP_MAX_FUSED];
  double* z[NVHIP_MAX_FUSED];
};

struct SAMOp {
  SAMArgs a;
  __device__ void one(int64_t i) const
  {
    const double xv = a.x[i];
    for (int k = 0; k < a.nv; k++) a.z[k][i] = a.a[k] * xv + a.y[k][i];
  }
  __device__ void pair(int64_t i) const
  {
    const double2 xv = ld2(a.x, i);
    for (int k = 0; k < a.nv; k++) {
      const double2 yv = ld2(a.y[k], i);
      st2(a.z[k], i, make_double2(a.a[k] * xv.x + yv.x, a.a[k] * xv.y + yv.y));
    }
  }
};

// ------------------------------------------------------------ reductions
enum RKind { RSUM = 0, RMAX = 1, RMIN = 2 };

template <int K>
__device__ inline double rcombine(double a, double b)
{
  if (K == RSUM) return a + b;
  if (K == RMAX) return fmax(a, b);
  return fmin(a, b);
}
template <int K>
__device__ inline double rinit()
{
  if (K == RSUM) return 0.0;
  if (K == RMAX) return 0.0;  // max-type reductions here are of non-negative values
  return DBL_MAX;
}

template <int K>
__device__ inline double wave_reduce(double v)
{
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v = rcombine<K>(v, __shfl_xor(v, off, kWave));
  return v;
}

// per-workgroup partials: one pass over the grid-stride pairs of each lane
template <int K, int NACC, class R>
__global__ __launch_bounds__(kBlock) void k_reduce(int64_t n, int nacc, R r, double* partials)
{
  double acc[NACC];
#pragma unroll
  for (int k = 0; k < NACC; k++) acc[k] = rinit<K>();
  const int64_t np = n >> 1;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < np; k += stride) r.pair(acc, 2 * k);
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) r.one(acc, n - 1);
  __shared__ double s[kBlock / kWave][NACC];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < NACC; k++) {
    if (k >= nacc) break;
    const double v = wave_reduce<K>(acc[k]);
    if (lane == 0) s[wave][k] = v;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nacc; k += kBlock) {
    double v = s[0][k];
#pragma unroll
    for (int w = 1; w < kBlock / kWave; w++) v = rcombine<K>(v, s[w][k]);
    partials[(int64_t)blockIdx.x * nacc + k] = v;
  }
}

template <int K>
__global__ __launch_bounds__(kBlock) void k_reduce_final(int nblocks, int nacc,
                                                         const double* partials, double* out)
{
  __shared__ double s[kBlock / kWave];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  for (int k = 0; k < nacc; k++) {
    double v = rinit<K>();
    for (int b = threadIdx.x; b < nblocks; b += kBlock) v = rcombine<K>(v, partials[(int64_t)b * nacc + k]);
    v = wave_reduce<K>(v);
    if (lane == 0) s[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = s[0];
      for (int w = 1; w < kBlock / kWave; w++) t = rcombine<K>(t, s[w]);
      out[k] = t;
    }
    __syncthreads();
  }
}

template <int K, int NACC, class R>
const double* run_reduce(Exec* ex, int64_t n, int nacc, const R& r)
{
  const unsigned g = reduce_grid(n);
  hipLaunchKernelGGL((k_reduce<K, NACC, R>), dim3(g), dim3(kBlock), 0, ex->stream, n, nacc, r, ex->d_partials);
  CEDA_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_reduce_final<K>, dim3(1), dim3(kBlock), 0, ex->stream, (int)g, nacc,
                     ex->d_partials, ex->d_result);
  CEDA_HIP_CHECK(hipGetLastError());
  CEDA_HIP_CHECK(hipMemcpyAsync(ex->h_result, ex->d_result, sizeof(double) * nacc,
                                hipMemcpyDeviceToHost, ex->stream));
  CEDA_HIP_CHECK(hipStreamSynchronize(ex->stream));
  return ex->h_result;
}

// reduction element functors: one(acc, i) and pair(acc, i)
struct RDot {
  const double *x, *y;
  __device__ void one(double* a, int64_t i) const { a[0] += x[i] * y[i]; }
  __device__ void pair(double* a, int64_t i) const
  {
    const double2 u = ld2(x, i), v = ld2(y, i);
    a[0] += u.x * v.x;
    a[0] += u.y * v.y;
  }
};
struct RWSq {
  const double *x, *w;
  __device__ void one(double* a, int64_t i) const
  {
    const double p = x[i] * w[i];
    a[0] += p * p;
  }
  __device__ void pair(double* a, int64_t i) const
  {
    const double2 u = ld2(x, i), v = ld2(w, i);
    const double p = u.x * v.x, q = u.y * v.y;
    a[0] += p * p;
    a[0] += q * q;
  }
};
struct RWSqMask {
  const double *x, *w, *id;
  __device__ void one(double* a, int64_t i) const
  {
    if (id[i] > 0.0) {
      const double p = x[i] * w[i];
      a[0] += p * p;
    }
  }
  __device__ void pair(double* a, int64_t i) const { one(a, i); one(a, i + 1); }
};
struct RMaxAbs {
  const double* x;
  __device__ void one(double* a, int64_t i) const { a[0] = fmax(a[0], fabs(x[i])); }
  __device__ void pair(double* a, int64_t i) const
  {
    const double2 u = ld2(x, i);
    a[0] = fmax(a[0], fabs(u.x));
    a[0] = fmax(a[0], fabs(u.y));
  }
};
struct RMinV {
  const double* x;
  __device__ void one(double* a, int64_t i) const { a[0] = fmin(a[0], x[i]); }
  __device__ void pair(double* a, int64_t i) const { one(a, i); one(a, i + 1); }
};
struct RL1 {
  const double* x;
  __device__ void one(double* a, int64_t i) const { a[0] += fabs(x[i]); }
  __device__ void pair(double* a, int64_t i) const { one(a, i); one(a, i + 1); }
};
struct RInvTest {  // z = 1/x where x != 0; counts zeros
  const double* x;
  double* z;
  __device__ void one(double* a, int64_t i) const
  {
    const double v = x[i];
    if (v == 0.0) a[0] = 1.0;
    else z[i] = 1.0 / v;
  }
  __device__ void pair(double* a, int64_t i) const { one(a, i); one(a, i + 1); }
};
struct RConstrMask {
  const double *c, *x;
  double* m;
  __device__ void one(double* a, int64_t i) const
  {
    m[i] = 0.0;
    const double ci = c[i];
    if (ci == 0.0) return;
    const bool test = (fabs(ci) > 1.5 && x[i] * ci <= 0.0) || (fabs(ci) > 0.5 && x[i] * ci < 0.0);
    if (test) {
      m[i] = 1.0;
      a[0] = 1.0;
    }
  }
  __device__ void pair(double* a, int64_t i) const { one(a, i); one(a, i + 1); }
};
struct RMinQuot {
  const double *num, *den;
  __device__ void one(double* a, int64_t i) const
  {
    if (den[i] != 0.0) a[0] = fmin(a[0], num[i] / den[i]);
  }
  __device__ void pair(double* a, int64_t i) const { one(a, i); one(a, i + 1); }
};
template <int NACC>
struct RDotMulti {
  const double* x;
  const double* y[NACC];
  int nv;
  __device__ void one(double* a, int64_t i) const
  {
    const double xv = x[i];
#pragma unroll
    for (int k = 0; k < NACC; k++) {
      if (k >= nv) break;
      a[k] += xv * y[k][i];
    }
  }
  __device__ void pair(double* a, int64_t i) const
  {
    const double2 xv = ld2(x, i);
#pragma unroll
    for (int k = 0; k < NACC; k++) {
      if (k >= nv) break;
      const double2 yv = ld2(y[k], i);
      a[k] += xv.x * yv.x;
      a[k] += xv.y * yv.y;
    }
  }
};

struct RStsErr {
  const double *yn, *ycur, *fn, *fs;
  double ch, rtol, atol;
  __device__ double term(double y, double yc, double f, double g) const
  {
    double e = 0.8 * y;
    e = e + -0.8 * yc;
    e = e + ch * f;
    e = e + ch * g;
    const double w = 1.0 / (rtol * fabs(y) + atol);
    return e * w;
  }
  __device__ void one(double* a, int64_t i) const
  {
    const double p = term(yn[i], ycur[i], fn[i], fs[i]);
    a[0] += p * p;
  }
  __device__ void pair(double* a, int64_t i) const
  {
    const double2 y = ld2(yn, i), yc = ld2(ycur, i), f = ld2(fn, i), g = ld2(fs, i);
    const double p = term(y.x, yc.x, f.x, g.x), q = term(y.y, yc.y, f.y, g.y);
    a[0] += p * p;
    a[0] += q * q;
  }
};

__global__ __launch_bounds__(kBlock) void k_seed(int64_t n, uint64_t seed, double* v)
{
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    uint64_t x = seed + (uint64_t)i;
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    x = x ^ (x >> 31);
    v[i] = (double)(x >> 11) * 0x1.0p-53 - 0.5;
  }
}

N_Vector make_empty(sunindextype length, SUNContext ctx)
{
  N_Vector v = N_VNewEmpty(ctx);
  if (!v) return nullptr;
  N_Vector_Ops o = v->ops;
  o->nvgetvectorid = N_VGetVectorID_HipNative;
  o->nvclone = N_VClone_HipNative;
  o->nvcloneempty = N_VCloneEmpty_HipNative;
  o->nvdestroy = N_VDestroy_HipNative;
  o->nvspace = N_VSpace_HipNative;
  o->nvgetarraypointer = N_VGetArrayPointer_HipNative;
  o->nvgetdevicearraypointer = N_VGetDeviceArrayPointer_HipNative;
  o->nvsetarraypointer = N_VSetArrayPointer_HipNative;
  o->nvgetlength = N_VGetLength_HipNative;
  o->nvgetlocallength = N_VGetLength_HipNative;
  o->nvlinearsum = N_VLinearSum_HipNative;
  o->nvconst = N_VConst_HipNative;
  o->nvprod = N_VProd_HipNative;
  o->nvdiv = N_VDiv_HipNative;
  o->nvscale = N_VScale_HipNative;
  o->nvabs = N_VAbs_HipNative;
  o->nvinv = N_VInv_HipNative;
  o->nvaddconst = N_VAddConst_HipNative;
  o->nvdotprod = N_VDotProd_HipNative;
  o->nvmaxnorm = N_VMaxNorm_HipNative;
  o->nvwrmsnorm = N_VWrmsNorm_HipNative;
  o->nvwrmsnormmask = N_VWrmsNormMask_HipNative;
  o->nvmin = N_VMin_HipNative;
  o->nvwl2norm = N_VWL2Norm_HipNative;
  o->nvl1norm = N_VL1Norm_HipNative;
  o->nvcompare = N_VCompare_HipNative;
  o->nvinvtest = N_VInvTest_HipNative;
  o->nvconstrmask = N_VConstrMask_HipNative;
  o->nvminquotient = N_VMinQuotient_HipNative;
  o->nvlinearcombination = N_VLinearCombination_HipNative;
  o->nvscaleaddmulti = N_VScaleAddMulti_HipNative;
  o->nvdotprodmulti = N_VDotProdMulti_HipNative;
  o->nvdotprodlocal = N_VDotProd_HipNative;
  o->nvmaxnormlocal = N_VMaxNorm_HipNative;
  o->nvminlocal = N_VMin_HipNative;
  o->nvl1normlocal = N_VL1Norm_HipNative;
  o->nvinvtestlocal = N_VInvTest_HipNative;
  o->nvconstrmasklocal = N_VConstrMask_HipNative;
  o->nvminquotientlocal = N_VMinQuotient_HipNative;
  o->nvwsqrsumlocal = N_VWSqrSumLocal_HipNative;
  o->nvwsqrsummasklocal = N_VWSqrSumMaskLocal_HipNative;
  o->nvdotprodmultilocal = N_VDotProdMulti_HipNative;
  o->nvbufsize = N_VBufSize_HipNative;
  o->nvbufpack = N_VBufPack_HipNative;
  o->nvbufunpack = N_VBufUnpack_HipNative;
  o->nvprint = N_VPrint_HipNative;
  o->nvprintfile = N_VPrintFile_HipNative;
  NVContent* c = new NVContent();
  c->length = length;
  c->exec = context_exec(ctx);
  v->content = c;
  return v;
}

template <int K, class R>
double reduce1(N_Vector x, const R& r)
{
  return run_reduce<K, 1, R>(content(x)->exec, vlen(x), 1, r)[0];
}

}  // namespace

namespace ceda {

void raw_linear_sum(hipStream_t s, int64_t n, double a, const double* x, double b,
                    const double* y, double* z)
{
  if (n == 0) return;
  MapOp<2, ELinSum> op;
  op.x[0] = x;
  op.x[1] = y;
  op.z = z;
  op.e = ELinSum{a, b};
  hipLaunchKernelGGL((k_map2<MapOp<2, ELinSum>>), dim3(stream_grid(n)), dim3(kBlock), 0, s, n, op);
  CEDA_HIP_CHECK(hipGetLastError());
}

double raw_sts_error_wsqr(Exec* ex, int64_t n, const double* yn, const double* ycur,
                          const double* fn, const double* fs, double h, double rtol,
                          double atol)
{
  RStsErr r{yn, ycur, fn, fs, 0.4 * h, rtol, atol};
  return run_reduce<RSUM, 1, RStsErr>(ex, n, 1, r)[0];
}

// Many problems' raw_sts_error_wsqr in two launches.  Problem q's reduction
// is the grid of reduce_grid(n_q) workgroups of k_reduce<RSUM, 1, RStsErr>,
// run as virtual workgroups vb0_q .. vb0_q + g_q - 1 of one launch (the same
// grid-stride partition with g_q in place of gridDim.x, the same wavefront
// and workgroup trees), then k_reduce_final's combine with one workgroup
// per problem: every sum is bitwise the single-problem one.
__global__ __launch_bounds__(kBlock) void k_sts_err_batch(const StsErrItem* __restrict__ items, int nitems,
                                                          double* __restrict__ partials)
{
  // the item whose virtual workgroups contain this one (vb0 ascending)
  int lo = 0, hi = nitems - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].vb0 <= (int)blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const StsErrItem it = items[lo];
  const int b = (int)blockIdx.x - it.vb0;
  const RStsErr r{it.yn, it.ycur, it.fn, it.fs, it.ch, it.rtol, it.atol};
  double acc[1] = {rinit<RSUM>()};
  const int64_t n = it.n, np = n >> 1;
  const int64_t stride = (int64_t)it.g * kBlock;
  for (int64_t k = (int64_t)b * kBlock + threadIdx.x; k < np; k += stride) r.pair(acc, 2 * k);
  if ((n & 1) && b == 0 && threadIdx.x == 0) r.one(acc, n - 1);
  __shared__ double sm[kBlock / kWave];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const double v = wave_reduce<RSUM>(acc[0]);
  if (lane == 0) sm[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = sm[0];
#pragma unroll
    for (int w = 1; w < kBlock / kWave; w++) t = rcombine<RSUM>(t, sm[w]);
    partials[it.vb0 + b] = t;
  }
}

__global__ __launch_bounds__(kBlock) void k_sts_err_batch_final(const StsErrItem* __restrict__ items,
                                                                const double* __restrict__ partials,
                                                                double* __restrict__ out)
{
  const StsErrItem it = items[blockIdx.x];
  __shared__ double sm[kBlock / kWave];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  double v = rinit<RSUM>();
  for (int b = threadIdx.x; b < it.g; b += kBlock) v = rcombine<RSUM>(v, partials[it.vb0 + b]);
  v = wave_reduce<RSUM>(v);
  if (lane == 0) sm[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = sm[0];
    for (int w = 1; w < kBlock / kWave; w++) t = rcombine<RSUM>(t, sm[w]);
    out[blockIdx.x] = t;
  }
}

int sts_err_blocks(int64_t n) { return (int)reduce_grid(n); }

void launch_sts_err_batch(const StsErrItem* items, int nitems, int nvblocks, double* partials, double* out,
                          hipStream_t s)
{
  if (nitems <= 0 || nvblocks <= 0) return;
  hipLaunchKernelGGL(k_sts_err_batch, dim3(nvblocks), dim3(kBlock), 0, s, items, nitems, partials);
  hipLaunchKernelGGL(k_sts_err_batch_final, dim3(nitems), dim3(kBlock), 0, s, items, partials, out);
  CEDA_HIP_CHECK(hipGetLastError());
}

}  // namespace ceda

// DotProdMulti with more than 8 vectors: groups of 8 vectors in one launch
// (blockIdx.y = group), 8 accumulators per thread instead of up to 64, so the
// kernel keeps enough wavefronts in flight to stream at HBM rate.  Every dot
// product is accumulated and combined exactly as by k_reduce with
// RDotMulti (same grid-stride partition, same wavefront and workgroup trees,
// same partial layout), so the results are bitwise those of the one-group
// kernel.
struct DotGroups {
  const double* x;
  const double* y[NVHIP_MAX_FUSED];
  int nv;
};
__global__ __launch_bounds__(kBlock) void k_dotmulti_groups(int64_t n, DotGroups r,
                                                            double* partials)
{
  const int base = blockIdx.y * 8, ntot = 8 * gridDim.y;
  double acc[8];
#pragma unroll
  for (int j = 0; j < 8; j++) acc[j] = 0.0;
  const int64_t np = n >> 1;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < np; k += stride) {
    const double2 xv = ld2(r.x, 2 * k);
#pragma unroll
    for (int j = 0; j < 8; j++) {
      if (base + j >= r.nv) break;
      const double2 yv = ld2(r.y[base + j], 2 * k);
      acc[j] += xv.x * yv.x;
      acc[j] += xv.y * yv.y;
    }
  }
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
    for (int j = 0; j < 8 && base + j < r.nv; j++) acc[j] += r.x[n - 1] * r.y[base + j][n - 1];
  }
  __shared__ double sh[kBlock / kWave][8];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const double v = wave_reduce<RSUM>(acc[j]);
    if (lane == 0) sh[wave][j] = v;
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    double v = sh[0][threadIdx.x];
#pragma unroll
    for (int w = 1; w < kBlock / kWave; w++) v = rcombine<RSUM>(v, sh[w][threadIdx.x]);
    partials[(int64_t)blockIdx.x * ntot + base + threadIdx.x] = v;
  }
}

static void dotmulti_groups(int nvec, N_Vector x, N_Vector* Y, sunrealtype* d)
{
  DotGroups r;
  r.x = dptr(x);
  r.nv = nvec;
  for (int k = 0; k < NVHIP_MAX_FUSED; k++) r.y[k] = (k < nvec) ? dptr(Y[k]) : nullptr;
  Exec* ex = content(x)->exec;
  const int64_t n = vlen(x);
  const unsigned g = reduce_grid(n);
  const int ngroups = (nvec + 7) / 8;
  hipLaunchKernelGGL(k_dotmulti_groups, dim3(g, ngroups), dim3(kBlock), 0, ex->stream, n, r, ex->d_partials);
  CEDA_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_reduce_final<RSUM>, dim3(1), dim3(kBlock), 0, ex->stream, (int)g, 8 * ngroups,
                     ex->d_partials, ex->d_result);
  CEDA_HIP_CHECK(hipGetLastError());
  CEDA_HIP_CHECK(hipMemcpyAsync(ex->h_result, ex->d_result, sizeof(double) * nvec, hipMemcpyDeviceToHost,
                                ex->stream));
  CEDA_HIP_CHECK(hipStreamSynchronize(ex->stream));
  for (int k = 0; k < nvec; k++) d[k] = ex->h_result[k];
}

template <int NACC>
static void dotmulti(int nvec, N_Vector x, N_Vector* Y, sunrealtype* d)
{
  RDotMulti<NACC> r;
  r.x = dptr(x);
  r.nv = nvec;
  for (int k = 0; k < NACC; k++) r.y[k] = (k < nvec) ? dptr(Y[k]) : nullptr;
  const double* out = run_reduce<RSUM, NACC, RDotMulti<NACC>>(content(x)->exec, vlen(x), nvec, r);
  for (int k = 0; k < nvec; k++) d[k] = out[k];
}

// =================================================================== C ABI
extern "C" {

N_Vector N_VNewEmpty_HipNative(sunindextype length, SUNContext sunctx)
{
  if (length < 0 || !sunctx) return nullptr;
  return make_empty(length, sunctx);
}

N_Vector N_VNew_HipNative(sunindextype length, SUNContext sunctx)
{
  N_Vector v = N_VNewEmpty_HipNative(length, sunctx);
  if (!v) return nullptr;
  NVContent* c = content(v);
  if (length > 0) CEDA_HIP_CHECK(ceda::dev_malloc(&c->dev, sizeof(double) * (size_t)length));
  c->own_data = true;
  return v;
}

N_Vector N_VMake_HipNative(sunindextype length, sunrealtype* d_data, SUNContext sunctx)
{
  N_Vector v = N_VNewEmpty_HipNative(length, sunctx);
  if (!v) return nullptr;
  content(v)->dev = d_data;
  content(v)->own_data = false;
  return v;
}

N_Vector_ID N_VGetVectorID_HipNative(N_Vector) { return SUNDIALS_NVEC_CUSTOM; }

N_Vector N_VCloneEmpty_HipNative(N_Vector w)
{
  if (!w) return nullptr;
  N_Vector v = N_VNewEmpty(w->sunctx);
  if (!v) return nullptr;
  if (N_VCopyOps(w, v)) {
    N_VDestroy(v);
    return nullptr;
  }
  NVContent* c = new NVContent();
  c->length = vlen(w);
  c->exec = content(w)->exec;
  v->content = c;
  return v;
}

N_Vector N_VClone_HipNative(N_Vector w)
{
  N_Vector v = N_VCloneEmpty_HipNative(w);
  if (!v) return nullptr;
  NVContent* c = content(v);
  if (c->length > 0) CEDA_HIP_CHECK(ceda::dev_malloc(&c->dev, sizeof(double) * (size_t)c->length));
  c->own_data = true;
  return v;
}

void N_VDestroy_HipNative(N_Vector v)
{
  if (!v) return;
  NVContent* c = content(v);
  if (c) {
    if (c->own_data && c->dev) {
      CEDA_HIP_CHECK(hipStreamSynchronize(c->exec->stream));
      CEDA_HIP_CHECK(ceda::dev_free(c->dev));
    }
    if (c->host) CEDA_HIP_CHECK(ceda::host_free(c->host));
    delete c;
    v->content = nullptr;
  }
  N_VFreeEmpty(v);
}

void N_VSpace_HipNative(N_Vector v, sunindextype* lrw, sunindextype* liw)
{
  *lrw = vlen(v);
  *liw = 2;
}

sunrealtype* N_VGetArrayPointer_HipNative(N_Vector v)
{
  NVContent* c = content(v);
  if (!c->host && c->length > 0)
    CEDA_HIP_CHECK(ceda::host_malloc(&c->host, sizeof(double) * (size_t)c->length, hipHostMallocDefault));
  return c->host;
}

sunrealtype* N_VGetDeviceArrayPointer_HipNative(N_Vector v) { return dptr(v); }

void N_VSetArrayPointer_HipNative(sunrealtype* d_data, N_Vector v)
{
  NVContent* c = content(v);
  if (c->own_data && c->dev) CEDA_HIP_CHECK(ceda::dev_free(c->dev));
  c->dev = d_data;
  c->own_data = false;
}

sunindextype N_VGetLength_HipNative(N_Vector v) { return vlen(v); }

SUNErrCode N_VCopyToDevice_HipNative(N_Vector v)
{
  NVContent* c = content(v);
  if (!c->host) return -1;
  CEDA_HIP_CHECK(hipMemcpyAsync(c->dev, c->host, sizeof(double) * (size_t)c->length,
                                hipMemcpyHostToDevice, c->exec->stream));
  CEDA_HIP_CHECK(hipStreamSynchronize(c->exec->stream));
  return 0;
}

SUNErrCode N_VCopyFromDevice_HipNative(N_Vector v)
{
  NVContent* c = content(v);
  N_VGetArrayPointer_HipNative(v);
  CEDA_HIP_CHECK(hipMemcpyAsync(c->host, c->dev, sizeof(double) * (size_t)c->length,
                                hipMemcpyDeviceToHost, c->exec->stream));
  CEDA_HIP_CHECK(hipStreamSynchronize(c->exec->stream));
  return 0;
}

void* N_VGetStream_HipNative(N_Vector v) { return (void*)vstream(v); }

SUNErrCode N_VSeedRandom_HipNative(N_Vector v, uint64_t seed)
{
  const int64_t n = vlen(v);
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_seed, dim3(stream_grid(2 * n)), dim3(kBlock), 0, vstream(v), n, seed, dptr(v));
  CEDA_HIP_CHECK(hipGetLastError());
  return 0;
}

void N_VLinearSum_HipNative(sunrealtype a, N_Vector x, sunrealtype b, N_Vector y, N_Vector z)
{
  const double* xs[2] = {dptr(x), dptr(y)};
  launch_map<2>(z, xs, ELinSum{a, b});
}
void N_VConst_HipNative(sunrealtype c, N_Vector z) { launch_map<0>(z, nullptr, EConst{c}); }
void N_VProd_HipNative(N_Vector x, N_Vector y, N_Vector z)
{
  const double* xs[2] = {dptr(x), dptr(y)};
  launch_map<2>(z, xs, EProd{});
}
void N_VDiv_HipNative(N_Vector x, N_Vector y, N_Vector z)
{
  const double* xs[2] = {dptr(x), dptr(y)};
  launch_map<2>(z, xs, EDiv{});
}
void N_VScale_HipNative(sunrealtype c, N_Vector x, N_Vector z)
{
  const double* xs[1] = {dptr(x)};
  launch_map<1>(z, xs, EScale{c});
}
void N_VAbs_HipNative(N_Vector x, N_Vector z)
{
  const double* xs[1] = {dptr(x)};
  launch_map<1>(z, xs, EAbs{});
}
void N_VInv_HipNative(N_Vector x, N_Vector z)
{
  const double* xs[1] = {dptr(x)};
  launch_map<1>(z, xs, EInv{});
}
void N_VAddConst_HipNative(N_Vector x, sunrealtype b, N_Vector z)
{
  const double* xs[1] = {dptr(x)};
  launch_map<1>(z, xs, EAddConst{b});
}
void N_VCompare_HipNative(sunrealtype c, N_Vector x, N_Vector z)
{
  const double* xs[1] = {dptr(x)};
  launch_map<1>(z, xs, ECompare{c});
}

sunrealtype N_VDotProd_HipNative(N_Vector x, N_Vector y)
{
  return reduce1<RSUM>(x, RDot{dptr(x), dptr(y)});
}
sunrealtype N_VMaxNorm_HipNative(N_Vector x) { return reduce1<RMAX>(x, RMaxAbs{dptr(x)}); }
sunrealtype N_VWSqrSumLocal_HipNative(N_Vector x, N_Vector w)
{
  return reduce1<RSUM>(x, RWSq{dptr(x), dptr(w)});
}
sunrealtype N_VWrmsNorm_HipNative(N_Vector x, N_Vector w)
{
  return sqrt(N_VWSqrSumLocal_HipNative(x, w) / (double)vlen(x));
}
sunrealtype N_VWSqrSumMaskLocal_HipNative(N_Vector x, N_Vector w, N_Vector id)
{
  return reduce1<RSUM>(x, RWSqMask{dptr(x), dptr(w), dptr(id)});
}
sunrealtype N_VWrmsNormMask_HipNative(N_Vector x, N_Vector w, N_Vector id)
{
  return sqrt(N_VWSqrSumMaskLocal_HipNative(x, w, id) / (double)vlen(x));
}
sunrealtype N_VMin_HipNative(N_Vector x) { return reduce1<RMIN>(x, RMinV{dptr(x)}); }
sunrealtype N_VWL2Norm_HipNative(N_Vector x, N_Vector w)
{
  return sqrt(N_VWSqrSumLocal_HipNative(x, w));
}
sunrealtype N_VL1Norm_HipNative(N_Vector x) { return reduce1<RSUM>(x, RL1{dptr(x)}); }
sunbooleantype N_VInvTest_HipNative(N_Vector x, N_Vector z)
{
  return reduce1<RMAX>(x, RInvTest{dptr(x), dptr(z)}) > 0.0 ? SUNFALSE : SUNTRUE;
}
sunbooleantype N_VConstrMask_HipNative(N_Vector c, N_Vector x, N_Vector m)
{
  return reduce1<RMAX>(x, RConstrMask{dptr(c), dptr(x), dptr(m)}) > 0.0 ? SUNFALSE : SUNTRUE;
}
sunrealtype N_VMinQuotient_HipNative(N_Vector num, N_Vector denom)
{
  return reduce1<RMIN>(num, RMinQuot{dptr(num), dptr(denom)});
}

SUNErrCode N_VLinearCombination_HipNative(int nvec, sunrealtype* c, N_Vector* X, N_Vector z)
{
  if (nvec < 1 || nvec > NVHIP_MAX_FUSED) return -1;
  LCArgs a;
  a.nv = nvec;
  bool al = aligned16(dptr(z));
  for (int k = 0; k < nvec; k++) {
    a.c[k] = c[k];
    a.x[k] = dptr(X[k]);
    al = al && aligned16(a.x[k]);
  }
  a.z = dptr(z);
  const int64_t n = vlen(z);
  if (n == 0) return 0;
  hipStream_t s = vstream(z);
  switch (nvec) {
  case 1: launch_lc<1>(a, n, s, al); break;
  case 2: launch_lc<2>(a, n, s, al); break;
  case 3: launch_lc<3>(a, n, s, al); break;
  case 4: launch_lc<4>(a, n, s, al); break;
  case 5: launch_lc<5>(a, n, s, al); break;
  case 6: launch_lc<6>(a, n, s, al); break;
  default: launch_lc<0>(a, n, s, al); break;
  }
  return 0;
}

SUNErrCode N_VScaleAddMulti_HipNative(int nvec, sunrealtype* a, N_Vector x, N_Vector* Y, N_Vector* Z)
{
  if (nvec < 1 || nvec > NVHIP_MAX_FUSED) return -1;
  SAMArgs s;
  s.nv = nvec;
  s.x = dptr(x);
  bool al = aligned16(s.x);
  for (int k = 0; k < nvec; k++) {
    s.a[k] = a[k];
    s.y[k] = dptr(Y[k]);
    s.z[k] = dptr(Z[k]);
    al = al && aligned16(s.y[k]) && aligned16(s.z[k]);
  }
  const int64_t n = vlen(x);
  if (n == 0) return 0;
  SAMOp op{s};
  if (al) hipLaunchKernelGGL(k_map2<SAMOp>, dim3(stream_grid(n)), dim3(kBlock), 0, vstream(x), n, op);
  else hipLaunchKernelGGL(k_map1<SAMOp>, dim3(stream_grid(2 * n)), dim3(kBlock), 0, vstream(x), n, op);
  CEDA_HIP_CHECK(hipGetLastError());
  return 0;
}

SUNErrCode N_VDotProdMulti_HipNative(int nvec, N_Vector x, N_Vector* Y, sunrealtype* dotprods)
{
  if (nvec < 1 || nvec > NVHIP_MAX_FUSED) return -1;
  if (nvec == 1) dotmulti<1>(nvec, x, Y, dotprods);
  else if (nvec == 2) dotmulti<2>(nvec, x, Y, dotprods);
  else if (nvec <= 4) dotmulti<4>(nvec, x, Y, dotprods);
  else if (nvec <= 8) dotmulti<8>(nvec, x, Y, dotprods);
  else dotmulti_groups(nvec, x, Y, dotprods);
  return 0;
}

SUNErrCode N_VBufSize_HipNative(N_Vector x, sunindextype* size)
{
  *size = vlen(x) * (sunindextype)sizeof(double);
  return 0;
}
SUNErrCode N_VBufPack_HipNative(N_Vector x, void* buf)
{
  CEDA_HIP_CHECK(hipMemcpyAsync(buf, dptr(x), sizeof(double) * (size_t)vlen(x), hipMemcpyDeviceToHost, vstream(x)));
  CEDA_HIP_CHECK(hipStreamSynchronize(vstream(x)));
  return 0;
}
SUNErrCode N_VBufUnpack_HipNative(N_Vector x, void* buf)
{
  CEDA_HIP_CHECK(hipMemcpyAsync(dptr(x), buf, sizeof(double) * (size_t)vlen(x), hipMemcpyHostToDevice, vstream(x)));
  CEDA_HIP_CHECK(hipStreamSynchronize(vstream(x)));
  return 0;
}
void N_VPrintFile_HipNative(N_Vector x, FILE* outfile)
{
  N_VCopyFromDevice_HipNative(x);
  const double* h = content(x)->host;
  for (sunindextype i = 0; i < vlen(x); i++) fprintf(outfile, "%.16e\n", h[i]);
  fprintf(outfile, "\n");
}
void N_VPrint_HipNative(N_Vector x) { N_VPrintFile_HipNative(x, stdout); }

}  // extern "C"
