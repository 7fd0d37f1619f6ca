// SUNContext and the generic N_Vector layer (dispatch through the ops table).
//
// Restates the public SUNDIALS interface the reference calls (N_VClone,
// N_VScale, N_VConst, N_VLinearSum, N_VDotProd, N_VMaxNorm, N_VProd and
// N_VGetArrayPointer at diffusion_2D/main.cpp:188-193,463-464,
// diffusion.cpp:24-31, preconditioner_jacobi.cpp:58, diffusion_2D.cpp:801).
// The context owns the HIP stream every vector created from it uses and the
// reduction workspace (device partials + pinned host result).
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>

#include "nvhip_internal.hpp"

struct SUNContext_ {
  ceda::Exec exec;
};

namespace ceda {
Exec* context_exec(SUNContext ctx) { return &ctx->exec; }
std::recursive_mutex& device_sync_mutex()
{
  static std::recursive_mutex mu;
  return mu;
}
}  // namespace ceda

extern "C" {

SUNErrCode SUNContext_Create(SUNComm, SUNContext* ctx)
{
  if (!ctx) return -1;
  SUNContext c = new SUNContext_();
  {
    std::lock_guard<std::recursive_mutex> g(ceda::device_sync_mutex());
    CEDA_HIP_CHECK(hipStreamCreateWithFlags(&c->exec.stream, hipStreamNonBlocking));
  }
  CEDA_HIP_CHECK(ceda::dev_malloc(&c->exec.d_partials,
                           sizeof(double) * ceda::kMaxReduceBlocks * ceda::kMaxAcc));
  CEDA_HIP_CHECK(ceda::dev_malloc(&c->exec.d_result, sizeof(double) * ceda::kMaxAcc));
  CEDA_HIP_CHECK(ceda::host_malloc(&c->exec.h_result, sizeof(double) * ceda::kMaxAcc,
                               hipHostMallocDefault));
  int dev = 0;
  CEDA_HIP_CHECK(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  CEDA_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
  c->exec.num_cu = prop.multiProcessorCount;
  *ctx = c;
  return 0;
}

SUNErrCode SUNContext_Free(SUNContext* ctx)
{
  if (!ctx || !*ctx) return 0;
  SUNContext c = *ctx;
  CEDA_HIP_CHECK(hipStreamSynchronize(c->exec.stream));
  CEDA_HIP_CHECK(ceda::dev_free(c->exec.d_partials));
  CEDA_HIP_CHECK(ceda::dev_free(c->exec.d_result));
  CEDA_HIP_CHECK(ceda::host_free(c->exec.h_result));
  {
    std::lock_guard<std::recursive_mutex> g(ceda::device_sync_mutex());
    CEDA_HIP_CHECK(hipStreamDestroy(c->exec.stream));
  }
  delete c;
  *ctx = nullptr;
  return 0;
}

// the HIP stream of a context (all its vectors' work is issued there)
void* SUNContext_GetHipStream(SUNContext ctx) { return (void*)ctx->exec.stream; }

N_Vector N_VNewEmpty(SUNContext sunctx)
{
  N_Vector v = (N_Vector)calloc(1, sizeof(*v));
  if (!v) return nullptr;
  v->ops = (N_Vector_Ops)calloc(1, sizeof(*v->ops));
  if (!v->ops) {
    free(v);
    return nullptr;
  }
  v->sunctx = sunctx;
  return v;
}

void N_VFreeEmpty(N_Vector v)
{
  if (!v) return;
  free(v->ops);
  free(v);
}

SUNErrCode N_VCopyOps(N_Vector w, N_Vector v)
{
  if (!w || !v || !w->ops || !v->ops) return -1;
  memcpy(v->ops, w->ops, sizeof(*w->ops));
  return 0;
}

N_Vector_ID N_VGetVectorID(N_Vector w) { return w->ops->nvgetvectorid(w); }
N_Vector N_VClone(N_Vector w) { return w->ops->nvclone(w); }
N_Vector N_VCloneEmpty(N_Vector w) { return w->ops->nvcloneempty(w); }
void N_VDestroy(N_Vector v)
{
  if (!v) return;
  if (v->ops && v->ops->nvdestroy) v->ops->nvdestroy(v);
  else N_VFreeEmpty(v);
}
void N_VSpace(N_Vector v, sunindextype* lrw, sunindextype* liw) { v->ops->nvspace(v, lrw, liw); }
sunrealtype* N_VGetArrayPointer(N_Vector v) { return v->ops->nvgetarraypointer(v); }
sunrealtype* N_VGetDeviceArrayPointer(N_Vector v)
{
  return v->ops->nvgetdevicearraypointer ? v->ops->nvgetdevicearraypointer(v) : nullptr;
}
void N_VSetArrayPointer(sunrealtype* d, N_Vector v) { v->ops->nvsetarraypointer(d, v); }
sunindextype N_VGetLength(N_Vector v) { return v->ops->nvgetlength(v); }
sunindextype N_VGetLocalLength(N_Vector v) { return v->ops->nvgetlocallength(v); }
void N_VLinearSum(sunrealtype a, N_Vector x, sunrealtype b, N_Vector y, N_Vector z)
{
  z->ops->nvlinearsum(a, x, b, y, z);
}
void N_VConst(sunrealtype c, N_Vector z) { z->ops->nvconst(c, z); }
void N_VProd(N_Vector x, N_Vector y, N_Vector z) { z->ops->nvprod(x, y, z); }
void N_VDiv(N_Vector x, N_Vector y, N_Vector z) { z->ops->nvdiv(x, y, z); }
void N_VScale(sunrealtype c, N_Vector x, N_Vector z) { z->ops->nvscale(c, x, z); }
void N_VAbs(N_Vector x, N_Vector z) { z->ops->nvabs(x, z); }
void N_VInv(N_Vector x, N_Vector z) { z->ops->nvinv(x, z); }
void N_VAddConst(N_Vector x, sunrealtype b, N_Vector z) { z->ops->nvaddconst(x, b, z); }
sunrealtype N_VDotProd(N_Vector x, N_Vector y) { return y->ops->nvdotprod(x, y); }
sunrealtype N_VMaxNorm(N_Vector x) { return x->ops->nvmaxnorm(x); }
sunrealtype N_VWrmsNorm(N_Vector x, N_Vector w) { return x->ops->nvwrmsnorm(x, w); }
sunrealtype N_VWrmsNormMask(N_Vector x, N_Vector w, N_Vector id)
{
  return x->ops->nvwrmsnormmask(x, w, id);
}
sunrealtype N_VMin(N_Vector x) { return x->ops->nvmin(x); }
sunrealtype N_VWL2Norm(N_Vector x, N_Vector w) { return x->ops->nvwl2norm(x, w); }
sunrealtype N_VL1Norm(N_Vector x) { return x->ops->nvl1norm(x); }
void N_VCompare(sunrealtype c, N_Vector x, N_Vector z) { z->ops->nvcompare(c, x, z); }
sunbooleantype N_VInvTest(N_Vector x, N_Vector z) { return z->ops->nvinvtest(x, z); }
sunbooleantype N_VConstrMask(N_Vector c, N_Vector x, N_Vector m)
{
  return x->ops->nvconstrmask(c, x, m);
}
sunrealtype N_VMinQuotient(N_Vector num, N_Vector denom)
{
  return num->ops->nvminquotient(num, denom);
}

SUNErrCode N_VLinearCombination(int nvec, sunrealtype* c, N_Vector* X, N_Vector z)
{
  if (z->ops->nvlinearcombination) return z->ops->nvlinearcombination(nvec, c, X, z);
  z->ops->nvscale(c[0], X[0], z);
  for (int i = 1; i < nvec; i++) z->ops->nvlinearsum(c[i], X[i], 1.0, z, z);
  return 0;
}

SUNErrCode N_VScaleAddMulti(int nvec, sunrealtype* a, N_Vector x, N_Vector* Y, N_Vector* Z)
{
  if (x->ops->nvscaleaddmulti) return x->ops->nvscaleaddmulti(nvec, a, x, Y, Z);
  for (int i = 0; i < nvec; i++) x->ops->nvlinearsum(a[i], x, 1.0, Y[i], Z[i]);
  return 0;
}

SUNErrCode N_VDotProdMulti(int nvec, N_Vector x, N_Vector* Y, sunrealtype* d)
{
  if (x->ops->nvdotprodmulti) return x->ops->nvdotprodmulti(nvec, x, Y, d);
  for (int i = 0; i < nvec; i++) d[i] = x->ops->nvdotprod(x, Y[i]);
  return 0;
}

SUNErrCode N_VLinearSumVectorArray(int nvec, sunrealtype a, N_Vector* X, sunrealtype b,
                                   N_Vector* Y, N_Vector* Z)
{
  for (int i = 0; i < nvec; i++) N_VLinearSum(a, X[i], b, Y[i], Z[i]);
  return 0;
}

SUNErrCode N_VScaleVectorArray(int nvec, sunrealtype* c, N_Vector* X, N_Vector* Z)
{
  for (int i = 0; i < nvec; i++) N_VScale(c[i], X[i], Z[i]);
  return 0;
}

SUNErrCode N_VConstVectorArray(int nvec, sunrealtype c, N_Vector* Z)
{
  for (int i = 0; i < nvec; i++) N_VConst(c, Z[i]);
  return 0;
}

SUNErrCode N_VWrmsNormVectorArray(int nvec, N_Vector* X, N_Vector* W, sunrealtype* nrm)
{
  for (int i = 0; i < nvec; i++) nrm[i] = N_VWrmsNorm(X[i], W[i]);
  return 0;
}

N_Vector* N_VCloneVectorArray(int count, N_Vector w)
{
  if (count <= 0) return nullptr;
  N_Vector* vs = (N_Vector*)calloc((size_t)count, sizeof(N_Vector));
  for (int j = 0; j < count; j++) {
    vs[j] = N_VClone(w);
    if (!vs[j]) {
      N_VDestroyVectorArray(vs, j);
      return nullptr;
    }
  }
  return vs;
}

void N_VDestroyVectorArray(N_Vector* vs, int count)
{
  if (!vs) return;
  for (int j = 0; j < count; j++) N_VDestroy(vs[j]);
  free(vs);
}

}  // extern "C"
