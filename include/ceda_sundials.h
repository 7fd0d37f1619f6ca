/* ---------------------------------------------------------------------------
 * SUNDIALS public ABI subset used by the diffusion_2D hot path.
 *
 * The reference drives its problem layer through SUNDIALS (an un-vendored
 * submodule, .gitmodules:1-3): the N_Vector object and ops table
 * (nvector_parallel, main.cpp:176), ARKODE's user callbacks (ARKRhsFn
 * diffusion_2D.hpp:250, ARKDomEigFn main.cpp:57-59, ARKLsPrecSetupFn /
 * ARKLsPrecSolveFn diffusion_2D.hpp:253-257) and the integrator API of
 * main.cpp:176-400.  This header restates that interface — the struct layout
 * of `struct _generic_N_Vector` and `struct _generic_N_Vector_Ops` follows the
 * SUNDIALS 7.x public header field for field, so a vector built by this
 * library can be handed to a SUNDIALS build of that series — with types fixed
 * to the configuration the path needs (sunrealtype = double, sunindextype =
 * int64_t, no MPI communicator).
 * ------------------------------------------------------------------------- */
#ifndef CEDA_SUNDIALS_H
#define CEDA_SUNDIALS_H

#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double sunrealtype;
typedef int64_t sunindextype;
typedef int sunbooleantype;
typedef int SUNErrCode;
typedef int SUNComm;
#define SUNFALSE 0
#define SUNTRUE 1
#define SUN_SUCCESS 0
#define SUN_COMM_NULL 0
#define SUN_RCONST(x) (x)

typedef struct SUNContext_* SUNContext;
typedef struct _generic_N_Vector* N_Vector;
typedef struct _generic_N_Vector_Ops* N_Vector_Ops;
typedef N_Vector* N_Vector_S;

typedef enum {
  SUNDIALS_NVEC_SERIAL, SUNDIALS_NVEC_PARALLEL, SUNDIALS_NVEC_OPENMP,
  SUNDIALS_NVEC_PTHREADS, SUNDIALS_NVEC_PARHYP, SUNDIALS_NVEC_PETSC,
  SUNDIALS_NVEC_CUDA, SUNDIALS_NVEC_HIP, SUNDIALS_NVEC_SYCL, SUNDIALS_NVEC_RAJA,
  SUNDIALS_NVEC_KOKKOS, SUNDIALS_NVEC_OPENMPDEV, SUNDIALS_NVEC_TRILINOS,
  SUNDIALS_NVEC_MANYVECTOR, SUNDIALS_NVEC_MPIMANYVECTOR, SUNDIALS_NVEC_MPIPLUSX,
  SUNDIALS_NVEC_CUSTOM
} N_Vector_ID;

struct _generic_N_Vector_Ops {
  /* constructors, destructors, utility */
  N_Vector_ID (*nvgetvectorid)(N_Vector);
  N_Vector (*nvclone)(N_Vector);
  N_Vector (*nvcloneempty)(N_Vector);
  void (*nvdestroy)(N_Vector);
  void (*nvspace)(N_Vector, sunindextype*, sunindextype*);
  sunrealtype* (*nvgetarraypointer)(N_Vector);
  sunrealtype* (*nvgetdevicearraypointer)(N_Vector);
  void (*nvsetarraypointer)(sunrealtype*, N_Vector);
  SUNComm (*nvgetcommunicator)(N_Vector);
  sunindextype (*nvgetlength)(N_Vector);
  sunindextype (*nvgetlocallength)(N_Vector);
  /* standard vector operations */
  void (*nvlinearsum)(sunrealtype, N_Vector, sunrealtype, N_Vector, N_Vector);
  void (*nvconst)(sunrealtype, N_Vector);
  void (*nvprod)(N_Vector, N_Vector, N_Vector);
  void (*nvdiv)(N_Vector, N_Vector, N_Vector);
  void (*nvscale)(sunrealtype, N_Vector, N_Vector);
  void (*nvabs)(N_Vector, N_Vector);
  void (*nvinv)(N_Vector, N_Vector);
  void (*nvaddconst)(N_Vector, sunrealtype, N_Vector);
  sunrealtype (*nvdotprod)(N_Vector, N_Vector);
  sunrealtype (*nvmaxnorm)(N_Vector);
  sunrealtype (*nvwrmsnorm)(N_Vector, N_Vector);
  sunrealtype (*nvwrmsnormmask)(N_Vector, N_Vector, N_Vector);
  sunrealtype (*nvmin)(N_Vector);
  sunrealtype (*nvwl2norm)(N_Vector, N_Vector);
  sunrealtype (*nvl1norm)(N_Vector);
  void (*nvcompare)(sunrealtype, N_Vector, N_Vector);
  sunbooleantype (*nvinvtest)(N_Vector, N_Vector);
  sunbooleantype (*nvconstrmask)(N_Vector, N_Vector, N_Vector);
  sunrealtype (*nvminquotient)(N_Vector, N_Vector);
  /* fused vector operations */
  SUNErrCode (*nvlinearcombination)(int, sunrealtype*, N_Vector*, N_Vector);
  SUNErrCode (*nvscaleaddmulti)(int, sunrealtype*, N_Vector, N_Vector*, N_Vector*);
  SUNErrCode (*nvdotprodmulti)(int, N_Vector, N_Vector*, sunrealtype*);
  /* vector array operations */
  SUNErrCode (*nvlinearsumvectorarray)(int, sunrealtype, N_Vector*, sunrealtype,
                                       N_Vector*, N_Vector*);
  SUNErrCode (*nvscalevectorarray)(int, sunrealtype*, N_Vector*, N_Vector*);
  SUNErrCode (*nvconstvectorarray)(int, sunrealtype, N_Vector*);
  SUNErrCode (*nvwrmsnormvectorarray)(int, N_Vector*, N_Vector*, sunrealtype*);
  SUNErrCode (*nvwrmsnormmaskvectorarray)(int, N_Vector*, N_Vector*, N_Vector,
                                          sunrealtype*);
  SUNErrCode (*nvscaleaddmultivectorarray)(int, int, sunrealtype*, N_Vector*,
                                           N_Vector**, N_Vector**);
  SUNErrCode (*nvlinearcombinationvectorarray)(int, int, sunrealtype*,
                                               N_Vector**, N_Vector*);
  /* local reductions */
  sunrealtype (*nvdotprodlocal)(N_Vector, N_Vector);
  sunrealtype (*nvmaxnormlocal)(N_Vector);
  sunrealtype (*nvminlocal)(N_Vector);
  sunrealtype (*nvl1normlocal)(N_Vector);
  sunbooleantype (*nvinvtestlocal)(N_Vector, N_Vector);
  sunbooleantype (*nvconstrmasklocal)(N_Vector, N_Vector, N_Vector);
  sunrealtype (*nvminquotientlocal)(N_Vector, N_Vector);
  sunrealtype (*nvwsqrsumlocal)(N_Vector, N_Vector);
  sunrealtype (*nvwsqrsummasklocal)(N_Vector, N_Vector, N_Vector);
  /* single-buffer reductions */
  SUNErrCode (*nvdotprodmultilocal)(int, N_Vector, N_Vector*, sunrealtype*);
  SUNErrCode (*nvdotprodmultiallreduce)(int, N_Vector, sunrealtype*);
  /* XBraid interface */
  SUNErrCode (*nvbufsize)(N_Vector, sunindextype*);
  SUNErrCode (*nvbufpack)(N_Vector, void*);
  SUNErrCode (*nvbufunpack)(N_Vector, void*);
  /* debugging */
  void (*nvprint)(N_Vector);
  void (*nvprintfile)(N_Vector, FILE*);
};

struct _generic_N_Vector {
  void* content;
  N_Vector_Ops ops;
  SUNContext sunctx;
};

/* ---- context (SUNContext_Create / SUNContext_Free) ---- */
SUNErrCode SUNContext_Create(SUNComm comm, SUNContext* ctx);
SUNErrCode SUNContext_Free(SUNContext* ctx);

/* ---- generic N_Vector layer (dispatch through the ops table) ---- */
N_Vector N_VNewEmpty(SUNContext sunctx);
void N_VFreeEmpty(N_Vector v);
SUNErrCode N_VCopyOps(N_Vector w, N_Vector v);
N_Vector_ID N_VGetVectorID(N_Vector w);
N_Vector N_VClone(N_Vector w);
N_Vector N_VCloneEmpty(N_Vector w);
void N_VDestroy(N_Vector v);
void N_VSpace(N_Vector v, sunindextype* lrw, sunindextype* liw);
sunrealtype* N_VGetArrayPointer(N_Vector v);
sunrealtype* N_VGetDeviceArrayPointer(N_Vector v);
void N_VSetArrayPointer(sunrealtype* v_data, N_Vector v);
sunindextype N_VGetLength(N_Vector v);
sunindextype N_VGetLocalLength(N_Vector v);
void N_VLinearSum(sunrealtype a, N_Vector x, sunrealtype b, N_Vector y, N_Vector z);
void N_VConst(sunrealtype c, N_Vector z);
void N_VProd(N_Vector x, N_Vector y, N_Vector z);
void N_VDiv(N_Vector x, N_Vector y, N_Vector z);
void N_VScale(sunrealtype c, N_Vector x, N_Vector z);
void N_VAbs(N_Vector x, N_Vector z);
void N_VInv(N_Vector x, N_Vector z);
void N_VAddConst(N_Vector x, sunrealtype b, N_Vector z);
sunrealtype N_VDotProd(N_Vector x, N_Vector y);
sunrealtype N_VMaxNorm(N_Vector x);
sunrealtype N_VWrmsNorm(N_Vector x, N_Vector w);
sunrealtype N_VWrmsNormMask(N_Vector x, N_Vector w, N_Vector id);
sunrealtype N_VMin(N_Vector x);
sunrealtype N_VWL2Norm(N_Vector x, N_Vector w);
sunrealtype N_VL1Norm(N_Vector x);
void N_VCompare(sunrealtype c, N_Vector x, N_Vector z);
sunbooleantype N_VInvTest(N_Vector x, N_Vector z);
sunbooleantype N_VConstrMask(N_Vector c, N_Vector x, N_Vector m);
sunrealtype N_VMinQuotient(N_Vector num, N_Vector denom);
SUNErrCode N_VLinearCombination(int nvec, sunrealtype* c, N_Vector* X, N_Vector z);
SUNErrCode N_VScaleAddMulti(int nvec, sunrealtype* a, N_Vector x, N_Vector* Y,
                            N_Vector* Z);
SUNErrCode N_VDotProdMulti(int nvec, N_Vector x, N_Vector* Y, sunrealtype* dotprods);
SUNErrCode N_VLinearSumVectorArray(int nvec, sunrealtype a, N_Vector* X,
                                   sunrealtype b, N_Vector* Y, N_Vector* Z);
SUNErrCode N_VScaleVectorArray(int nvec, sunrealtype* c, N_Vector* X, N_Vector* Z);
SUNErrCode N_VConstVectorArray(int nvec, sunrealtype c, N_Vector* Z);
SUNErrCode N_VWrmsNormVectorArray(int nvec, N_Vector* X, N_Vector* W,
                                  sunrealtype* nrm);
N_Vector* N_VCloneVectorArray(int count, N_Vector w);
void N_VDestroyVectorArray(N_Vector* vs, int count);

/* ---- ARKODE user callbacks ---- */
typedef int (*ARKRhsFn)(sunrealtype t, N_Vector y, N_Vector ydot, void* user_data);
typedef int (*ARKDomEigFn)(sunrealtype t, N_Vector y, N_Vector fn,
                           sunrealtype* lambdaR, sunrealtype* lambdaI,
                           void* user_data, N_Vector temp1, N_Vector temp2,
                           N_Vector temp3);
typedef int (*ARKLsPrecSetupFn)(sunrealtype t, N_Vector y, N_Vector fy,
                                sunbooleantype jok, sunbooleantype* jcurPtr,
                                sunrealtype gamma, void* user_data);
typedef int (*ARKLsPrecSolveFn)(sunrealtype t, N_Vector y, N_Vector fy,
                                N_Vector r, N_Vector z, sunrealtype gamma,
                                sunrealtype delta, int lr, void* user_data);
typedef int (*ARKLsJacTimesSetupFn)(sunrealtype t, N_Vector y, N_Vector fy, void* user_data);
typedef int (*ARKLsJacTimesVecFn)(N_Vector v, N_Vector Jv, sunrealtype t, N_Vector y,
                                  N_Vector fy, void* user_data, N_Vector tmp);

#ifdef __cplusplus
}
#endif
#endif
