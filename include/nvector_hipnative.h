/* ---------------------------------------------------------------------------
 * Native HIP N_Vector for MI355X (gfx950).
 *
 * Replaces the serial/MPI vector the reference builds with
 * N_VNew_Parallel(comm, nodes_loc, nodes, ctx) (diffusion_2D/main.cpp:176) and
 * follows the ownership conventions of the reference's own ops-table filling
 * (gkeyll_diffusion/src/nvector_gkylzero.c:32-82 attach, :111-203 clone /
 * own_vector / destroy): Clone allocates and owns device memory, Make wraps
 * caller memory without owning it, Destroy frees only what it owns.
 *
 * All operations run on the vector's HIP stream in issue order.  Reductions
 * return a host scalar and therefore synchronise that stream; element-wise and
 * fused operations never do, so a sequence of them can be captured in a HIP
 * graph.  Reductions are deterministic (fixed grid, fixed combine order).
 * N_VGetArrayPointer returns a pinned host mirror that is synchronised only by
 * N_VCopyToDevice_HipNative / N_VCopyFromDevice_HipNative.
 * ------------------------------------------------------------------------- */
#ifndef NVECTOR_HIPNATIVE_H
#define NVECTOR_HIPNATIVE_H

#include "ceda_sundials.h"

#ifdef __cplusplus
extern "C" {
#endif

/* replaces N_VNew_Parallel (main.cpp:176): owned device vector of `length` */
N_Vector N_VNew_HipNative(sunindextype length, SUNContext sunctx);
/* empty vector with the ops table attached (nvector_gkylzero.c:32) */
N_Vector N_VNewEmpty_HipNative(sunindextype length, SUNContext sunctx);
/* wraps caller device memory without owning it (nvector_gkylzero.c:93) */
N_Vector N_VMake_HipNative(sunindextype length, sunrealtype* d_data, SUNContext sunctx);
/* host mirror <-> device */
SUNErrCode N_VCopyToDevice_HipNative(N_Vector v);
SUNErrCode N_VCopyFromDevice_HipNative(N_Vector v);
/* the HIP stream (hipStream_t) all operations of v are issued on */
void* N_VGetStream_HipNative(N_Vector v);
/* device-side generator of the SplitMix64 start vector: v_i = (splitmix64(seed+i)>>11)*2^-53 - 0.5 */
SUNErrCode N_VSeedRandom_HipNative(N_Vector v, uint64_t seed);

/* the implementations behind the ops table */
N_Vector_ID N_VGetVectorID_HipNative(N_Vector v);
N_Vector N_VClone_HipNative(N_Vector w);
N_Vector N_VCloneEmpty_HipNative(N_Vector w);
void N_VDestroy_HipNative(N_Vector v);
void N_VSpace_HipNative(N_Vector v, sunindextype* lrw, sunindextype* liw);
sunrealtype* N_VGetArrayPointer_HipNative(N_Vector v);
sunrealtype* N_VGetDeviceArrayPointer_HipNative(N_Vector v);
void N_VSetArrayPointer_HipNative(sunrealtype* d_data, N_Vector v);
sunindextype N_VGetLength_HipNative(N_Vector v);
void N_VLinearSum_HipNative(sunrealtype a, N_Vector x, sunrealtype b, N_Vector y, N_Vector z);
void N_VConst_HipNative(sunrealtype c, N_Vector z);
void N_VProd_HipNative(N_Vector x, N_Vector y, N_Vector z);
void N_VDiv_HipNative(N_Vector x, N_Vector y, N_Vector z);
void N_VScale_HipNative(sunrealtype c, N_Vector x, N_Vector z);
void N_VAbs_HipNative(N_Vector x, N_Vector z);
void N_VInv_HipNative(N_Vector x, N_Vector z);
void N_VAddConst_HipNative(N_Vector x, sunrealtype b, N_Vector z);
sunrealtype N_VDotProd_HipNative(N_Vector x, N_Vector y);
sunrealtype N_VMaxNorm_HipNative(N_Vector x);
sunrealtype N_VWrmsNorm_HipNative(N_Vector x, N_Vector w);
sunrealtype N_VWrmsNormMask_HipNative(N_Vector x, N_Vector w, N_Vector id);
sunrealtype N_VMin_HipNative(N_Vector x);
sunrealtype N_VWL2Norm_HipNative(N_Vector x, N_Vector w);
sunrealtype N_VL1Norm_HipNative(N_Vector x);
void N_VCompare_HipNative(sunrealtype c, N_Vector x, N_Vector z);
sunbooleantype N_VInvTest_HipNative(N_Vector x, N_Vector z);
sunbooleantype N_VConstrMask_HipNative(N_Vector c, N_Vector x, N_Vector m);
sunrealtype N_VMinQuotient_HipNative(N_Vector num, N_Vector denom);
SUNErrCode N_VLinearCombination_HipNative(int nvec, sunrealtype* c, N_Vector* X, N_Vector z);
SUNErrCode N_VScaleAddMulti_HipNative(int nvec, sunrealtype* a, N_Vector x,
                                      N_Vector* Y, N_Vector* Z);
SUNErrCode N_VDotProdMulti_HipNative(int nvec, N_Vector x, N_Vector* Y,
                                     sunrealtype* dotprods);
sunrealtype N_VWSqrSumLocal_HipNative(N_Vector x, N_Vector w);
sunrealtype N_VWSqrSumMaskLocal_HipNative(N_Vector x, N_Vector w, N_Vector id);
SUNErrCode N_VBufSize_HipNative(N_Vector x, sunindextype* size);
SUNErrCode N_VBufPack_HipNative(N_Vector x, void* buf);
SUNErrCode N_VBufUnpack_HipNative(N_Vector x, void* buf);
void N_VPrint_HipNative(N_Vector x);
void N_VPrintFile_HipNative(N_Vector x, FILE* outfile);

/* the largest vector count a single fused kernel accepts */
#define NVHIP_MAX_FUSED 64

#ifdef __cplusplus
}
#endif
#endif
