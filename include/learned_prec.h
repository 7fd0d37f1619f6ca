/* ---------------------------------------------------------------------------
 * [ext] Learned preconditioner for the implicit path: a small convolutional
 * network (3x3 convolutions, 16 hidden channels, BF16 weights and
 * activations, FP32 accumulation on MFMA) that maps a residual r to a
 * correction z ~ (I - gamma J)^{-1} r inside FGMRES.  Two forms: the ReLU
 * network (LearnedPrec_Create) and the symmetric network K^T K
 * (LearnedPrec_CreateSymmetric), a fixed SPD operator that also serves PCG.
 *
 * There is no reference counterpart.  It plugs into the preconditioner slot
 * of the reference's implicit path (ARKodeSetPreconditioner, main.cpp:
 * 307-313; the Jacobi pair of preconditioner_jacobi.cpp:9-62 has the same
 * signatures).  The ReLU network is not a fixed linear operator and needs a
 * flexible Krylov method (SUNLinSol_SPFGMR).  The FP64 Krylov residual
 * decides convergence; when FGMRES does not converge within its iteration
 * cap the system is re-solved by Jacobi-PCG (ARKodeSetLSFallbackPCG).
 *
 * ReLU network.  nlayer layers (2..8): layer 0 maps 1 -> 16 channels, the
 * hidden layers 16 -> 16, the last 16 -> 1; layer l has dilation dil[l].  Zero
 * padding at the grid edge (the Dirichlet ghosts).  Applied to an image r:
 *   s = RMS(r);  x_0 = bf16(r / s);  x_{l+1} = bf16(relu(W_l * x_l + b_l));
 *   N(r) = s * out_scale * (W_L * x_L + b_L)      (FP64 output)
 * Parameters (FP32, NumParams of them): for each layer W[cout][cin][3][3]
 * (row-major taps, dy outer) then b[cout].
 *
 * Symmetric network (LearnedPrec_CreateSymmetric).  nlayer layers K_l (1..7):
 * layer 0 maps 1 -> 16 channels, the others 16 -> 16, no activation function
 * and no biases (their slots stay in the layout and at zero); the network is
 * K followed by its adjoint:
 *   h_0 = bf16(r / s);  h_{l+1} = bf16(W_l * h_l);  u_nl = h_nl;
 *   u_l = bf16(W_l^T * u_{l+1});  N(r) = s * out_scale * (W_0^T * u_1)
 * where W^T * is the adjoint convolution (taps flipped, channels exchanged).
 * N = K^T K is linear, symmetric and positive semi-definite up to the BF16
 * roundings, which keeps every eigenvalue of the preconditioned operator
 * positive (a ReLU network can map some residuals to near-zero or opposing
 * corrections, which stalls the Krylov method).  With 4 layers of dilations
 * 1, 3, 9, 27 its receptive field is 161 x 161 nodes (each layer is applied
 * twice: 8 convolutions per application).
 * ------------------------------------------------------------------------- */
#ifndef LEARNED_PREC_H
#define LEARNED_PREC_H

#include "arkode_hip.h"
#include "diffusion_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct LearnedPrec LearnedPrec;

LearnedPrec* LearnedPrec_Create(int nlayer, const int* dil, SUNContext ctx);
/* the symmetric network N = K^T K of nlayer linear layers (see above) */
LearnedPrec* LearnedPrec_CreateSymmetric(int nlayer, const int* dil, SUNContext ctx);
/* 1 for a symmetric network, 0 otherwise, -1 for NULL */
int LearnedPrec_IsSymmetric(const LearnedPrec* lp);
void LearnedPrec_Destroy(LearnedPrec* lp);
int LearnedPrec_NumParams(const LearnedPrec* lp);
/* host FP32 parameters in the layout above */
int LearnedPrec_SetParams(LearnedPrec* lp, const float* params);
int LearnedPrec_GetParams(const LearnedPrec* lp, float* params);
/* deterministic start: the identity map through the ReLUs (channel 0 carries
   relu(x), channel 1 relu(-x), the last layer takes their difference) plus
   uniform noise of the given amplitude (seeded) on every weight; for the
   symmetric network every layer passes channel 0 through (K^T K = I) */
int LearnedPrec_InitIdentity(LearnedPrec* lp, double noise, unsigned long seed);
/* multiplier of the network output (1/diag of the operator it is trained for) */
int LearnedPrec_SetOutputScale(LearnedPrec* lp, double out_scale);
double LearnedPrec_GetOutputScale(const LearnedPrec* lp);

/* the convolution kernels (every network in the process): 1 (default)
   sweeps rows of one residue class mod the dilation with the input rows
   staged in LDS (dilations up to 64; the 16-input-channel layers and the
   first layer), 2 does the same and, in the symmetric network's
   application, runs two consecutive layers per kernel with the intermediate
   rows in LDS where one dilation divides the other (both at most 32), 0 the
   direct kernel that reads every tap through the caches.  All give the same
   results bit for bit; -1 for another value. */
int LearnedPrec_SetConvKernel(int kind);
/* tuning of the two-layer kernel: chain rows per workgroup (default 16;
   1..1024) and input rows in flight per workgroup (default 2; 1..3); -1 for
   other values.  The results do not depend on them. */
int LearnedPrec_SetPairTuning(int rows, int ahead);
/* the row-chain kernels give each workgroup 16 consecutive rows of a chain,
   or 8, 4 or 2 on images where 16 would launch fewer than min_tasks
   workgroups (default 1024; 1 keeps 16 everywhere; 1..2^20, -1 otherwise).
   The results do not depend on it. */
int LearnedPrec_SetChainTasks(int min_tasks);

/* out = N(in) for nb images of nx*ny pixels (image-major, x fastest), FP64
   device arrays; addto = 1 adds N(in) to out */
int LearnedPrec_Apply(LearnedPrec* lp, int nb, int nx, int ny, const double* in, double* out,
                      int addto);

/* LearnedPrec_Apply without any allocation, for callers that cannot
   allocate device memory (code under HIP graph capture, serving loops,
   harnesses that forbid hipMalloc); it has no counterpart in the reference.
   The workspace (two activation buffers, per-image scales and partial sums)
   is a fixed device array of 65 MiB per GPU, held until the process ends:
   at most 64 images and 2^20 pixels in all, rows padded to multiples of 16
   (one 1024 x 1024 image); larger inputs return -2 with a message and launch
   nothing.  Same kernels and operands as LearnedPrec_Apply, so the same
   result bit for bit.  All calls on one GPU share the array: calls that
   overlap on one GPU (other streams or host threads) race. */
int LearnedPrec_ApplyNoAlloc(LearnedPrec* lp, int nb, int nx, int ny, const double* in, double* out,
                             int addto);

/* training objective on a batch of pairs (r_b, e_b) with (I - ratio*L) e = r
   on nb Dirichlet grids of nx*ny nodes (L the unit-spacing 5-point Laplacian,
   zero boundary), in units normalised by s_b = RMS(r_b):
     loss = mean_b  d_b^T A d_b / (e_b^T r_b),  d_b = N(r_b) - e_b
   (the squared energy-norm error of the correction, relative to that of e);
   writes the loss and, if grad is not NULL, its gradient with respect to the
   parameters (host FP32, NumParams values).  r, e: FP64 device arrays. */
int LearnedPrec_LossGrad(LearnedPrec* lp, int nb, int nx, int ny, double ratio, const double* r,
                         const double* e, double* loss, float* grad);
/* the same objective for the anisotropic operator A = I - ratio_x Dxx -
   ratio_y Dyy (Dxx, Dyy the unit-spacing second differences along x and y);
   ratio_x == ratio_y is LearnedPrec_LossGrad bit for bit.  Training a
   network per anisotropy class: a problem with diffusion coefficients kx, ky
   has ratio_x / ratio_y = (kx/dx^2) / (ky/dy^2). */
int LearnedPrec_LossGradXY(LearnedPrec* lp, int nb, int nx, int ny, double ratio_x, double ratio_y,
                           const double* r, const double* e, double* loss, float* grad);

/* training pairs on the GPU: nsamp smooth random fields r_i (sum of sine
   modes with amplitude ~ N(0,1)/|k| up to a random cut-off, seed + i) on n*n
   Dirichlet grids, e_i from Jacobi-PCG on (I - ratio*L) e = r to relative
   residual reltol (all pairs solved together, per-pair step lengths); r, e:
   device arrays of nsamp*n*n doubles.  Writes the largest iteration count. */
int LearnedPrec_MakeTrainingSet(int nsamp, int n, double ratio, double reltol,
                                unsigned long seed, double* r, double* e, int* max_iters,
                                SUNContext ctx);
/* the same pairs for (I - ratio_x Dxx - ratio_y Dyy) e = r (Jacobi diagonal
   1 + 2 ratio_x + 2 ratio_y); ratio_x == ratio_y is LearnedPrec_MakeTrainingSet
   bit for bit */
int LearnedPrec_MakeTrainingSetXY(int nsamp, int n, double ratio_x, double ratio_y, double reltol,
                                  unsigned long seed, double* r, double* e, int* max_iters,
                                  SUNContext ctx);

/* Adam (beta 0.9/0.999) on minibatches of `batch` pairs drawn (seeded) from
   the training set, learning rate lr0 with cosine decay over `steps`;
   loss_hist (host, may be NULL) receives every step's loss.  All on the GPU
   with the forward and backward kernels of this file. */
int LearnedPrec_Train(LearnedPrec* lp, int nsamp, int n, double ratio, const double* r,
                      const double* e, int steps, int batch, double lr0, unsigned long seed,
                      double* loss_hist);
/* LearnedPrec_Train with the objective of LearnedPrec_LossGradXY */
int LearnedPrec_TrainXY(LearnedPrec* lp, int nsamp, int n, double ratio_x, double ratio_y,
                        const double* r, const double* e, int steps, int batch, double lr0,
                        unsigned long seed, double* loss_hist);

/* attach a trained network to a problem (NULL detaches); the problem does
   not own it.  The network must have been created on the problem's
   SUNContext (its kernels run on that context's stream, in order with the
   problem's vectors); -1 otherwise.  Several problems may share one network:
   each keeps its own output scale. */
int DiffusionProblem_SetLearnedPrec(DiffusionProblem* prob, LearnedPrec* lp);
/* preconditioner callbacks for ARKodeSetPreconditioner with the same
   user_data as PSetup/PSolve (the DiffusionProblem, which also carries the
   RHS, as the reference's UserData does, main.cpp:266, 307-313).  Setup
   refreshes the Jacobi diagonal (the fallback's preconditioner, so the pair
   PSetup/PSolve stays valid) and sets the problem's network output scale to
   1/(1 + gamma*(2kx/dx^2 + 2ky/dy^2)), the inverse diagonal of I - gamma*J
   for constant coefficients (the network's own scale is not changed);
   solve: z = N(r) with that scale. */
int LearnedPSetup(sunrealtype t, N_Vector u, N_Vector f, sunbooleantype jok,
                  sunbooleantype* jcurPtr, sunrealtype gamma, void* user_data);
int LearnedPSolve(sunrealtype t, N_Vector u, N_Vector f, N_Vector r, N_Vector z,
                  sunrealtype gamma, sunrealtype delta, int lr, void* user_data);

#ifdef __cplusplus
}
#endif
#endif
