// Internal interface between the diffusion problem kernels and the
// integrator's fused super-time-stepping path.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "diffusion_hip.h"
#include "nvhip_internal.hpp"

// launch configuration of one stencil kernel kind (the plain RHS or the fused
// STS stage): variant 0 = row-interleaved tiles, rows = rows per wavefront;
// variant 1 = y-marching, rows = consecutive rows per wavefront; nt = 1 uses
// non-temporal loads/stores for the streamed operands
struct StencilCfg {
  int variant = 0;
  int rows = 1;
  int nt = 0;
};

struct LearnedPrec;
struct MGHierarchy;

struct DiffusionProblem {
  DiffusionParams p;
  int64_t n = 0;
  double dx = 0, dy = 0;
  // host copies of the HBM tables (built with the same expressions as the
  // reference so device results are bitwise reproducible)
  std::vector<double> fx, fy, px, py, sx, sy, cx, cy, ax, by;
  double *d_fx = nullptr, *d_fy = nullptr, *d_px = nullptr, *d_py = nullptr;
  double *d_sx = nullptr, *d_sy = nullptr, *d_cx = nullptr, *d_cy = nullptr;
  double *d_ax = nullptr, *d_by = nullptr;
  double* d_tables = nullptr;  // the one allocation the d_ tables above point into
  double* d_tf = nullptr;   // device time factors (T1, T2) for the plain RHS
  double* h_tf = nullptr;   // pinned staging for d_tf
  N_Vector diag = nullptr;  // Jacobi inverse diagonal (udata->diag, main.cpp:223)
  double diag_gamma = 0.0;       // the gamma of the last PSetup, which computed diag
  bool diag_gamma_set = false;   // diag was last written by PSetup
  SUNContext ctx = nullptr;
  ceda::Exec* exec = nullptr;
  long nrhs = 0;
  // [0] RHS kernel, [1] fused stage kernel, [2] multi-stage pass (variant =
  // stages per pass K, 0 = single-stage kernels only; rows; nt)
  StencilCfg cfg[3];
  LearnedPrec* learned = nullptr;  // network of LearnedPSolve (not owned)
  double learned_out_scale = 1.0;  // its output scale for this problem (LearnedPSetup)
  MGHierarchy* mg = nullptr;       // multigrid hierarchy of MGPSetup/MGPSolve (owned)
  int mg_relax = 2, mg_nrelax = 2;  // V-cycle smoother: PFMG relax type 2|3, sweeps
  bool tuned = false;
};

namespace ceda {

void mg_destroy(MGHierarchy* H);

// One entry of the device-resident stage table of the fused STS path.  Every
// kernel of a step reads its coefficients, time factors and operand pointers
// from its entry, so a captured HIP graph stays valid when h, t and the buffer
// roles change between steps; only the stage count s selects the graph.
struct alignas(16) StageEntry {
  double c[5];          // combination coefficients
  double T1, T2, pad;   // MMS time factors of the RHS evaluated in this kernel
  const double* yjm1;   // stencil input
  const double* yjm2;
  const double* yn;
  const double* fn;
  double* out;
  const void* pad2[3];
};
static_assert(sizeof(StageEntry) == 128, "StageEntry layout");

// out = c0*yn + c1*fn (the first STS stage, N_VLinearSum order)
void launch_stage1_entry(DiffusionProblem* P, const StageEntry* e, hipStream_t s);
// out = c0*F(yjm1) + c1*yjm2 + c2*yn + c3*yjm1 + c4*fn
void launch_fused_entry(DiffusionProblem* P, const StageEntry* e, hipStream_t s);
// out = F(yjm1) with the entry's time factors
void launch_rhs_entry(DiffusionProblem* P, const StageEntry* e, hipStream_t s);

// K consecutive fused stages in one sweep (temporal blocking): stages read
// e[0].yjm1, e[0].yjm2, e[0].yn, e[0].fn; stage K-1 is written to e[K-2].out
// and stage K to e[K-1].out (K = 1..4).  rows = output rows per wavefront
// tile, nt = non-temporal streams.  Bitwise equal to K single-stage launches.
void launch_pass(DiffusionProblem* P, int K, int rows, int nt, const StageEntry* e, hipStream_t s);
// halo lanes of a K-stage pass (each side of a wavefront strip)
int pass_halo_lanes(int K, int cpl);

// ---- batched super-time-stepping (many independent problems, one launch
// per pass index) ----
constexpr int kMaxPassStages = 6;  // stages per multi-stage pass (one problem)
constexpr int kBatchMaxK = 5;  // stages per pass the batch tile geometry covers
struct BatchDesc {             // per problem, static
  const double *fx, *fy, *sx, *sy, *cx, *cy;
  int64_t nx, ny;
  int per, cpl, rows, strips;
};
struct BatchTile {             // one wavefront tile, problems in slot order
  int slot, prob, strip, rb;
};
int batch_halo_lanes(int cpl);
// column layout and boundary shared by every problem of a batch, so that
// the phase kernel instance holds only that layout's code and registers
enum { kLayoutAny = 0, kLayoutC2Mms = 1, kLayoutC2Per = 2 };
// phase 0: stage 1 of every slot; phase q >= 1: stages j = 2+(q-1)*kmax ..
// j+kmax-1 (clipped at s = stages[slot]) of the slots with j <= s, and the RHS
// at t_n + h (adaptive) of the slots whose last pass was phase q-1; entries of
// slot k start at ent + k*stride; tiles [0, ntiles) must cover exactly those
// slots (a prefix of the slots when they are sorted by s, descending)
void launch_batch_phase(const BatchDesc* desc, const BatchTile* tiles, int ntiles,
                        const StageEntry* ent, int stride, const int* stages, int phase, int kmax,
                        int adaptive, int layout, hipStream_t s);


// one STS stage of the LSRKStep recurrence, fused with the RHS evaluation:
//   out = c0*F(yjm1) + c1*yjm2 + c2*yn + c3*yjm1 + c4*fn
// coef points at 8 device doubles {c0..c4, T1, T2, unused}; F is the diffusion
// RHS with MMS time factors (T1, T2).  Bitwise equal to diffusion() followed
// by N_VLinearCombination(5, ...) in that association order.
void launch_fused_stage(DiffusionProblem* P, const double* coef, const double* yjm1,
                        const double* yjm2, const double* yn, const double* fn,
                        double* out, hipStream_t s);
// f = F(u) with the time factors read from the device (coef[5], coef[6])
void launch_rhs_devtime(DiffusionProblem* P, const double* coef, const double* u,
                        double* f, hipStream_t s);
// host evaluation of the MMS time factors of the forcing (README.md:28)
void mms_time_factors(double t, double* T1, double* T2);

}  // namespace ceda
