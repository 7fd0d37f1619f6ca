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
  double* d_tf = nullptr;   // device time factors (T1, T2) for the plain RHS
  double* h_tf = nullptr;   // pinned staging for d_tf
  N_Vector diag = nullptr;  // Jacobi inverse diagonal (udata->diag, main.cpp:223)
  SUNContext ctx = nullptr;
  ceda::Exec* exec = nullptr;
  long nrhs = 0;
  StencilCfg cfg[2];        // [0] RHS kernel, [1] fused stage kernel
  bool tuned = false;
};

namespace ceda {

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

// ---- batched super-time-stepping (many independent problems, one launch
// per stage index) ----
struct BatchDesc {           // per problem, static
  const double *fx, *fy, *sx, *sy, *cx, *cy;
  int64_t nx, ny;
  int per, cpl, rows, tiles_x;
};
struct BatchTile {           // per tile of a step group, problems in slot order
  int slot, prob, tx, ty;
};
struct BatchErr {            // per slot: operands of the embedded error norm
  const double *yn, *ys, *fn, *fs;
  double ch, rtol, atol, pad;
};
// phase j of a step group: for every tile t < ntiles, with s = stages[slot]:
// j == 1 first stage, 2 <= j <= s fused stage, j == s + 1 the RHS at t_n + h;
// entries of slot k start at ent + k*stride
void launch_batch_phase(const BatchDesc* desc, const BatchTile* tiles, const StageEntry* ent,
                        int stride, const int* stages, int j, int ntiles, hipStream_t s);
// per-tile sums of (e*ewt)^2, then per-slot sums in tile order (deterministic)
void launch_batch_err(const BatchDesc* desc, const BatchTile* tiles, const BatchErr* err,
                      int ntiles, double* partial, const int* tile_begin, const int* tile_count,
                      int nslots, double* dsq, hipStream_t s);

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
