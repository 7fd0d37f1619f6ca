// Internal interface of the integrator layer: the per-problem state and the
// step logic, shared by the single-problem ARKODE entry points and the
// batched super-time-stepping engine (so both take identical steps).
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <vector>

#include "arkode_hip.h"
#include "diffusion_internal.hpp"

struct SUNAdaptController_ {
  int type;
};
struct SUNMatrix_ {
  int unused;
};
struct SUNDomEigEstimator_ {
  long maxiters = 100;
  double rtol = 0.01;
  N_Vector V = nullptr, q = nullptr, work = nullptr;
  int started = 0;
};
enum { LS_PCG = 0, LS_SPGMR = 1, LS_SPFGMR = 2, LS_FPCG = 3 };
struct SUNLinearSolver_ {
  int type = LS_PCG;
  int pretype = SUN_PREC_NONE;
  int maxl = 5;
  int mrestart = 30;
  N_Vector r = nullptr, p = nullptr, z = nullptr, Ap = nullptr;
  std::vector<N_Vector> V, Z;
  double* vz_pool = nullptr;  // one allocation holding every V and Z vector's data
  // fused PCG: per-workgroup partial sums and the solve state (device), and
  // the state's host copy (pinned); allocated on the first fused solve
  double* d_ws = nullptr;
  double* h_ws = nullptr;
  double* d_p2 = nullptr;  // its second direction buffer (length of the vectors)
  long last_iters = 0;     // iterations of the last fused solve (sizes its first batch)
};

namespace ceda_ark {

using ceda::StageEntry;

enum { M_RKC = 0, M_RKL = 1, M_DIRK = 2 };
constexpr uint64_t kDeeSeed = 20240917ull;
constexpr double kTiny = 1.0e-10;
constexpr double kOneMSM = 0.999999;
constexpr double kOnePSM = 1.000001;
constexpr int kRetry = 100;
constexpr int kConvRecover = 101;

struct ArkMem {
  SUNContext ctx = nullptr;
  int method = M_RKC;
  ARKRhsFn f = nullptr;
  void* ud = nullptr;
  int64_t n = 0;
  double rtol = 1.0e-5, atol = 1.0e-10;
  double tn = 0, h = 0, h0u = 0, hfixed = 0, hold = 0, h0_used = 0;
  double tstop = 0;
  bool tstopset = false;
  int fixedstep = 0, initialized = 0;
  long mxstep = 500;
  N_Vector yn = nullptr, ycur = nullptr, fn = nullptr, tv1 = nullptr, tv2 = nullptr,
           ewt = nullptr;
  N_Vector tv3 = nullptr;  // fourth scratch buffer of the multi-stage fused path
  int fn_is_current = 0, ewt_current = 0;
  // controller
  int ctype = ARK_ADAPT_PID, p_order = 2, firststeps = 0;
  double k1 = 0.58, k2 = 0.21, k3 = 0.1, bias = 1.5, ep = 0, epp = 0, hp = 0, hpp = 0;
  // adaptivity
  double safety = 0.96, growth = 20.0, etamx1 = 10000.0, etamxf = 0.3, etamin = 0.1,
         etacf = 0.25, lbound = 1.0, ubound = 1.5, eta = 1.0;
  int small_nef = 2, maxnef = 7, maxncf = 10;
  // LSRK
  double rho = 0, rho_max = 0, dom_eig_safety = 1.01;
  long dom_eig_freq = 25, dom_eig_nst = 0;
  int dom_eig_update = 1, smax = 200;
  ARKDomEigFn domeig = nullptr;
  SUNDomEigEstimator dee = nullptr;
  int dee_init_pre = 0, dee_pre = 0;
  // fused path
  DiffusionProblem* fast = nullptr;
  DiffusionProblem* fast_pcg = nullptr;  // fused PCG (ARKodeSetDiffusionFastPCG)
  double pcg_bulk = 0.75;                // ARKodeSetFusedPCGBulk
  int use_graphs = 0;
  StageEntry* h_tab = nullptr;
  StageEntry* d_tab = nullptr;
  hipEvent_t tab_ev = nullptr;
  int tab_cap = 0;
  std::map<long, hipGraphExec_t> graphs;
  // DIRK
  int order = 3, stages = 0, linear = 1, msbp = 20, psetup_done = 0, maxcor = 3;
  double A[16], b[4], bh[4], c[4];
  N_Vector F[4] = {nullptr, nullptr, nullptr, nullptr};
  N_Vector sdata = nullptr, zcur = nullptr, fz = nullptr, rhsv = nullptr, dz = nullptr,
           jwork = nullptr, jv = nullptr;
  double epslin = 0, ls_reltol = 0, gammap = 0, nlscoef = 0.1;
  // fallback Jacobi-PCG of a GMRES solve that misses its tolerance
  SUNLinearSolver fbLS = nullptr;
  ARKLsPrecSolveFn fb_psolve = nullptr;
  void* fb_ud = nullptr;
  long nls_fallback = 0;
  long nstlp = 0;
  SUNLinearSolver LS = nullptr;
  ARKLsPrecSetupFn psetup = nullptr;
  ARKLsPrecSolveFn psolve = nullptr;
  ARKLsJacTimesSetupFn jtsetup = nullptr;
  ARKLsJacTimesVecFn jtimes = nullptr;
  // statistics (same layout as the CPU restatement)
  long nst = 0, attempts = 0, nfe = 0, netf = 0, nstlim = 0, ndomeig = 0, ndee_iters = 0,
       stage_max = 0, last_stages = 0, nfe_dee = 0, nli = 0, npsetup = 0, npsolve = 0,
       nni = 0, ncfn = 0, nlsfail = 0, njtimes = 0, nfe_ls = 0;
};

struct StagePlan {
  int s;
  double mus1;                 // stage 1: Y1 = yn + (h*mus1) fn
  std::vector<double> c;       // 5 per stage j = 2..s
  std::vector<double> t;       // RHS time of stage j
};

struct AttemptState {
  int nef = 0, ncf = 0, force_setup = 0, clipped = 0;
};

int call_rhs(ArkMem* A, double t, N_Vector y, N_Vector f);
int compute_dom_eig(ArkMem* A);
void plan_stages(const ArkMem* A, int s, double h, double tn, StagePlan& P);
int lsrk_plan(ArkMem* A, StagePlan& SP);
void lsrk_after(ArkMem* A, double dsm);
// fills table entries 1..s+1 for one s-stage attempt and assigns the buffer
// roles exactly as the N_Vector path does (Y_s ends in ycur, F_s in tv2)
// stage table of one attempt for passes of up to kp stages (kp = 1: one
// kernel per stage)
void fused_fill(ArkMem* A, const StagePlan& SP, StageEntry* T, int kp);
// PCG in fused kernels with the scalars on the device (pcg_fused.hip);
// pmode 0 none, 1 Jacobi (P->diag), 2 the preconditioner solve ps
int pcg_fused(ArkMem* A, SUNLinearSolver S, DiffusionProblem* P, int pmode, ARKLsPrecSolveFn ps, void* ud,
              double t, double gamma, N_Vector b, N_Vector x);
// true (after a message on stderr) when a Krylov scalar is NaN or infinite:
// the solve then fails unrecoverably instead of iterating to its cap
bool krylov_nonfinite(const char* solver, const char* what, double v);
// stages per pass of a multi-stage configuration (1 when disabled)
inline int pass_stages(const StencilCfg& c) { return c.variant >= 2 ? c.variant : 1; }
void attempt_begin(ArkMem* A, double tout, AttemptState& S);
int attempt_end(ArkMem* A, int ret, double dsm, AttemptState& S);
int step_complete(ArkMem* A, double dsm, double tout, const AttemptState& S);
int evolve_init(ArkMem* A, double tout);

}  // namespace ceda_ark
