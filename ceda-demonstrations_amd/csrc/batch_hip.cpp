// Batched super-time-stepping engine: many independent diffusion problems,
// each with its own adaptive step and stage count, advanced in step groups.
//
// Every problem is an ordinary LSRKStep integrator (ArkMem) on its own
// DiffusionProblem, and takes exactly the decisions the single-problem path
// takes (attempt_begin / lsrk_plan / lsrk_after / attempt_end /
// step_complete are shared).  What changes is the execution of the stage
// loops: in a step group every active problem plans one attempt, the planned
// problems are sorted by stage count (descending) and their tiles are listed
// in that order, so the tiles still in their stage loop at a given pass index
// are a prefix of the list; a phase is one launch over that prefix.  Each
// phase runs up to kpass stages of every problem in it (the multi-stage pass
// kernel: intermediate stages stay in registers), or its first stage, or its
// closing RHS.  Every problem's error norm is then reduced in the very tree
// of the single-problem path (raw_sts_error_wsqr, as many virtual workgroups
// of one launch), so every decision equals that path's bit for bit; one
// device-to-host copy per group.  Phase launches of small problems, which
// would each be limited by launch latency when the problems are solved one
// after another, are amortised over the whole batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "arkode_internal.hpp"
#include "batch_hip.h"

using namespace ceda_ark;
using ceda::BatchDesc;
using ceda::BatchTile;
using ceda::StageEntry;

namespace {

struct BProb {
  DiffusionProblem* P = nullptr;
  N_Vector y0 = nullptr;
  ArkMem* A = nullptr;
  AttemptState S;
  long nst0 = 0;
  bool done = false, new_step = true;
  int err = 0;
  int rows = 0, cpl = 1, strips = 0, rbs = 0;
  int rows_f = 0, rbs_f = 0;  // the fine tiles of thin phases
  StagePlan SP;
};

int tiles_of(const BProb& b) { return b.strips * b.rbs; }
int fine_tiles_of(const BProb& b) { return b.strips * b.rbs_f; }

}  // namespace

struct DiffusionBatch {
  SUNContext ctx = nullptr;
  ceda::Exec* ex = nullptr;
  std::vector<BProb> probs;
  int method = M_RKC;
  int smax = 200, stride = 202;
  int tile_rows = 0;
  // thin phases (fewer than thin_tiles wavefront tiles, late in a step group
  // where only the problems with the most stages are left) run on tiles of
  // thin_rows rows: more, shorter column marches, each re-computing the same
  // 2(K-1) halo rows, so the phase ends sooner on an underfilled GPU; the
  // cells' values do not depend on the tiling.  thin_rows = 0: coarse tiles only
  int thin_rows = 12, thin_tiles = 2048;
  int kpass = ceda::kBatchMaxK;
  int layout = ceda::kLayoutAny;  // shared column layout and boundary of the problems
  bool tuned = false;
  long ngroups = 0, nlaunches = 0;
  // algorithmic HBM bytes of the phase kernels (FP64 cells): stage 1 reads
  // y_n, F_n and writes Y_1 (24 B); a pass of K stages reads Y_{j-1},
  // Y_{j-2}, y_n, F_n and writes two stages (48 B); the closing RHS 16 B; the
  // error norm reads y_n, Y_s, F_n, F_s (32 B)
  double alg_bytes = 0.0;
  // device/pinned staging (grown on demand)
  BatchDesc* d_desc = nullptr;
  BatchDesc* d_desc_f = nullptr;  // the same problems with the fine tiles' geometry
  StageEntry *d_ent = nullptr, *h_ent = nullptr;
  BatchTile *d_tiles = nullptr, *h_tiles = nullptr;
  BatchTile *d_tiles_f = nullptr, *h_tiles_f = nullptr;
  ceda::StsErrItem *d_err = nullptr, *h_err = nullptr;  // per slot: its error-norm reduction
  int *d_s = nullptr, *h_s = nullptr;
  double *d_part = nullptr, *d_dsq = nullptr, *h_dsq = nullptr;
  int cap_slots = 0, cap_tiles = 0, cap_vblocks = 0, cap_tiles_f = 0;
};

namespace {

template <class T>
void dev_alloc(T** d, size_t n)
{
  CEDA_HIP_CHECK(ceda::dev_malloc(d, sizeof(T) * (n ? n : 1)));
}
template <class T>
void host_alloc(T** h, size_t n)
{
  CEDA_HIP_CHECK(ceda::host_malloc(h, sizeof(T) * (n ? n : 1), hipHostMallocDefault));
}

void free_staging(DiffusionBatch* B)
{
  void* ds[] = {B->d_ent, B->d_tiles, B->d_tiles_f, B->d_err, B->d_s, B->d_part, B->d_dsq};
  for (void* p : ds)
    if (p) CEDA_HIP_CHECK(ceda::dev_free(p));
  void* hs[] = {B->h_ent, B->h_tiles, B->h_tiles_f, B->h_err, B->h_s, B->h_dsq};
  for (void* p : hs)
    if (p) CEDA_HIP_CHECK(ceda::host_free(p));
  B->d_ent = B->h_ent = nullptr;
  B->d_tiles = B->h_tiles = nullptr;
  B->d_tiles_f = B->h_tiles_f = nullptr;
  B->d_err = B->h_err = nullptr;
  B->d_s = B->h_s = nullptr;
  B->d_part = B->d_dsq = B->h_dsq = nullptr;
}

void ensure_staging(DiffusionBatch* B, int slots, int tiles, int vblocks = 0, int tiles_f = 0)
{
  if (slots <= B->cap_slots && tiles <= B->cap_tiles && vblocks <= B->cap_vblocks && tiles_f <= B->cap_tiles_f)
    return;
  CEDA_HIP_CHECK(hipStreamSynchronize(B->ex->stream));
  free_staging(B);
  B->cap_slots = std::max(slots, B->cap_slots);
  B->cap_tiles = std::max(tiles, B->cap_tiles);
  B->cap_vblocks = std::max(vblocks, B->cap_vblocks);
  B->cap_tiles_f = std::max(tiles_f, B->cap_tiles_f);
  const size_t ns = B->cap_slots, nt = B->cap_tiles;
  dev_alloc(&B->d_tiles_f, B->cap_tiles_f);
  host_alloc(&B->h_tiles_f, B->cap_tiles_f);
  dev_alloc(&B->d_ent, ns * B->stride);
  host_alloc(&B->h_ent, ns * B->stride);
  memset(B->h_ent, 0, sizeof(StageEntry) * ns * B->stride);
  dev_alloc(&B->d_tiles, nt);
  host_alloc(&B->h_tiles, nt);
  dev_alloc(&B->d_err, ns);
  host_alloc(&B->h_err, ns);
  dev_alloc(&B->d_s, ns);
  host_alloc(&B->h_s, ns);
  dev_alloc(&B->d_part, B->cap_vblocks);  // the error norms' per-workgroup partials
  dev_alloc(&B->d_dsq, ns);
  host_alloc(&B->h_dsq, ns);
}

// wavefront tiles: strips of (64 - 2L)*cpl owned columns, `rows` rows
void set_geometry(BProb& b, int rows)
{
  const DiffusionParams& p = b.P->p;
  b.cpl = (p.nx % 2 == 0) ? 2 : 1;
  b.rows = rows;
  const int64_t ow = (int64_t)(64 - 2 * ceda::batch_halo_lanes(b.cpl)) * b.cpl;
  b.strips = (int)((p.nx + ow - 1) / ow);
  b.rbs = (int)((p.ny + rows - 1) / rows);
}

void list_tiles(const BProb& b, int slot, int prob, BatchTile* out, int& t)
{
  for (int rb = 0; rb < b.rbs; rb++)
    for (int x = 0; x < b.strips; x++) out[t++] = BatchTile{slot, prob, x, rb};
}
void list_fine_tiles(const BProb& b, int slot, int prob, BatchTile* out, int& t)
{
  for (int rb = 0; rb < b.rbs_f; rb++)
    for (int x = 0; x < b.strips; x++) out[t++] = BatchTile{slot, prob, x, rb};
}

// the problems' descriptors with the coarse tiles' geometry and with the
// fine one (at most thin_rows rows, never more than the coarse tiles')
void upload_desc(DiffusionBatch* B)
{
  std::vector<BatchDesc> h(B->probs.size()), hf(B->probs.size());
  for (size_t k = 0; k < B->probs.size(); k++) {
    BProb& b = B->probs[k];
    const DiffusionProblem* P = b.P;
    b.rows_f = B->thin_rows > 0 ? std::min(B->thin_rows, b.rows) : b.rows;
    b.rbs_f = (int)((P->p.ny + b.rows_f - 1) / b.rows_f);
    h[k] = BatchDesc{P->d_fx, P->d_fy, P->d_sx, P->d_sy, P->d_cx, P->d_cy, P->p.nx, P->p.ny,
                     P->p.kind == DIFFUSION_CEDA_PERIODIC ? 1 : 0, b.cpl, b.rows, b.strips};
    hf[k] = h[k];
    hf[k].rows = b.rows_f;
  }
  if (!B->d_desc) dev_alloc(&B->d_desc, h.size());
  if (!B->d_desc_f) dev_alloc(&B->d_desc_f, h.size());
  CEDA_HIP_CHECK(ceda::copy_sync(B->d_desc, h.data(), sizeof(BatchDesc) * h.size(), hipMemcpyHostToDevice));
  CEDA_HIP_CHECK(ceda::copy_sync(B->d_desc_f, hf.data(), sizeof(BatchDesc) * hf.size(), hipMemcpyHostToDevice));
}

// the tile autotuner: for every size class (nx, ny) time a full-length pass
// phase over that class's tiles for each candidate height and keep the
// fastest
void autotune_tiles(DiffusionBatch* B)
{
  std::map<std::pair<int64_t, int64_t>, std::vector<int>> classes;
  for (size_t k = 0; k < B->probs.size(); k++)
    classes[{B->probs[k].P->p.nx, B->probs[k].P->p.ny}].push_back((int)k);
  hipStream_t s = B->ex->stream;
  hipEvent_t e0, e1;
  CEDA_HIP_CHECK(hipEventCreate(&e0));
  CEDA_HIP_CHECK(hipEventCreate(&e1));
  const int K = B->kpass;
  for (auto& cl : classes) {
    const std::vector<int>& ks = cl.second;
    int best_rows = 16;
    float best = 1e30f;
    for (int rows : {4, 8, 16, 24, 32, 48, 64}) {
      if (rows > 2 * cl.first.second) continue;
      for (int k : ks) set_geometry(B->probs[k], rows);
      upload_desc(B);
      int nt = 0;
      for (int k : ks) nt += tiles_of(B->probs[k]);
      ensure_staging(B, (int)ks.size(), nt);
      int t = 0;
      for (size_t q = 0; q < ks.size(); q++) {
        BProb& b = B->probs[ks[q]];
        // stages 2..K+1 read yn only and write tv1 (stage K) and tv2
        // (stage K+1); scratch contents at a step boundary
        for (int j = 2; j <= K + 1; j++) {
          StageEntry& e = B->h_ent[q * B->stride + j];
          memset(&e, 0, sizeof(e));
          e.yjm1 = e.yjm2 = e.yn = e.fn = N_VGetDeviceArrayPointer(b.A->yn);
          e.out = N_VGetDeviceArrayPointer(j == K + 1 ? b.A->tv2 : b.A->tv1);
        }
        B->h_s[q] = K + 1;
        list_tiles(b, (int)q, ks[q], B->h_tiles, t);
      }
      CEDA_HIP_CHECK(hipMemcpyAsync(B->d_ent, B->h_ent, sizeof(StageEntry) * ks.size() * B->stride,
                                    hipMemcpyHostToDevice, s));
      CEDA_HIP_CHECK(hipMemcpyAsync(B->d_tiles, B->h_tiles, sizeof(BatchTile) * nt,
                                    hipMemcpyHostToDevice, s));
      CEDA_HIP_CHECK(hipMemcpyAsync(B->d_s, B->h_s, sizeof(int) * ks.size(), hipMemcpyHostToDevice, s));
      ceda::launch_batch_phase(B->d_desc, B->d_tiles, nt, B->d_ent, B->stride, B->d_s, 1, K, 0, B->layout, s);
      CEDA_HIP_CHECK(hipEventRecord(e0, s));
      for (int r = 0; r < 3; r++)
        ceda::launch_batch_phase(B->d_desc, B->d_tiles, nt, B->d_ent, B->stride, B->d_s, 1, K, 0, B->layout, s);
      CEDA_HIP_CHECK(hipEventRecord(e1, s));
      CEDA_HIP_CHECK(hipEventSynchronize(e1));
      float ms = 0.0f;
      CEDA_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) {
        best = ms;
        best_rows = rows;
      }
    }
    for (int k : ks) set_geometry(B->probs[k], best_rows);
  }
  CEDA_HIP_CHECK(hipEventDestroy(e0));
  CEDA_HIP_CHECK(hipEventDestroy(e1));
  upload_desc(B);
  B->tuned = true;
}

// one step group: every active problem plans an attempt, all planned stage
// loops run together, then every problem finishes its attempt
int step_group(DiffusionBatch* B, double tout, bool* progress)
{
  *progress = false;
  std::vector<int> planned;
  for (size_t k = 0; k < B->probs.size(); k++) {
    BProb& b = B->probs[k];
    if (b.done) continue;
    ArkMem* A = b.A;
    if (A->tn >= tout) {
      b.done = true;
      continue;
    }
    if (b.new_step && A->mxstep > 0 && A->nst - b.nst0 >= A->mxstep) {
      b.err = ARK_TOO_MUCH_WORK;
      b.done = true;
      continue;
    }
    if (b.new_step) {
      b.S = AttemptState();
      b.new_step = false;
    }
    for (;;) {
      attempt_begin(A, tout, b.S);
      const int r = lsrk_plan(A, b.SP);
      if (r == 0) {
        planned.push_back((int)k);
        break;
      }
      const int r2 = attempt_end(A, r, 0.0, b.S);
      if (r2 == 1) continue;
      b.err = r2 < 0 ? r2 : r;
      b.done = true;
      break;
    }
  }
  if (planned.empty()) return 0;
  *progress = true;
  std::stable_sort(planned.begin(), planned.end(), [&](int a, int c) {
    return B->probs[a].SP.s > B->probs[c].SP.s;
  });
  const bool adaptive = !B->probs[planned[0]].A->fixedstep;
  const bool thin = B->thin_rows > 0;
  int ntiles = 0, nvblocks = 0, ntiles_f = 0;
  for (int k : planned) {
    ntiles += tiles_of(B->probs[k]);
    if (thin) ntiles_f += fine_tiles_of(B->probs[k]);
    nvblocks += ceda::sts_err_blocks(B->probs[k].P->n);
  }
  const int nslots = (int)planned.size();
  ensure_staging(B, nslots, ntiles, nvblocks, ntiles_f);
  int t = 0, vb = 0, tf = 0;
  std::vector<int> tile_end(nslots), tile_end_f(nslots);
  for (int q = 0; q < nslots; q++) {
    BProb& b = B->probs[planned[q]];
    ArkMem* A = b.A;
    StageEntry* T = B->h_ent + (size_t)q * B->stride;
    fused_fill(A, b.SP, T, B->kpass);
    B->h_s[q] = b.SP.s;
    list_tiles(b, q, planned[q], B->h_tiles, t);
    tile_end[q] = t;
    if (thin) list_fine_tiles(b, q, planned[q], B->h_tiles_f, tf);
    tile_end_f[q] = tf;
    // the single-problem path's reduction (raw_sts_error_wsqr), bit for bit
    const int g = ceda::sts_err_blocks(A->n);
    B->h_err[q] = ceda::StsErrItem{N_VGetDeviceArrayPointer(A->yn), N_VGetDeviceArrayPointer(A->ycur),
                                   N_VGetDeviceArrayPointer(A->fn), N_VGetDeviceArrayPointer(A->tv2),
                                   0.4 * A->h, A->rtol, A->atol, A->n, vb, g};
    vb += g;
  }
  hipStream_t s = B->ex->stream;
  CEDA_HIP_CHECK(hipMemcpyAsync(B->d_ent, B->h_ent, sizeof(StageEntry) * nslots * B->stride,
                                hipMemcpyHostToDevice, s));
  CEDA_HIP_CHECK(hipMemcpyAsync(B->d_tiles, B->h_tiles, sizeof(BatchTile) * ntiles,
                                hipMemcpyHostToDevice, s));
  if (thin)
    CEDA_HIP_CHECK(hipMemcpyAsync(B->d_tiles_f, B->h_tiles_f, sizeof(BatchTile) * ntiles_f,
                                  hipMemcpyHostToDevice, s));
  CEDA_HIP_CHECK(hipMemcpyAsync(B->d_s, B->h_s, sizeof(int) * nslots, hipMemcpyHostToDevice, s));
  CEDA_HIP_CHECK(hipMemcpyAsync(B->d_err, B->h_err, sizeof(ceda::StsErrItem) * nslots, hipMemcpyHostToDevice, s));
  // phase 0: stage 1 of every slot; phase q >= 1 starts at stage
  // j = 2 + (q-1)*kpass and covers the slots with s >= j (their next pass) or,
  // adaptive, s >= j - kpass (their closing RHS): a prefix of the slots
  const int kp = B->kpass;
  ceda::launch_batch_phase(B->d_desc, B->d_tiles, ntiles, B->d_ent, B->stride, B->d_s, 0, kp,
                           adaptive ? 1 : 0, B->layout, s);
  B->nlaunches++;
  int active = nslots;
  for (int q = 1;; q++) {
    const int j = 2 + (q - 1) * kp;
    const int need = adaptive ? j - kp : j;
    while (active > 0 && B->h_s[active - 1] < need) active--;
    if (active == 0) break;
    if (thin && tile_end[active - 1] < B->thin_tiles)
      ceda::launch_batch_phase(B->d_desc_f, B->d_tiles_f, tile_end_f[active - 1], B->d_ent, B->stride,
                               B->d_s, q, kp, adaptive ? 1 : 0, B->layout, s);
    else
      ceda::launch_batch_phase(B->d_desc, B->d_tiles, tile_end[active - 1], B->d_ent, B->stride,
                               B->d_s, q, kp, adaptive ? 1 : 0, B->layout, s);
    B->nlaunches++;
  }
  if (adaptive) {
    ceda::launch_sts_err_batch(B->d_err, nslots, nvblocks, B->d_part, B->d_dsq, s);
    CEDA_HIP_CHECK(hipMemcpyAsync(B->h_dsq, B->d_dsq, sizeof(double) * nslots,
                                  hipMemcpyDeviceToHost, s));
  }
  CEDA_HIP_CHECK(hipGetLastError());
  for (int q = 0; q < nslots; q++) {
    const BProb& b = B->probs[planned[q]];
    const double cells = (double)b.P->n;
    const int passes = (b.SP.s - 1 + kp - 1) / kp;
    const int last_k = (b.SP.s - 1) - kp * (passes - 1);  // a one-stage pass writes one stage: 40 B
    B->alg_bytes += cells * (24.0 + 48.0 * (passes - 1) + (last_k >= 2 ? 48.0 : 40.0) +
                             (adaptive ? 16.0 + 32.0 : 0.0));
  }
  CEDA_HIP_CHECK(hipStreamSynchronize(s));
  B->ngroups++;
  for (int q = 0; q < nslots; q++) {
    BProb& b = B->probs[planned[q]];
    ArkMem* A = b.A;
    const long nrhs = (b.SP.s - 1) + (adaptive ? 1 : 0);
    A->nfe += nrhs;
    b.P->nrhs += nrhs;
    const double dsm = adaptive ? sqrt(B->h_dsq[q] / (double)A->n) : 0.0;
    lsrk_after(A, dsm);
    const int r = attempt_end(A, 0, dsm, b.S);
    if (r == 0) {
      const int rc = step_complete(A, dsm, tout, b.S);
      if (rc < 0) {
        b.err = rc;
        b.done = true;
      }
      b.new_step = true;
    } else if (r < 0) {
      b.err = r;
      b.done = true;
    }
  }
  return 0;
}

}  // namespace

extern "C" {

DiffusionBatch* DiffusionBatch_Create(int nprob, const DiffusionParams* params,
                                      ARKODE_LSRKMethodType method, SUNContext ctx)
{
  if (nprob < 1 || !params || !ctx) return nullptr;
  DiffusionBatch* B = new DiffusionBatch();
  B->ctx = ctx;
  B->ex = ceda::context_exec(ctx);
  B->method = (method == ARKODE_LSRK_RKL_2) ? M_RKL : M_RKC;
  B->probs.resize(nprob);
  for (int k = 0; k < nprob; k++) {
    BProb& b = B->probs[k];
    b.P = DiffusionProblem_Create(&params[k], ctx);
    if (!b.P) {
      DiffusionBatch_Destroy(B);
      return nullptr;
    }
    b.y0 = N_VNew_HipNative(b.P->n, ctx);
    Initial(0.0, b.y0, b.P);
    void* mem = LSRKStepCreateSTS(diffusion, 0.0, b.y0, ctx);
    b.A = static_cast<ArkMem*>(mem);
    LSRKStepSetSTSMethod(mem, method);
    ARKodeSetUserData(mem, b.P);
    LSRKStepSetDomEigFn(mem, dom_eig);
    b.A->fast = b.P;  // error weights only when needed, as in the fused path
    b.A->tv3 = N_VClone(b.A->yn);
    set_geometry(b, 16);
  }
  // one kernel instance per shared layout (two columns per lane: even sides)
  bool c2 = true, all_per = true, all_mms = true;
  for (const BProb& b : B->probs) {
    c2 = c2 && b.cpl == 2;
    const bool per = b.P->p.kind == DIFFUSION_CEDA_PERIODIC;
    all_per = all_per && per;
    all_mms = all_mms && !per;
  }
  if (c2 && all_mms) B->layout = ceda::kLayoutC2Mms;
  else if (c2 && all_per) B->layout = ceda::kLayoutC2Per;
  upload_desc(B);
  return B;
}

void DiffusionBatch_Destroy(DiffusionBatch* B)
{
  if (!B) return;
  CEDA_HIP_CHECK(hipStreamSynchronize(B->ex->stream));
  free_staging(B);
  if (B->d_desc) CEDA_HIP_CHECK(ceda::dev_free(B->d_desc));
  if (B->d_desc_f) CEDA_HIP_CHECK(ceda::dev_free(B->d_desc_f));
  for (BProb& b : B->probs) {
    if (b.A) {
      void* m = b.A;
      ARKodeFree(&m);
    }
    if (b.y0) N_VDestroy(b.y0);
    if (b.P) DiffusionProblem_Destroy(b.P);
  }
  delete B;
}

int DiffusionBatch_SetTolerances(DiffusionBatch* B, sunrealtype rtol, sunrealtype atol)
{
  if (!B) return ARK_MEM_NULL;
  for (BProb& b : B->probs) ARKodeSStolerances(b.A, rtol, atol);
  return 0;
}

int DiffusionBatch_SetController(DiffusionBatch* B, int controller)
{
  if (!B || controller < 0 || controller > 5) return ARK_ILL_INPUT;
  for (BProb& b : B->probs) {
    b.A->ctype = controller;
    SUNAdaptController C = SUNAdaptController_PID(B->ctx);
    C->type = controller;
    ARKodeSetAdaptController(b.A, C);
    SUNAdaptController_Destroy(C);
  }
  return 0;
}

int DiffusionBatch_SetMaxNumStages(DiffusionBatch* B, int smax)
{
  if (!B || smax < 2) return ARK_ILL_INPUT;
  for (BProb& b : B->probs) LSRKStepSetMaxNumStages(b.A, smax);
  B->smax = smax;
  CEDA_HIP_CHECK(hipStreamSynchronize(B->ex->stream));
  free_staging(B);
  B->cap_slots = B->cap_tiles = 0;
  B->stride = smax + 2;
  return 0;
}

int DiffusionBatch_SetMaxNumSteps(DiffusionBatch* B, long mxsteps)
{
  if (!B) return ARK_MEM_NULL;
  for (BProb& b : B->probs) ARKodeSetMaxNumSteps(b.A, mxsteps);
  return 0;
}

int DiffusionBatch_SetFixedStep(DiffusionBatch* B, sunrealtype h)
{
  if (!B) return ARK_MEM_NULL;
  for (BProb& b : B->probs) ARKodeSetFixedStep(b.A, h);
  return 0;
}

int DiffusionBatch_SetTileRows(DiffusionBatch* B, int rows)
{
  if (!B || rows < 0 || rows > 256) return ARK_ILL_INPUT;
  B->tile_rows = rows;
  if (rows > 0) {
    for (BProb& b : B->probs) set_geometry(b, rows);
    upload_desc(B);
    B->tuned = true;
  } else B->tuned = false;
  return 0;
}

int DiffusionBatch_SetThinPhases(DiffusionBatch* B, int rows, int min_tiles)
{
  if (!B || rows < 0 || rows > 256 || min_tiles < 1) return ARK_ILL_INPUT;
  B->thin_rows = rows;
  B->thin_tiles = min_tiles;
  if (B->tuned) upload_desc(B);
  return 0;
}

int DiffusionBatch_SetStagesPerPass(DiffusionBatch* B, int kpass)
{
  if (!B || kpass < 1 || kpass > ceda::kBatchMaxK) return ARK_ILL_INPUT;
  B->kpass = kpass;
  if (!B->tile_rows) B->tuned = false;
  return 0;
}

int DiffusionBatch_Evolve(DiffusionBatch* B, sunrealtype tout)
{
  if (!B) return ARK_MEM_NULL;
  for (BProb& b : B->probs) {
    if (!b.A->initialized) {
      const int r = evolve_init(b.A, tout);
      if (r) return r;
    }
    b.nst0 = b.A->nst;
    b.done = (b.A->tn >= tout);
    b.new_step = true;
    b.err = 0;
  }
  if (!B->tuned) autotune_tiles(B);
  bool progress = true;
  while (progress) {
    const int r = step_group(B, tout, &progress);
    if (r) return r;
  }
  for (BProb& b : B->probs)
    if (b.err) return b.err;
  return 0;
}

int DiffusionBatch_GetSolution(DiffusionBatch* B, int p, N_Vector y)
{
  if (!B || p < 0 || p >= (int)B->probs.size()) return ARK_ILL_INPUT;
  if (!y || N_VGetLength(y) != B->probs[p].A->n) return ARK_ILL_INPUT;
  N_VScale(1.0, B->probs[p].A->yn, y);
  return 0;
}

int DiffusionBatch_GetCurrentTime(DiffusionBatch* B, int p, sunrealtype* t)
{
  if (!B || p < 0 || p >= (int)B->probs.size()) return ARK_ILL_INPUT;
  *t = B->probs[p].A->tn;
  return 0;
}

int DiffusionBatch_GetStatsArray(DiffusionBatch* B, int p, long* out, int nout)
{
  if (!B || p < 0 || p >= (int)B->probs.size()) return ARK_ILL_INPUT;
  return ARKodeGetStatsArray(B->probs[p].A, out, nout);
}

int DiffusionBatch_GetAlgorithmicBytes(DiffusionBatch* B, double* bytes)
{
  if (!B || !bytes) return ARK_ILL_INPUT;
  *bytes = B->alg_bytes;
  return 0;
}

int DiffusionBatch_GetNumGroups(DiffusionBatch* B, long* ngroups, long* nlaunches)
{
  if (!B) return ARK_MEM_NULL;
  if (ngroups) *ngroups = B->ngroups;
  if (nlaunches) *nlaunches = B->nlaunches;
  return 0;
}

}  // extern "C"
