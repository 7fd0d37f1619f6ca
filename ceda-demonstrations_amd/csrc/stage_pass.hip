// Multi-stage super-time-stepping passes: K consecutive stages of the
// LSRKStep recurrence
//   Y_j = c0_j*F(t_j, Y_{j-1}) + c1_j*Y_{j-2} + c2_j*Y_n + c3_j*Y_{j-1} + c4_j*F_n
// in one sweep over the grid (temporal blocking).  A stage needs its
// predecessor only on the 5-point neighbourhood, so a wavefront that walks
// down a column strip can keep every intermediate stage in registers: the
// three most recent rows of each level form the stencil window of the next
// level, which lags one row behind.  Only Y_{j-1}, Y_{j-2}, Y_n, F_n are read
// and only the last two stages are written (the next pass needs both), so a
// pass moves 48 bytes per cell for K stages instead of 40 per stage.
//
// Geometry.  Lane l of a wavefront holds CPL adjacent columns; the strip of a
// wavefront is 64*CPL columns wide but owns only the middle (64-2L)*CPL: the
// outermost L lanes on each side recompute their neighbours' columns, because
// a level is valid one column less on each side than the level below it.
// The first level gets its west/east neighbours from two loads (lanes 0 and
// 63), the others by wavefront shuffles.  In y the march starts K-1 rows
// above the tile and ends K-1 rows below it for the same reason.  Out-of-
// domain cells are the zero Dirichlet ghosts, or the periodic wrap of the
// one-rank reference (diffusion_2D.cpp:242-327).
//
// Every stage evaluates exactly cell(), forcing() and lc5() of the single-
// stage kernels in the same order, so a K-stage pass is bitwise identical to
// K single-stage launches.  A pass must not write a buffer it reads (other
// wavefronts still need the halo rows of Y_{j-1}, Y_{j-2}): the stage plan
// rotates four buffers.
#include "diffusion_internal.hpp"
#include "stencil_ops.hpp"

namespace ceda {

namespace {

struct PassGeo {
  const double *fx, *fy, *sx, *sy, *cx, *cy;
  int64_t nx, ny;
  int rows;    // output rows of a wavefront tile
  int strips;  // wavefront strips across x
};

constexpr int kWavesPerBlock = kBlock / kWave;

// index r of a row or column wrapped into [0, n) (periodic; r is at most a
// few strip widths outside) or flagged outside (Dirichlet ghost)
template <bool PER>
__device__ inline bool map_index(int64_t r, int64_t n, int64_t& rw)
{
  rw = r;
  if (PER) {
    while (rw < 0) rw += n;
    while (rw >= n) rw -= n;
    return true;
  }
  return r >= 0 && r < n;
}

// operands in the global address space: global_load/global_store (not flat
// accesses, which would also count against the LDS/shuffle wait counter)
typedef __attribute__((address_space(1))) const double gcdouble;
typedef __attribute__((address_space(1))) double gdouble;
typedef __attribute__((address_space(1))) const d2v gcd2v;
typedef __attribute__((address_space(1))) d2v gd2v;
// neighbour exchange across the 64 lanes by DPP wave shifts (VALU moves, no
// LDS round trip): lane i receives lane i-1 (from_west) or lane i+1
// (from_east); the end lanes keep their own value, as __shfl_up/down do
__device__ inline double from_west(double x)
{
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(x), __double2loint(x), 0x138, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(x), __double2hiint(x), 0x138, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
__device__ inline double from_east(double x)
{
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(x), __double2loint(x), 0x130, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(x), __double2hiint(x), 0x130, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

// read-only per-row coefficient tables: constant address space, so that a
// wavefront-uniform index compiles to a scalar load
typedef __attribute__((address_space(4))) const double ccdouble;

template <bool NT>
__device__ inline void gload(gcdouble* p, double* v, int cpl)
{
  if (cpl == 2) {
    const d2v t = NT ? __builtin_nontemporal_load(reinterpret_cast<gcd2v*>(p))
                     : *reinterpret_cast<gcd2v*>(p);
    v[0] = t.x;
    v[1] = t.y;
  } else v[0] = NT ? __builtin_nontemporal_load(p) : *p;
}
template <bool NT>
__device__ inline void gstore(gdouble* p, const double* v, int cpl)
{
  if (cpl == 2) {
    d2v w;
    w.x = v[0];
    w.y = v[1];
    if (NT) __builtin_nontemporal_store(w, reinterpret_cast<gd2v*>(p));
    else *reinterpret_cast<gd2v*>(p) = w;
  } else if (NT) __builtin_nontemporal_store(v[0], p);
  else *p = v[0];
}

// K stages (FUSED) or the plain RHS F(t, Y) (K = 1, !FUSED) on one wavefront
// tile: output columns [strip*OW, strip*OW+OW) x rows [rb*rows, rb*rows+rows).
// e[m] is the stage-table entry of the (m+1)-th stage of the pass; the input
// operands come from e[0], the outputs go to e[K-2].out and e[K-1].out.
// strip and rb must be wavefront-uniform (scalar registers): the row indices
// then stay scalar and the per-row coefficient tables come through the scalar
// cache, so the only vector-memory waits are those of the row prefetch.
//
// SCALAR_E: e is wavefront-uniform (scalar registers).  The level's T1 is then
// re-read into a scalar register every row step through an empty asm, so the
// compiler cannot hoist the K x CPL products T1*sx out of the march: they
// would hold 4K more VGPRs, which at K = 5 costs a wavefront per SIMD.
// CLDS: the stage coefficients (5 per level, with T1 and T2) are read from
// LDS (lc[8m + k]) at every row step instead of being held in scalar
// registers for the whole march, where K = 5 spills 46 of them to vector
// lanes (read back by v_readlane on the VALU at every use).
typedef __attribute__((address_space(3))) const double lcdouble;

template <int CPL, bool PER, bool MMS, int K, int L, bool FUSED, bool NT, bool SCALAR_E = false, bool CLDS = false>
__device__ __forceinline__ void pass_wave(const PassGeo& g, const StageEntry* __restrict__ e,
                                          int strip, int rb, lcdouble* __restrict__ lc = nullptr)
{
  static_assert(K >= 1 && K <= kMaxPassStages, "1..6 stages per pass");
  static_assert(L * CPL >= K - 1, "halo lanes must cover the shrinking levels");
  static_assert(FUSED || K == 1, "the plain RHS is a single level");
  constexpr int OW = (kWave - 2 * L) * CPL;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t nx = g.nx, ny = g.ny;
  const int64_t c = (int64_t)strip * OW - (int64_t)(L * CPL) + (int64_t)lane * CPL;
  int64_t cw;
  const bool in = map_index<PER>(c, nx, cw);
  const bool own = (lane >= L) && (lane < kWave - L) && (c < nx);
  int64_t hw, he;
  const bool hw_ok = map_index<PER>(c - 1, nx, hw);
  const bool he_ok = map_index<PER>(c + CPL, nx, he);
  const bool ld_w = (lane == 0) && hw_ok, ld_e = (lane == kWave - 1) && he_ok;

  double Dw0 = 0.0, De0 = 0.0, De1 = 0.0, sx0 = 0.0, sx1 = 0.0, cx0 = 0.0, cx1 = 0.0;
  if (in) {
    Dw0 = g.fx[cw];
    De0 = g.fx[cw + 1];
    if (CPL == 2) De1 = g.fx[cw + 2];
    if (MMS) {
      sx0 = g.sx[cw];
      cx0 = g.cx[cw];
      if (CPL == 2) {
        sx1 = g.sx[cw + 1];
        cx1 = g.cx[cw + 1];
      }
    }
  }
  constexpr int KC = CLDS ? 1 : K;  // coefficient registers (CLDS: unused)
  double cc[KC][5], T1[KC], T2[KC];
#pragma unroll
  for (int m = 0; m < KC; m++) {
#pragma unroll
    for (int k = 0; k < 5; k++) cc[m][k] = (FUSED && !CLDS) ? e[m].c[k] : 0.0;
    T1[m] = CLDS ? 0.0 : e[m].T1;
    T2[m] = CLDS ? 0.0 : e[m].T2;
  }
  gcdouble* __restrict__ y0 = (gcdouble*)e[0].yjm1;
  gcdouble* __restrict__ ym = (gcdouble*)e[0].yjm2;
  gcdouble* __restrict__ yn = (gcdouble*)e[0].yn;
  gcdouble* __restrict__ fn = (gcdouble*)e[0].fn;
  gdouble* __restrict__ outA = (gdouble*)((K >= 2) ? e[K >= 2 ? K - 2 : 0].out : nullptr);
  gdouble* __restrict__ outB = (gdouble*)e[K - 1].out;
  ccdouble* fy = (ccdouble*)g.fy;
  ccdouble* sy = (ccdouble*)g.sy;
  ccdouble* cy = (ccdouble*)g.cy;

  auto ld_row = [&](gcdouble* p, int64_t r, double* v, bool nt) {
    int64_t rw;
    const bool ok = map_index<PER>(r, ny, rw) && in;
#pragma unroll
    for (int k = 0; k < CPL; k++) v[k] = 0.0;
    if (ok) {
      if (nt) gload<NT>(p + rw * nx + cw, v, CPL);
      else gload<false>(p + rw * nx + cw, v, CPL);
    }
  };
  auto ld_edges = [&](int64_t r, double& w, double& ee) {
    w = 0.0;
    ee = 0.0;
    int64_t rw;
    if (!map_index<PER>(r, ny, rw)) return;
    if (ld_w) w = y0[rw * nx + hw];
    if (ld_e) ee = y0[rw * nx + he];
  };

  const int64_t j0 = (int64_t)rb * g.rows;
  const int64_t j1 = (j0 + g.rows < ny) ? j0 + g.rows : ny;
  const int64_t ts = j0 - (K - 1), te = j1 + (K - 2);  // rows of level 1

  double W[K][3][CPL];  // levels 0..K-1: rows south, centre, north
  double YN[K][CPL], FN[K][CPL];
#pragma unroll
  for (int m = 0; m < K; m++)
#pragma unroll
    for (int k = 0; k < CPL; k++) {
      W[m][0][k] = W[m][1][k] = W[m][2][k] = 0.0;
      YN[m][k] = FN[m][k] = 0.0;
    }
  ld_row(y0, ts - 1, W[0][1], false);
  ld_row(y0, ts, W[0][2], false);
  double PR[CPL], PA[CPL], PY[CPL], PG[CPL], Pw, Pe;
  ld_row(y0, ts + 1, PR, false);
  ld_edges(ts, Pw, Pe);
  if (FUSED) {
    ld_row(ym, ts, PA, true);
    ld_row(yn, ts, PY, true);
    ld_row(fn, ts, PG, true);
  }

  for (int64_t t = ts; t <= te; t++) {
    if (CLDS) asm volatile("" ::: "memory");  // the coefficients are read again at every row step
    double A1[CPL];
#pragma unroll
    for (int k = 0; k < CPL; k++) {
      W[0][0][k] = W[0][1][k];
      W[0][1][k] = W[0][2][k];
      W[0][2][k] = PR[k];
      A1[k] = FUSED ? PA[k] : 0.0;
    }
    const double we = Pw, ee = Pe;
    if (FUSED) {
#pragma unroll
      for (int m = K - 1; m > 0; m--)
#pragma unroll
        for (int k = 0; k < CPL; k++) {
          YN[m][k] = YN[m - 1][k];
          FN[m][k] = FN[m - 1][k];
        }
#pragma unroll
      for (int k = 0; k < CPL; k++) {
        YN[0][k] = PY[k];
        FN[0][k] = PG[k];
      }
    }
    if (t < te) {  // operands of the next iteration
      ld_row(y0, t + 2, PR, false);
      ld_edges(t + 1, Pw, Pe);
      if (FUSED) {
        ld_row(ym, t + 1, PA, true);
        ld_row(yn, t + 1, PY, true);
        ld_row(fn, t + 1, PG, true);
      }
    }
#pragma unroll
    for (int m = 1; m <= K; m++) {
      const int64_t rho = t - (m - 1);
      const double* S = W[m - 1][0];
      const double* C = W[m - 1][1];
      const double* N = W[m - 1][2];
      double wv = from_west(C[CPL - 1]);
      double ev = from_east(C[0]);
      if (m == 1) {
        if (lane == 0) wv = we;
        if (lane == kWave - 1) ev = ee;
      }
      int64_t rw;
      const bool rok = map_index<PER>(rho, ny, rw);
      double f[CPL];
#pragma unroll
      for (int k = 0; k < CPL; k++) f[k] = 0.0;
      // wavefront-uniform row coefficients, read before the lane test
      const double Ds = rok ? fy[rw] : 0.0, Dn = rok ? fy[rw + 1] : 0.0;
      const double syj = (MMS && rok) ? sy[rw] : 0.0, cyj = (MMS && rok) ? cy[rw] : 0.0;
      if (rok && in) {
        if (CPL == 2) {
          f[0] = cell(Dw0, De0, Ds, Dn, C[0], wv, C[CPL - 1], S[0], N[0]);
          f[CPL - 1] = cell(De0, De1, Ds, Dn, C[CPL - 1], C[0], ev, S[CPL - 1], N[CPL - 1]);
        } else f[0] = cell(Dw0, De0, Ds, Dn, C[0], wv, ev, S[0], N[0]);
        if (MMS) {
          double t1 = CLDS ? lc[8 * (m - 1) + 5] : T1[CLDS ? 0 : m - 1];
          const double t2 = CLDS ? lc[8 * (m - 1) + 6] : T2[CLDS ? 0 : m - 1];
          if (SCALAR_E && !CLDS) asm volatile("" : "+s"(t1));
          f[0] = f[0] + forcing(t1, t2, sx0, cx0, syj, cyj);
          if (CPL == 2) f[CPL - 1] = f[CPL - 1] + forcing(t1, t2, sx1, cx1, syj, cyj);
        }
        if (FUSED) {
          const double* Am = (m == 1) ? A1 : W[m >= 2 ? m - 2 : 0][0];
          double cm[5];
#pragma unroll
          for (int k = 0; k < 5; k++) cm[k] = CLDS ? lc[8 * (m - 1) + k] : cc[CLDS ? 0 : m - 1][k];
#pragma unroll
          for (int k = 0; k < CPL; k++) f[k] = lc5(cm, f[k], Am[k], YN[m - 1][k], C[k], FN[m - 1][k]);
        }
      }
      if (own && rho >= j0 && rho < j1 && (m == K || m == K - 1))
        gstore<NT>(((m == K) ? outB : outA) + rho * nx + c, f, CPL);
      if (m < K) {
#pragma unroll
        for (int k = 0; k < CPL; k++) {
          W[m < K ? m : 0][0][k] = W[m < K ? m : 0][1][k];
          W[m < K ? m : 0][1][k] = W[m < K ? m : 0][2][k];
          W[m < K ? m : 0][2][k] = f[k];
        }
      }
    }
  }
}

// stage 1 (out = c0*Y_n + c1*F_n, N_VLinearSum order) over the owned cells of
// a wavefront tile of the same geometry
template <int CPL, int L>
__device__ __forceinline__ void axpy_wave(const PassGeo& g, const StageEntry* __restrict__ e,
                                          int strip, int rb)
{
  constexpr int OW = (kWave - 2 * L) * CPL;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t nx = g.nx, ny = g.ny;
  const int64_t c = (int64_t)strip * OW - (int64_t)(L * CPL) + (int64_t)lane * CPL;
  if (lane < L || lane >= kWave - L || c >= nx) return;
  const double a0 = e->c[0], a1 = e->c[1];
  const double* __restrict__ x = e->yn;
  const double* __restrict__ y = e->fn;
  double* __restrict__ z = e->out;
  const int64_t j0 = (int64_t)rb * g.rows;
  const int64_t j1 = (j0 + g.rows < ny) ? j0 + g.rows : ny;
  for (int64_t r = j0; r < j1; r++) {
    const int64_t o = r * nx + c;
    if (CPL == 2) {
      const double2 X = *reinterpret_cast<const double2*>(x + o);
      const double2 Y = *reinterpret_cast<const double2*>(y + o);
      *reinterpret_cast<double2*>(z + o) = make_double2(a0 * X.x + a1 * Y.x, a0 * X.y + a1 * Y.y);
    } else z[o] = a0 * x[o] + a1 * y[o];
  }
}

// ---------------------------------------------------------- one problem
template <int CPL, bool PER, bool MMS, int K, bool NT, bool CLDS>
__global__ __launch_bounds__(kBlock) void k_pass(const PassGeo g, const StageEntry* __restrict__ e)
{
  constexpr int L = (K - 1 + CPL - 1) / CPL;
  __shared__ double scoef[CLDS ? 8 * K : 1];
  if (CLDS) {  // every wavefront of the workgroup reads the same table
    const int i = threadIdx.x;
    if (i < 8 * K) {
      const int m = i >> 3, k = i & 7;
      scoef[i] = k < 5 ? e[m].c[k] : (k == 5 ? e[m].T1 : (k == 6 ? e[m].T2 : 0.0));
    }
    __syncthreads();
  }
  // XCD-aware: the 8 XCDs receive workgroups round-robin, so workgroup b is
  // given tile (b mod 8)*ceil(T/8) + (b div 8) and each XCD sweeps a
  // contiguous band of row blocks (their halo rows stay in its L2)
  const unsigned B = gridDim.x, b = blockIdx.x;
  const unsigned q = B >> 3, r = B & 7u, xcd = b & 7u, loc = b >> 3;
  const unsigned tile = xcd * q + (xcd < r ? xcd : r) + loc;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
  const int64_t wt = (int64_t)tile * kWavesPerBlock + wave;
  const int64_t nrb = (g.ny + g.rows - 1) / g.rows;
  if (wt >= (int64_t)g.strips * nrb) return;
  const int strip = (int)(wt % g.strips), rb = (int)(wt / g.strips);
  pass_wave<CPL, PER, MMS, K, L, true, NT, true, CLDS>(g, e, strip, rb, (lcdouble*)scoef);
}

PassGeo geo_of(const DiffusionProblem* P, int cpl, int L, int rows)
{
  PassGeo g;
  g.fx = P->d_fx;
  g.fy = P->d_fy;
  g.sx = P->d_sx;
  g.sy = P->d_sy;
  g.cx = P->d_cx;
  g.cy = P->d_cy;
  g.nx = P->p.nx;
  g.ny = P->p.ny;
  g.rows = rows;
  const int64_t ow = (int64_t)(kWave - 2 * L) * cpl;
  g.strips = (int)((P->p.nx + ow - 1) / ow);
  return g;
}

// variant: bit 0 non-temporal operand streams, bit 1 stage coefficients in LDS
template <int CPL, bool PER, bool MMS>
void launch_pass_t(const PassGeo& g, int K, int variant, const StageEntry* e, hipStream_t s)
{
  const int64_t waves = (int64_t)g.strips * ((g.ny + g.rows - 1) / g.rows);
  const unsigned grid = (unsigned)((waves + kWavesPerBlock - 1) / kWavesPerBlock);
#define CEDA_PASS(KK)                                                                                     \
  switch (variant & 3) {                                                                                  \
    case 0: hipLaunchKernelGGL((k_pass<CPL, PER, MMS, KK, false, false>), dim3(grid), dim3(kBlock), 0, s, g, e); break; \
    case 1: hipLaunchKernelGGL((k_pass<CPL, PER, MMS, KK, true, false>), dim3(grid), dim3(kBlock), 0, s, g, e); break;  \
    case 2: hipLaunchKernelGGL((k_pass<CPL, PER, MMS, KK, false, true>), dim3(grid), dim3(kBlock), 0, s, g, e); break;  \
    default: hipLaunchKernelGGL((k_pass<CPL, PER, MMS, KK, true, true>), dim3(grid), dim3(kBlock), 0, s, g, e); break; \
  }
  switch (K) {
    case 1: CEDA_PASS(1) break;
    case 2: CEDA_PASS(2) break;
    case 3: CEDA_PASS(3) break;
    case 4: CEDA_PASS(4) break;
    case 5: CEDA_PASS(5) break;
    default: CEDA_PASS(6) break;
  }
#undef CEDA_PASS
}

// ---------------------------------------------------------------- batch
// Batch tiles use one geometry for every phase of a step group: the halo
// lanes of a kBatchMaxK-stage pass (2 with two columns per lane, 4 with one).
template <int CPL>
constexpr int batch_halo() { return (kBatchMaxK - 1 + CPL - 1) / CPL; }

// KMAX: the longest pass of the launch (one kernel instance per KMAX, so a
// short-pass configuration keeps its own register allocation)
template <int CPL, bool PER, bool MMS, int KMAX>
__device__ void batch_wave(int mode, int K, const PassGeo& g, const StageEntry* e, int strip, int rb)
{
  constexpr int L = batch_halo<CPL>();
  if (mode == 1) {
    axpy_wave<CPL, L>(g, e, strip, rb);
    return;
  }
  if (mode == 3) {
    pass_wave<CPL, PER, MMS, 1, L, false, false>(g, e, strip, rb);
    return;
  }
  // 1 <= K <= KMAX (the launch's longest pass)
  if (K == 1) pass_wave<CPL, PER, MMS, 1, L, true, false>(g, e, strip, rb);
  if constexpr (KMAX >= 2)
    if (K == 2) pass_wave<CPL, PER, MMS, 2, L, true, false>(g, e, strip, rb);
  if constexpr (KMAX >= 3)
    if (K == 3) pass_wave<CPL, PER, MMS, 3, L, true, false>(g, e, strip, rb);
  if constexpr (KMAX >= 4)
    if (K == 4) pass_wave<CPL, PER, MMS, 4, L, true, false>(g, e, strip, rb);
  if constexpr (KMAX >= 5)
    if (K == 5) pass_wave<CPL, PER, MMS, 5, L, true, false>(g, e, strip, rb);
}

__device__ inline PassGeo geo_of(const BatchDesc& D)
{
  PassGeo g;
  g.fx = D.fx;
  g.fy = D.fy;
  g.sx = D.sx;
  g.sy = D.sy;
  g.cx = D.cx;
  g.cy = D.cy;
  g.nx = D.nx;
  g.ny = D.ny;
  g.rows = D.rows;
  g.strips = D.strips;
  return g;
}

// LAYOUT: the problems' column layout and boundary when the whole batch
// shares them (kLayoutC2Mms: two columns per lane, Dirichlet with forcing;
// kLayoutC2Per: two columns per lane, periodic), so that the instance holds
// only that layout's code and registers; kLayoutAny dispatches per tile.

template <int KMAX, int LAYOUT>
__global__ __launch_bounds__(kBlock) void k_batch_pass(const BatchDesc* __restrict__ desc,
                                                       const BatchTile* __restrict__ tiles,
                                                       int ntiles, const StageEntry* __restrict__ ent,
                                                       int stride, const int* __restrict__ stages,
                                                       int phase, int adaptive)
{
  constexpr int kmax = KMAX;
  const int w = blockIdx.x * kWavesPerBlock +
                __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
  if (w >= ntiles) return;
  const BatchTile t = tiles[w];
  const int s = stages[t.slot];
  const StageEntry* base = ent + (int64_t)t.slot * stride;
  int mode, K = 1;
  const StageEntry* e;
  if (phase == 0) {
    mode = 1;
    e = base + 1;
  } else {
    const int j = 2 + (phase - 1) * kmax;
    if (j <= s) {
      mode = 2;
      K = (s - j + 1 < kmax) ? s - j + 1 : kmax;
      e = base + j;
    } else if (adaptive && j - kmax <= s) {
      mode = 3;
      e = base + s + 1;
    } else return;
  }
  const BatchDesc D = desc[t.prob];
  const PassGeo g = geo_of(D);
  if constexpr (LAYOUT == kLayoutC2Mms) {
    batch_wave<2, false, true, KMAX>(mode, K, g, e, t.strip, t.rb);
    return;
  }
  if constexpr (LAYOUT == kLayoutC2Per) {
    batch_wave<2, true, false, KMAX>(mode, K, g, e, t.strip, t.rb);
    return;
  }
  if (D.cpl == 2) {
    if (D.per) batch_wave<2, true, false, KMAX>(mode, K, g, e, t.strip, t.rb);
    else batch_wave<2, false, true, KMAX>(mode, K, g, e, t.strip, t.rb);
  } else {
    if (D.per) batch_wave<1, true, false, KMAX>(mode, K, g, e, t.strip, t.rb);
    else batch_wave<1, false, true, KMAX>(mode, K, g, e, t.strip, t.rb);
  }
}

}  // namespace

int pass_halo_lanes(int K, int cpl) { return (K - 1 + cpl - 1) / cpl; }

int batch_halo_lanes(int cpl) { return cpl == 2 ? batch_halo<2>() : batch_halo<1>(); }

void launch_pass(DiffusionProblem* P, int K, int rows, int nt, const StageEntry* e, hipStream_t s)
{
  // the kernel reads e[0..K-1]; callers validate K, this keeps a bad one
  // from reaching the device
  if (K < 1 || K > kMaxPassStages || !e || rows < 1) return;
  const int cpl = (P->p.nx % 2 == 0) ? 2 : 1;
  const PassGeo g = geo_of(P, cpl, pass_halo_lanes(K, cpl), rows);
  const bool per = (P->p.kind == DIFFUSION_CEDA_PERIODIC);
  if (cpl == 2) {
    if (per) launch_pass_t<2, true, false>(g, K, nt, e, s);
    else launch_pass_t<2, false, true>(g, K, nt, e, s);
  } else {
    if (per) launch_pass_t<1, true, false>(g, K, nt, e, s);
    else launch_pass_t<1, false, true>(g, K, nt, e, s);
  }
}

void launch_batch_phase(const BatchDesc* desc, const BatchTile* tiles, int ntiles,
                        const StageEntry* ent, int stride, const int* stages, int phase, int kmax,
                        int adaptive, int layout, hipStream_t s)
{
  if (ntiles <= 0 || kmax < 1 || kmax > kBatchMaxK) return;
  const unsigned grid = (unsigned)((ntiles + kWavesPerBlock - 1) / kWavesPerBlock);
#define CEDA_BATCH(KM, LY)                                                                          \
  hipLaunchKernelGGL((k_batch_pass<KM, LY>), dim3(grid), dim3(kBlock), 0, s, desc, tiles, ntiles, ent, \
                     stride, stages, phase, adaptive);
#define CEDA_BATCH_K(LY)                                                                            \
  switch (kmax) {                                                                                   \
    case 1: CEDA_BATCH(1, LY) break;                                                                \
    case 2: CEDA_BATCH(2, LY) break;                                                                \
    case 3: CEDA_BATCH(3, LY) break;                                                                \
    case 4: CEDA_BATCH(4, LY) break;                                                                \
    default: CEDA_BATCH(5, LY) break;                                                               \
  }
  if (layout == kLayoutC2Mms) {
    CEDA_BATCH_K(kLayoutC2Mms)
  } else if (layout == kLayoutC2Per) {
    CEDA_BATCH_K(kLayoutC2Per)
  } else {
    CEDA_BATCH_K(kLayoutAny)
  }
#undef CEDA_BATCH_K
#undef CEDA_BATCH
}

}  // namespace ceda
