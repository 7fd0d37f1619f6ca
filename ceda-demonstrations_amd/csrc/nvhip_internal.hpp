// Internal declarations shared by the HIP N_Vector, the problem kernels and
// the integrator.  gfx950 only: 64-lane wavefronts, 256-thread workgroups.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <mutex>

#include "ceda_sundials.h"
#include "nvector_hipnative.h"

#define CEDA_HIP_CHECK(expr)                                                   \
  do {                                                                         \
    hipError_t e__ = (expr);                                                   \
    if (e__ != hipSuccess) {                                                   \
      fprintf(stderr, "HIP error %s at %s:%d: %s\n", hipGetErrorName(e__),     \
              __FILE__, __LINE__, #expr);                                      \
      abort();                                                                 \
    }                                                                          \
  } while (0)

namespace ceda {

// HIP rejects a call that may synchronise the whole device (an allocation or
// free, a synchronous copy or set on the null stream, a stream destruction)
// while any thread's stream is being captured into a graph
// (hipErrorStreamCaptureImplicit).  Every graph capture of this library and
// every such call hold this lock, so that integrators on separate contexts
// can be driven from separate host threads.
std::recursive_mutex& device_sync_mutex();
template <class T>
inline hipError_t dev_malloc(T** p, size_t bytes)
{
  std::lock_guard<std::recursive_mutex> g(device_sync_mutex());
  return hipMalloc(p, bytes);
}
inline hipError_t dev_free(void* p)
{
  std::lock_guard<std::recursive_mutex> g(device_sync_mutex());
  return hipFree(p);
}
template <class T>
inline hipError_t host_malloc(T** p, size_t bytes, unsigned flags)
{
  std::lock_guard<std::recursive_mutex> g(device_sync_mutex());
  return hipHostMalloc(p, bytes, flags);
}
inline hipError_t host_free(void* p)
{
  std::lock_guard<std::recursive_mutex> g(device_sync_mutex());
  return hipHostFree(p);
}
// Copies and sets on the null stream, complete on return.  The contexts'
// streams are non-blocking, so nothing orders their kernels after null-stream
// work, and hipMemset, like a hipMemcpy from pageable host memory to the
// device, may return before the device memory is written: without the wait
// a kernel on a context's stream can read the destination first (a network
// created while other threads drive the GPU could have its zero fill land
// after its weights were packed).
inline hipError_t copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind)
{
  std::lock_guard<std::recursive_mutex> g(device_sync_mutex());
  const hipError_t e = hipMemcpy(dst, src, bytes, kind);
  return e != hipSuccess ? e : hipStreamSynchronize(nullptr);
}
inline hipError_t set_sync(void* dst, int value, size_t bytes)
{
  std::lock_guard<std::recursive_mutex> g(device_sync_mutex());
  const hipError_t e = hipMemset(dst, value, bytes);
  return e != hipSuccess ? e : hipStreamSynchronize(nullptr);
}
inline hipError_t graph_exec_destroy(hipGraphExec_t ge)
{
  std::lock_guard<std::recursive_mutex> g(device_sync_mutex());
  return hipGraphExecDestroy(ge);
}

constexpr int kBlock = 256;          // threads per workgroup (4 wavefronts)
constexpr int kWave = 64;            // gfx950 wavefront width
constexpr int kMaxReduceBlocks = 2048;
constexpr int kMaxAcc = 64;          // accumulators per reduction (DotProdMulti)

// One per SUNContext: the stream every vector of the context uses and the
// reduction workspace (device partials, device result, pinned host result).
struct Exec {
  hipStream_t stream = nullptr;
  double* d_partials = nullptr;      // kMaxReduceBlocks * kMaxAcc
  double* d_result = nullptr;        // kMaxAcc
  double* h_result = nullptr;        // pinned, kMaxAcc
  int num_cu = 256;
};

struct NVContent {
  sunindextype length = 0;
  bool own_data = false;
  double* dev = nullptr;
  double* host = nullptr;            // pinned mirror, lazily allocated
  Exec* exec = nullptr;
};

inline NVContent* content(N_Vector v) { return static_cast<NVContent*>(v->content); }
inline double* dptr(N_Vector v) { return content(v)->dev; }
inline sunindextype vlen(N_Vector v) { return content(v)->length; }
inline hipStream_t vstream(N_Vector v) { return content(v)->exec->stream; }

Exec* context_exec(SUNContext ctx);

// grid for a streaming kernel handling two elements per thread per pass
inline unsigned stream_grid(int64_t n)
{
  const int64_t pairs = (n + 1) / 2;
  int64_t g = (pairs + kBlock - 1) / kBlock;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;  // grid-stride beyond this
  return (unsigned)g;
}

inline unsigned reduce_grid(int64_t n)
{
  int64_t g = (n + 2 * kBlock - 1) / (2 * kBlock);
  if (g < 1) g = 1;
  if (g > kMaxReduceBlocks) g = kMaxReduceBlocks;
  return (unsigned)g;
}

// raw-pointer helpers used by the integrator's fused STS path; identical
// arithmetic (and, for the reduction, identical combine order) to the
// corresponding N_Vector operations
void raw_linear_sum(hipStream_t s, int64_t n, double a, const double* x, double b,
                    const double* y, double* z);
// sum_i (e_i * ewt_i)^2 with e = 0.8 yn - 0.8 ycur + 0.4h fn + 0.4h fs and
// ewt = 1/(rtol |yn| + atol): the LSRKStep error test without materialising
// e or ewt (same values, same reduction tree as LinearCombination + WrmsNorm)
double raw_sts_error_wsqr(Exec* ex, int64_t n, const double* yn, const double* ycur,
                          const double* fn, const double* fs, double h, double rtol,
                          double atol);
// the same sum for many problems in two launches, bitwise raw_sts_error_wsqr
// of each: problem q owns the g = sts_err_blocks(n) partials from vb0 on
struct StsErrItem {
  const double *yn, *ycur, *fn, *fs;
  double ch, rtol, atol;  // ch = 0.4 h
  int64_t n;
  int vb0, g;
};
int sts_err_blocks(int64_t n);
void launch_sts_err_batch(const StsErrItem* items, int nitems, int nvblocks, double* partials, double* out,
                          hipStream_t s);

}  // namespace ceda
