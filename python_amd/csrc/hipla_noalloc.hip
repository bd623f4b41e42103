// hipla_noalloc.hip -- solve_matrix on device pointers with no device allocation at all.
//
// The drop-in solve_matrix stages the caller's host arrays through device buffers and a cached
// plan, both allocated on first use.  Callers that may not allocate (code under hipGraph capture,
// tight serving loops, harnesses that forbid hipMalloc) use hipla_solve_matrix_noalloc instead: the
// same FP64 GSL-order kernels on the caller's device arrays, with the one-cell descriptors and the
// factor store of the workgroup kernels in `scratch`, a file-scope device array that every GPU
// holds from module load to process end.  Its address is taken once per device with
// hipGetSymbolAddress; the array's name is never passed to a kernel or a copy from host code.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <mutex>

#include "hipla_kernels.h"
#include "hipla_plan.h"

#define NOALLOC_MAXN 1024                        // the blocked FP64 kernel's largest order
#define NOALLOC_HDR 64                           // doubles of cell descriptors
#define NOALLOC_DOUBLES (NOALLOC_HDR + (size_t) NOALLOC_MAXN * NOALLOC_MAXN + NOALLOC_MAXN)

// [0, 64): n (int), offA / offV (int64, both 0), list (int 0); then the factor store (n x n) and
// the forward-substituted right-hand side (n) of the workgroup kernels.  8.4 MB per GPU.
static __device__ __align__ (256) double scratch[NOALLOC_DOUBLES];

#define NOALLOC_MAXDEV 64
static std::mutex g_noalloc_mu;
static double *g_scratch_ptr[NOALLOC_MAXDEV];    // per device: the address of `scratch`

static double *scratch_on_current_device ()
{
  int dev = 0;
  if (hipGetDevice (&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAXDEV)
  {
    hipla_set_detail ("hipla_solve_matrix_noalloc: no current device (or device index >= %d)", NOALLOC_MAXDEV);
    return nullptr;
  }
  std::lock_guard<std::mutex> lk (g_noalloc_mu);
  if (!g_scratch_ptr[dev])
  {
    double *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress ((void **) &ptr, HIP_SYMBOL (scratch));
    if (e != hipSuccess || !ptr)
    {
      hipla_set_detail ("hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString (e));
      return nullptr;
    }
    g_scratch_ptr[dev] = ptr;
  }
  return g_scratch_ptr[dev];
}

// the one-cell descriptors, written by a kernel (no copy, no memset: capturable)
__global__ void k_noalloc_cell (double *hdr, int n)
{
  if (threadIdx.x == 0)
  {
    int *ni = reinterpret_cast<int *> (hdr);
    int64_t *off = reinterpret_cast<int64_t *> (hdr + 1);
    ni[0] = n;
    off[0] = 0;                                // offA
    off[1] = 0;                                // offV
    reinterpret_cast<int *> (hdr + 3)[0] = 0;  // list
  }
}

extern "C" int hipla_solve_matrix_noalloc (const double *dA, const double *dB, int n, double *dX, int *dcode,
                                           void *stream)
{
  if (!dA || !dB || !dX || !dcode || n <= 0 || n > NOALLOC_MAXN)
  {
    hipla_set_detail ("hipla_solve_matrix_noalloc: bad argument (1 <= n <= %d, device pointers)", NOALLOC_MAXN);
    return 5;
  }
  int rc = hipla_ensure_init ();
  if (rc)
    return rc;
  double *base = scratch_on_current_device ();
  if (!base)
    return 5;
  hipStream_t st = (hipStream_t) stream;
  hipLaunchKernelGGL (k_noalloc_cell, dim3 (1), dim3 (64), 0, st, base, n);
  HIPLA_CHECK (hipGetLastError ());
  SolveArgs a;
  a.n = reinterpret_cast<const int *> (base);
  a.offA = reinterpret_cast<const int64_t *> (base + 1);
  a.offV = reinterpret_cast<const int64_t *> (base + 2);
  a.list = reinterpret_cast<const int *> (base + 3);
  a.nlist = 1;
  a.A = dA;
  a.B = dB;
  a.X = dX;
  a.codes = dcode;
  a.LU = nullptr;
  a.piv = nullptr;
  a.dcount = nullptr;
  a.F = base + NOALLOC_HDR;
  a.Y = base + NOALLOC_HDR + (size_t) n * n;
  int kind = 0, param = 0;
  plan_route_fp64 (n, &kind, &param);
  switch (kind)
  {
  case KK_WAVE_F64:
    HIPLA_CHECK (launch_wave_f64 (a, param, false, 1, st));
    break;
  case KK_WGR_F64:
    HIPLA_CHECK (launch_wgr_f64 (a, n, false, 1, st));
    break;
  default:
    HIPLA_CHECK (launch_blocked_f64 (a, n, 1, false, st));
    break;
  }
  return 0;
}
