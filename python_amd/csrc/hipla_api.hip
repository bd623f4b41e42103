// hipla_api.hip -- C ABI of libhipla.so: drop-in entry points, context, batched plans.
//
// Drop-in functions keep the reference's names, signatures and codes (source/templates.h:358-360,
// source/matrix_cpu.c:273-349); hipla_init / hipla_finalize take the place of cusolver_create /
// cusolver_destroy (source/matrix_gpu.cu:124-169).  Unlike the reference GPU path
// (source/matrix_gpu.cu:263-307: five cudaMalloc, transposes, getrf, getrs and frees per call)
// device buffers are pooled per process, matrices are consumed row-major as given, and the
// factorisation, solve and residual check run in one kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/hipla.h"
#include "hipla_kernels.h"
#include "hipla_plan.h"

// --------------------------------------------------------------------------------------------
// error detail
// --------------------------------------------------------------------------------------------
static thread_local char g_detail[512] = "";

void hipla_set_detail (const char *fmt, ...)
{
  va_list ap;
  va_start (ap, fmt);
  vsnprintf (g_detail, sizeof (g_detail), fmt, ap);
  va_end (ap);
}

extern "C" const char *hipla_last_error_detail (void)
{
  return g_detail;
}

// --------------------------------------------------------------------------------------------
// process context
// --------------------------------------------------------------------------------------------
static std::recursive_mutex g_mu;   // entry points call each other (plan creation re-enters init)
static bool g_init = false;
static DeviceInfo g_dev;

const DeviceInfo &hipla_device_info ()
{
  return g_dev;
}

static int init_locked ()
{
  if (g_init)
    return 0;
  int ndev = 0;
  if (hipGetDeviceCount (&ndev) != hipSuccess || ndev <= 0)
  {
    hipla_set_detail ("no HIP device visible");
    return 5;
  }
  int dev = 0;
  HIPLA_CHECK (hipGetDevice (&dev));
  hipDeviceProp_t prop;
  HIPLA_CHECK (hipGetDeviceProperties (&prop, dev));
  g_dev.device = dev;
  g_dev.num_cu = prop.multiProcessorCount;
  g_dev.max_smem = prop.sharedMemPerBlock;
  if (std::strncmp (prop.gcnArchName, "gfx950", 6) != 0)
  {
    hipla_set_detail ("libhipla is built for gfx950; device is %s", prop.gcnArchName);
    return 5;
  }
  if (setup_solve_f64_attrs (g_dev.max_smem) != hipSuccess || setup_mixed_attrs (g_dev.max_smem) != hipSuccess ||
      setup_invert_attrs (g_dev.max_smem) != hipSuccess)
  {
    hipla_set_detail ("hipFuncSetAttribute(MaxDynamicSharedMemorySize=%zu) failed", g_dev.max_smem);
    return 5;
  }
  g_init = true;
  return 0;
}

int hipla_ensure_init ()
{
  std::lock_guard<std::recursive_mutex> lk (g_mu);
  return init_locked ();
}

extern "C" int hipla_init (void)
{
  return hipla_ensure_init ();
}

// --------------------------------------------------------------------------------------------
// drop-in single-matrix context: one pooled plan, grown as needed
// --------------------------------------------------------------------------------------------
struct DropIn
{
  hipStream_t stream = nullptr;
  double *dA = nullptr, *dB = nullptr, *dX = nullptr;
  int *dcode = nullptr;
  size_t capA = 0, capV = 0;
  hipla_plan plan = nullptr;
  int plan_n = -1, plan_kind = -1;
};
static DropIn g_drop;

static int drop_reserve (int n)
{
  const size_t nn = (size_t) n * n, nv = (size_t) n;
  if (!g_drop.stream)
    HIPLA_CHECK (hipStreamCreateWithFlags (&g_drop.stream, hipStreamNonBlocking));
  if (!g_drop.dcode)
    HIPLA_CHECK (hipMalloc (&g_drop.dcode, sizeof (int) * 2));
  if (nn > g_drop.capA)
  {
    if (g_drop.dA)
      HIPLA_CHECK (hipFree (g_drop.dA));
    g_drop.dA = nullptr;
    HIPLA_CHECK (hipMalloc (&g_drop.dA, sizeof (double) * nn * 2));
    g_drop.capA = nn;
  }
  if (nv > g_drop.capV)
  {
    if (g_drop.dB)
      HIPLA_CHECK (hipFree (g_drop.dB));
    g_drop.dB = nullptr;
    HIPLA_CHECK (hipMalloc (&g_drop.dB, sizeof (double) * nv * 2));
    g_drop.dX = g_drop.dB + nv;
    g_drop.capV = nv;
  }
  return 0;
}

static int drop_plan (int n, int kind)
{
  if (g_drop.plan && g_drop.plan_n == n && g_drop.plan_kind == kind)
    return 0;
  if (g_drop.plan)
    hipla_plan_destroy (g_drop.plan);
  g_drop.plan = nullptr;
  int64_t off = 0;
  int rc = hipla_plan_create (&g_drop.plan, 1, &n, &off, &off, HIPLA_MODE_FP64, 0);
  if (rc)
    return rc;
  g_drop.plan_n = n;
  g_drop.plan_kind = kind;
  return 0;
}

extern "C" int solve_matrix (double *a_matrix, double *b_matrix, int size, double *x_matrix, int nplasma)
{
  (void) nplasma;
  if (size <= 0 || !a_matrix || !b_matrix || !x_matrix)
  {
    hipla_set_detail ("solve_matrix: bad argument (size=%d)", size);
    return 5;
  }
  std::lock_guard<std::recursive_mutex> lk (g_mu);
  int rc = init_locked ();
  if (rc)
    return rc;
  if ((rc = drop_reserve (size)) || (rc = drop_plan (size, 0)))
    return rc;
  const size_t nn = (size_t) size * size;
  HIPLA_CHECK (hipMemcpyAsync (g_drop.dA, a_matrix, sizeof (double) * nn, hipMemcpyHostToDevice, g_drop.stream));
  HIPLA_CHECK (hipMemcpyAsync (g_drop.dB, b_matrix, sizeof (double) * size, hipMemcpyHostToDevice, g_drop.stream));
  rc = hipla_solve (g_drop.plan, g_drop.dA, g_drop.dB, g_drop.dX, g_drop.dcode, nullptr, g_drop.stream);
  if (rc)
    return rc;
  int code = 5;
  HIPLA_CHECK (hipMemcpyAsync (&code, g_drop.dcode, sizeof (int), hipMemcpyDeviceToHost, g_drop.stream));
  HIPLA_CHECK (hipStreamSynchronize (g_drop.stream));
  if (code == 0 || code == 2 || code == 3)
  {
    HIPLA_CHECK (hipMemcpyAsync (x_matrix, g_drop.dX, sizeof (double) * size, hipMemcpyDeviceToHost, g_drop.stream));
    HIPLA_CHECK (hipStreamSynchronize (g_drop.stream));
  }
  return code;
}

extern "C" int invert_matrix (double *matrix, double *inverted_matrix, int num_rows)
{
  if (num_rows <= 0 || !matrix || !inverted_matrix)
  {
    hipla_set_detail ("invert_matrix: bad argument (num_rows=%d)", num_rows);
    return 5;
  }
  std::lock_guard<std::recursive_mutex> lk (g_mu);
  int rc = init_locked ();
  if (rc)
    return rc;
  if ((rc = drop_reserve (num_rows)) || (rc = drop_plan (num_rows, 1)))
    return rc;
  const size_t nn = (size_t) num_rows * num_rows;
  double *dInv = g_drop.dA + nn;
  HIPLA_CHECK (hipMemcpyAsync (g_drop.dA, matrix, sizeof (double) * nn, hipMemcpyHostToDevice, g_drop.stream));
  rc = hipla_invert (g_drop.plan, g_drop.dA, dInv, g_drop.dcode, nullptr, g_drop.stream);
  if (rc)
    return rc;
  int code = 5;
  HIPLA_CHECK (hipMemcpyAsync (&code, g_drop.dcode, sizeof (int), hipMemcpyDeviceToHost, g_drop.stream));
  HIPLA_CHECK (hipStreamSynchronize (g_drop.stream));
  if (code == 0)
  {
    HIPLA_CHECK (hipMemcpyAsync (inverted_matrix, dInv, sizeof (double) * nn, hipMemcpyDeviceToHost, g_drop.stream));
    HIPLA_CHECK (hipStreamSynchronize (g_drop.stream));
  }
  return code;
}

extern "C" const char *get_matrix_error_string (int error_code)
{
  // codes and texts of source/matrix_cpu.c:58-72, plus this library's 5
  switch (error_code)
  {
  case 2:
    return "some matrix rows failing relative error check";
  case 3:
    return "some matrix rows failing absolute error check";
  case 4:
    return "Unsolvable matrix! Determinant is zero. Defaulting to no change";
  case 5:
    return "GPU linear algebra library failure (see hipla_last_error_detail)";
  default:
    return "bad return from solve_matrix";
  }
}

extern "C" int hipla_finalize (void)
{
  std::lock_guard<std::recursive_mutex> lk (g_mu);
  if (g_drop.plan)
    hipla_plan_destroy (g_drop.plan);
  if (g_drop.dA)
    (void) hipFree (g_drop.dA);
  if (g_drop.dB)
    (void) hipFree (g_drop.dB);
  if (g_drop.dcode)
    (void) hipFree (g_drop.dcode);
  if (g_drop.stream)
    (void) hipStreamDestroy (g_drop.stream);
  g_drop = DropIn ();
  return 0;
}
