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
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
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
// allocation accounting
// --------------------------------------------------------------------------------------------
static std::atomic<long long> g_allocs{ 0 }, g_live_bytes{ 0 };
static std::mutex g_size_mu;
static std::unordered_map<void *, size_t> g_sizes;   // live device allocations and their sizes

hipError_t hipla_dev_alloc (void **ptr, size_t bytes)
{
  hipError_t e = hipMalloc (ptr, bytes);
  if (e == hipSuccess)
  {
    g_allocs++;
    g_live_bytes += (long long) bytes;
    std::lock_guard<std::mutex> lk (g_size_mu);
    g_sizes[*ptr] = bytes;
  }
  return e;
}

hipError_t hipla_dev_free (void *ptr)
{
  if (!ptr)
    return hipSuccess;
  {
    std::lock_guard<std::mutex> lk (g_size_mu);
    auto it = g_sizes.find (ptr);
    if (it != g_sizes.end ())
    {
      g_live_bytes -= (long long) it->second;
      g_sizes.erase (it);
    }
  }
  return hipFree (ptr);
}

void hipla_count_resource ()
{
  g_allocs++;
}

extern "C" int hipla_alloc_stats (int64_t *device_allocations, int64_t *device_bytes_live)
{
  if (device_allocations)
    *device_allocations = (int64_t) g_allocs.load ();
  if (device_bytes_live)
    *device_bytes_live = (int64_t) g_live_bytes.load ();
  return 0;
}

// --------------------------------------------------------------------------------------------
// drop-in single-matrix context: staging buffers grown to the largest order seen, and a small
// least-recently-used cache of one-cell FP64 plans keyed by the order alone (solve and invert
// share a plan; its inverse workspace is made on the first invert).  The reference's callers
// cycle through a few orders -- macro_pops solves one system per macro-atom element
// (source/macro_gov.c:366-444) between the ionization solves (source/matrix_ion.c:276) and the
// macro-atom inversions (source/macro_accelerate.c:315) -- so after the first pass over them a
// call allocates nothing.
// --------------------------------------------------------------------------------------------
struct DropPlan
{
  int n = -1;
  hipla_plan plan = nullptr;
  unsigned long long used = 0;
};

struct DropIn
{
  static constexpr int kPlans = 8;
  hipStream_t stream = nullptr;
  double *dA = nullptr, *dB = nullptr, *dX = nullptr;
  int *dcode = nullptr;
  size_t capA = 0, capV = 0;
  DropPlan plans[kPlans];
  unsigned long long tick = 0;
};
static DropIn g_drop;

static int drop_reserve (int n)
{
  const size_t nn = (size_t) n * n, nv = (size_t) n;
  if (!g_drop.stream)
  {
    HIPLA_CHECK (hipStreamCreateWithFlags (&g_drop.stream, hipStreamNonBlocking));
    hipla_count_resource ();
  }
  if (!g_drop.dcode)
    HIPLA_CHECK (hipla_dev_alloc (&g_drop.dcode, sizeof (int) * 2));
  if (nn > g_drop.capA)
  {
    if (g_drop.dA)
      HIPLA_CHECK (hipla_dev_free (g_drop.dA));
    g_drop.dA = nullptr;
    HIPLA_CHECK (hipla_dev_alloc (&g_drop.dA, sizeof (double) * nn * 2));
    g_drop.capA = nn;
  }
  if (nv > g_drop.capV)
  {
    if (g_drop.dB)
      HIPLA_CHECK (hipla_dev_free (g_drop.dB));
    g_drop.dB = nullptr;
    HIPLA_CHECK (hipla_dev_alloc (&g_drop.dB, sizeof (double) * nv * 2));
    g_drop.dX = g_drop.dB + nv;
    g_drop.capV = nv;
  }
  return 0;
}

// The cached plan for order n (made on a miss, replacing the least recently used one).
static int drop_plan (int n, hipla_plan &out)
{
  DropPlan *slot = nullptr;
  for (DropPlan &d : g_drop.plans)
    if (d.plan && d.n == n)
      slot = &d;
  if (!slot)
  {
    slot = &g_drop.plans[0];
    for (DropPlan &d : g_drop.plans)
    {
      if (!d.plan)
      {
        slot = &d;
        break;
      }
      if (d.used < slot->used)
        slot = &d;
    }
    if (slot->plan)
      hipla_plan_destroy (slot->plan);
    slot->plan = nullptr;
    slot->n = -1;
    int64_t off = 0;
    int rc = hipla_plan_create (&slot->plan, 1, &n, &off, &off, HIPLA_MODE_FP64, 0);
    if (rc)
      return rc;
    slot->n = n;
  }
  slot->used = ++g_drop.tick;
  out = slot->plan;
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
  hipla_plan plan = nullptr;
  if ((rc = drop_reserve (size)) || (rc = drop_plan (size, plan)))
    return rc;
  const size_t nn = (size_t) size * size;
  HIPLA_CHECK (hipMemcpyAsync (g_drop.dA, a_matrix, sizeof (double) * nn, hipMemcpyHostToDevice, g_drop.stream));
  HIPLA_CHECK (hipMemcpyAsync (g_drop.dB, b_matrix, sizeof (double) * size, hipMemcpyHostToDevice, g_drop.stream));
  rc = hipla_solve (plan, g_drop.dA, g_drop.dB, g_drop.dX, g_drop.dcode, nullptr, g_drop.stream);
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
  hipla_plan plan = nullptr;
  if ((rc = drop_reserve (num_rows)) || (rc = drop_plan (num_rows, plan)))
    return rc;
  const size_t nn = (size_t) num_rows * num_rows;
  double *dInv = g_drop.dA + nn;
  HIPLA_CHECK (hipMemcpyAsync (g_drop.dA, matrix, sizeof (double) * nn, hipMemcpyHostToDevice, g_drop.stream));
  rc = hipla_invert (plan, g_drop.dA, dInv, g_drop.dcode, nullptr, g_drop.stream);
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
  for (DropPlan &d : g_drop.plans)
    if (d.plan)
      hipla_plan_destroy (d.plan);
  (void) hipla_dev_free (g_drop.dA);
  (void) hipla_dev_free (g_drop.dB);
  (void) hipla_dev_free (g_drop.dcode);
  if (g_drop.stream)
    (void) hipStreamDestroy (g_drop.stream);
  g_drop = DropIn ();
  return 0;
}
