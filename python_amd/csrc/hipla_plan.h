// hipla_plan.h -- host-side structures of libhipla: device info, error macro, batch plans.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/hipla.h"

struct DeviceInfo
{
  int device = 0;
  int num_cu = 256;
  size_t max_smem = 65536;
};

const DeviceInfo &hipla_device_info ();
int hipla_ensure_init ();
void hipla_set_detail (const char *fmt, ...);

// Device allocations of the library go through these two (hipMalloc / hipFree, counted), and
// every stream or event it creates is counted by hipla_count_resource: hipla_alloc_stats reports
// the totals, so a caller can see that warm calls allocate nothing.
hipError_t hipla_dev_alloc (void **ptr, size_t bytes);
hipError_t hipla_dev_free (void *ptr);
void hipla_count_resource ();
template <typename T>
inline hipError_t hipla_dev_alloc (T **ptr, size_t bytes)
{
  return hipla_dev_alloc (reinterpret_cast<void **> (ptr), bytes);
}

#define HIPLA_CHECK(expr)                                                                        \
  do                                                                                             \
  {                                                                                              \
    hipError_t _e = (expr);                                                                      \
    if (_e != hipSuccess)                                                                        \
    {                                                                                            \
      hipla_set_detail ("%s failed: %s (%s:%d)", #expr, hipGetErrorString (_e), __FILE__, __LINE__); \
      return 5;                                                                                  \
    }                                                                                            \
  } while (0)

// Kernel shapes a size class can be routed to.
enum KernelKind
{
  KK_WAVE_F64 = 0,     // k_wave_f64<NC>            one wave per matrix, FP64 GSL order
  KK_WGR_F64 = 1,      // k_tile_f64 + k_bsub_f64   workgroup per matrix, [A|b] in registers
  KK_BLOCKED_F64 = 2,  // k_blocked_f64 + k_bsub_f64  blocked, panel in LDS, factors in place
  KK_WAVE_MIX = 3,     // k_wave_mixed<NC>          FP32 LU in registers + FP64 refinement
  KK_LDS_MIX = 4,      // k_lds_mixed               FP32 LU in LDS + FP64 refinement
  KK_BLOCKED_MIX = 5,  // k_blocked_mixed           FP32 blocked LU in HBM workspace + refinement
  KK_COUNT = 6
};

// A group of cells that one launch handles (cells sorted largest first).
struct Group
{
  int kind = 0;
  int param = 0;      // NC for wave kernels, panel width for blocked ones
  int nmax = 0;       // largest order in the group
  int first = 0;      // offset into the plan's cell list
  int count = 0;
  int grid = 0;       // persistent grid size (blocked kernels)
  int stream = 0;     // fork-join branch
  int direct = 0;     // mixed plans: 1 = the tuner routes this size class straight to FP64
  int variant = 0;    // FP64 kernel shape chosen by the tuner (launch_fp64_variant)
  double t_mixed = -1.0, t_direct = -1.0;   // tuner measurements (ms), -1 if not tuned
};

struct hipla_plan_s
{
  int ncells = 0;
  int mode = 0;
  int flags = 0;
  int max_refine = 5;
  double tol = 1e-12;
  std::vector<int> n;
  std::vector<int64_t> offA, offV;
  int64_t totA = 0, totV = 0;
  std::vector<int> list;     // cells, grouped
  std::vector<Group> groups;
  // device copies
  int *d_n = nullptr;
  int64_t *d_offA = nullptr, *d_offV = nullptr;
  int *d_list = nullptr;
  // workspaces
  int *d_fallback = nullptr;  // mixed: 5 lists [count, cells...] of cells to re-solve in FP64
  int fb_route_used[5] = { 0, 0, 0, 0, 0 };
  int *d_list_desc = nullptr;  // every cell, largest order first (inverse launches)
  double *d_inv_work = nullptr; // inverse: per-workgroup LU workspace
  // inverse: size classes swept by the MFMA Gauss-Jordan kernels (n, first, count into d_gj_list),
  // the cells left to the pivoting kernel, per-cell flags and the hand-off list
  std::vector<int> gj_n, gj_first, gj_count;
  int *d_gj_list = nullptr, *d_gj_flag = nullptr, *d_gj_fb = nullptr, *d_inv_rest = nullptr;
  int inv_nrest = 0, inv_rest_nmax = 0;
  float *d_gx_Mf = nullptr, *d_gx_Rf = nullptr;   // mixed inverse: FP32 sweep and residual stores (A's layout)
  bool inv_ready = false;
  int64_t inv_wstride = 0;
  int inv_slots = 0;
  double *d_LU = nullptr;     // keep-factors store (A's layout)
  int *d_piv = nullptr;       // keep-factors pivots (B's layout)
  double *d_F = nullptr;      // factor store of the workgroup kernels (d_LU when factors are kept)
  bool own_F = false;
  double *d_Y = nullptr;      // forward-substituted right-hand sides (B's layout)
  double *d_lam = nullptr;    // keep-factors: transposed-solve results of hipla_macro_emiss (B's layout)
  // the plan's own stream (graph capture, tuning)
  std::vector<hipStream_t> aux;
  // a second stream for the back substitution of the largest class, forked after its factor
  // kernel and joined at the end of the solve (default priority like every stream here)
  hipStream_t fork_st = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // graph cache
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  const void *gkey[6] = { nullptr, nullptr, nullptr, nullptr, nullptr, nullptr };
  int gkind = -1;
  // last solve (adjoint)
  const double *last_A = nullptr;
  int solved = 0;
};

// the FP64 kernel shape (KernelKind, param) an order n is routed to
void plan_route_fp64 (int n, int *kind, int *param);

// launch every group of a solve / invert on `st`, one after another
int plan_enqueue_solve (hipla_plan p, const double *dA, const double *dB, double *dX, int *dcodes,
                        hipla_cell_stats *dstats, hipStream_t st);
int plan_enqueue_mixed (hipla_plan p, const double *dA, const double *dB, double *dX, int *dcodes,
                        hipla_cell_stats *dstats, hipStream_t st);
int plan_enqueue_invert (hipla_plan p, const double *dA, double *dInv, int *dcodes, hipla_cell_stats *dstats,
                         hipStream_t st);
int plan_invert_prepare (hipla_plan p);   // allocates the inverse workspace (outside any capture)

// mixed plans: one group alone, the fallback-counter reset, the FP64 re-solve pass
int enqueue_mixed_group (hipla_plan p, int gi, const double *dA, const double *dB, double *dX, int *dcodes,
                         hipla_cell_stats *dstats, hipStream_t s);
int enqueue_mixed_reset (hipla_plan p, hipStream_t st);
int enqueue_mixed_fallback (hipla_plan p, const double *dA, const double *dB, double *dX, int *dcodes, hipStream_t st);

// kernel-family attribute setup (dynamic LDS above 64 KB)
hipError_t setup_mixed_attrs (size_t max_smem);
hipError_t setup_invert_attrs (size_t max_smem);
