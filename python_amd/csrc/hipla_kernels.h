// hipla_kernels.h -- host-side launchers of the gfx950 kernels (one per kernel family).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "hipla_dev.h"

#define BLK_NB 16  // panel width of the blocked FP64 kernel (n <= 512; 8 above)
#define BLK_BS 512 // threads per workgroup of the blocked FP64 kernel
#define INV_NB 8    // panel width / backward block of the inverse kernel
#define BLK_MAXM 16 // blocked FP64 kernel: n <= 64 * BLK_MAXM (1024)

// ---- FP64, GSL operation order (k_solve_f64.hip) ----
size_t wave_f64_smem (int nc, int wpb, bool keep);
size_t blocked_f64_smem (int nmax);
hipError_t launch_wave_f64 (const SolveArgs &a, int nc, bool keep, int grid, hipStream_t st);
// workgroup-resident route (64 < n <= wgr_f64_limit ()): register-tiled factorization
// (k_tile.hip) + back substitution
hipError_t launch_wgr_f64 (const SolveArgs &a, int nmax, bool keep, int grid, hipStream_t st);
hipError_t launch_blocked_f64 (const SolveArgs &a, int nmax, int grid, bool keep, hipStream_t st,
                               bool with_bsub = true);
// the look-ahead route alone (lists known to hold at most lookahead_f64_max_cells () cells)
#define LA_ROUTE_MAXN 256
hipError_t launch_lookahead_route_f64 (const SolveArgs &a, int nmax, bool keep, hipStream_t st);
// back substitution + residual check of cells a workgroup factor kernel has factored
hipError_t launch_bsub_cells (const SolveArgs &a, int nmax, hipStream_t st);
int blocked_f64_wgs_per_cu (int nmax);   // resident workgroups per CU (persistent grid size)
hipError_t setup_solve_f64_attrs (size_t max_smem);
// one-wave panel + look-ahead factorization, n <= 256 (k_lookahead.hip); back substitution by
// k_bsub_f64 as for the blocked kernel
size_t lookahead_f64_smem ();
hipError_t launch_lookahead_f64 (const SolveArgs &a, int grid, bool keep, hipStream_t st);
hipError_t setup_lookahead_attrs (size_t max_smem);
int lookahead_f64_max_cells ();   // list lengths up to this run k_lookahead_f64 (n <= 256)
int wgr_f64_limit ();   // largest n of the workgroup-resident FP64 route
int tile_f64_limit ();
hipError_t launch_tile_f64 (const SolveArgs &a, int nmax, bool keep, int grid, hipStream_t st);
void tile_kernel_symbol (int nmax, bool keep, char *buf, size_t len);
// kernel shapes of an FP64 family for the autotuner (k_solve_f64.hip)
int fp64_family_of (int nmax);
int fp64_variants (int family, int nmax);
hipError_t launch_fp64_variant (const SolveArgs &a, int nmax, int variant, bool keep, int grid, hipStream_t st,
                                bool with_bsub = true);
// count: the group's cell count (-1: unknown), which decides between the look-ahead and the
// blocked kernel for n <= 256
void fp64_variant_symbol (int nmax, int variant, bool keep, char *buf, size_t len, int count = -1);
void fp64_kernel_symbol (int nmax, bool keep, char *buf, size_t len, int count = -1);

// ---- FP64 MFMA Gauss-Jordan inverse of row-diagonally dominant cells (k_gj.hip) ----
struct GjArgs
{
  const int64_t *offA;
  const int *list;      // the cells of one size class
  int ncells;           // in list
  int nsz;              // their order (a multiple of 64, >= 128)
  int sb;               // first diagonal block of the step
  int sw;               // diagonal blocks of the step (1 or 2)
  int r0, nr;           // the step's block range [r0, r0 + nr) (a sub-sweep of a 2 x 2 block: nr = 2)
  const double *A;      // input
  double *M;            // output, inverted in place
  int *codes;
  int *flag;            // per cell: 0 swept here, 1 handed to k_invert_f64, 2 singular
  int *fb;              // [count, cells...] handed to k_invert_f64
};
bool gj_eligible (int n);
hipError_t gj_reset (int *fb, hipStream_t st);
hipError_t gj_enqueue_class (const GjArgs &g, hipStream_t st);

// ---- mixed-precision inverse: BF16x3 MFMA sweep in FP32 + FP64 refinement (k_gjx.hip) ----
struct GjxArgs
{
  const int64_t *offA;
  const int *list;
  int ncells;
  int nsz;
  int sb;
  const double *A;      // input (M = A read column-major)
  float *Mf;            // FP32 sweep store: X0 ~ M^-1 after the sweep
  float *Rf;            // FP32 residual store
  double *X;            // output: the refined inverse
  int *codes;
  int *flag;
  int *fb;
};
#define GX_REFINE_STEPS 2   // FP64 residual refinement steps of the mixed inverse
hipError_t gjx_enqueue_class (const GjxArgs &g, int refine_steps, hipStream_t st);
