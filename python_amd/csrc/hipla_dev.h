// hipla_dev.h -- device helpers shared by the gfx950 kernels of libhipla.
//
// Wave64 reductions, the pivot key of GSL's idamax, GSL's multiplier rule, and the argument
// blocks the host side passes to the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#define HIPLA_WAVE 64

// ----------------------------------------------------------------------------------------
// Kernel argument blocks
// ----------------------------------------------------------------------------------------
struct SolveArgs
{
  const int *n;          // per-cell order
  const int64_t *offA;   // per-cell offset of A (doubles)
  const int64_t *offV;   // per-cell offset of b / x (doubles)
  const double *A;
  const double *B;
  double *X;
  int *codes;
  const int *list;       // cells handled by this launch, largest first
  int nlist;
  // optional LU factor store (HIPLA_FLAG_KEEP_FACTORS): LU row-major at LU + offA[c],
  // GSL ipiv (row swapped with step k) at piv + offV[c]
  double *LU;
  int *piv;
  // optional device-side list length (FP64 re-solve of cells the mixed solver hands back):
  // the kernels then process min(*dcount, nlist) cells, persistent over a fixed grid
  const int *dcount;
  // factor store (A's layout: the factored matrix in GSL's row order) and forward-substituted
  // right-hand side (b's layout), written by the workgroup factor kernels and read by the
  // back-substitution kernel; with HIPLA_FLAG_KEEP_FACTORS, F is the kept LU store
  double *F;
  double *Y;
  // a launch does its work only when its list length lies in [cnt_lo, cnt_hi] (otherwise every
  // workgroup returns at once): two kernel shapes launched back to back on one device-side list
  // length, each for the lengths it is faster at
  int cnt_lo = 0;
  int cnt_hi = 0x7fffffff;
};

__device__ __forceinline__ int list_len (const SolveArgs &s)
{
  return s.dcount ? min (*s.dcount, s.nlist) : s.nlist;
}

__device__ __forceinline__ bool list_in_window (const SolveArgs &s, int len)
{
  return len >= s.cnt_lo && len <= s.cnt_hi;
}

// ----------------------------------------------------------------------------------------
// Pivot key.  GSL's idamax (cblas/source_iamax_r.h:30-35) starts from max = 0 at the first
// index and replaces it only on a strictly larger |x|; NaN never wins.  Mapping |x| to its
// IEEE bit pattern (monotone for non-negative doubles) with NaN and 0 both mapped to 0, the
// GSL choice is: the largest key, ties to the smallest current row position.
// ----------------------------------------------------------------------------------------
__device__ __forceinline__ long long pivot_key (double v)
{
  const double a = fabs (v);
  return (a > 0.0) ? __double_as_longlong (a) : 0LL;
}

__device__ __forceinline__ int pivot_key32 (float v)
{
  const float a = fabsf (v);
  return (a > 0.0f) ? __float_as_int (a) : 0;
}

// ----------------------------------------------------------------------------------------
// Wave64 butterfly reductions (every lane gets the result).
// ----------------------------------------------------------------------------------------
__device__ __forceinline__ long long wave_max_i64 (long long v)
{
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1)
  {
    long long o = __shfl_xor (v, m);
    v = o > v ? o : v;
  }
  return v;
}

__device__ __forceinline__ int wave_min_i32 (int v)
{
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1)
  {
    int o = __shfl_xor (v, m);
    v = o < v ? o : v;
  }
  return v;
}

__device__ __forceinline__ int wave_max_i32 (int v)
{
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1)
  {
    int o = __shfl_xor (v, m);
    v = o > v ? o : v;
  }
  return v;
}

__device__ __forceinline__ double wave_max_f64 (double v)
{
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1)
  {
    double o = __shfl_xor (v, m);
    v = o > v ? o : v;
  }
  return v;
}

__device__ __forceinline__ double wave_sum_f64 (double v)
{
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1)
    v += __shfl_xor (v, m);
  return v;
}

__device__ __forceinline__ float wave_max_f32 (float v)
{
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1)
  {
    float o = __shfl_xor (v, m);
    v = o > v ? o : v;
  }
  return v;
}

// (key, pos) lexicographic: larger key first, then smaller position.
__device__ __forceinline__ void wave_argmax_keypos (long long &key, int &pos)
{
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1)
  {
    long long ok = __shfl_xor (key, m);
    int op = __shfl_xor (pos, m);
    if (ok > key || (ok == key && op < pos))
    {
      key = ok;
      pos = op;
    }
  }
}

__device__ __forceinline__ void wave_argmax_keypos32 (int &key, int &pos)
{
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1)
  {
    int ok = __shfl_xor (key, m);
    int op = __shfl_xor (pos, m);
    if (ok > key || (ok == key && op < pos))
    {
      key = ok;
      pos = op;
    }
  }
}

// Orders this wave's LDS accesses: every LDS op issued before completes before any after.
__device__ __forceinline__ void wave_lds_fence ()
{
  __asm__ volatile ("s_waitcnt lgkmcnt(0)" ::: "memory");
}

// GSL's multiplier: reciprocal multiply when |pivot| >= DBL_MIN, else divide (linalg/lu.c:162-176).
__device__ __forceinline__ double gsl_multiplier (double a, double piv)
{
  if (fabs (piv) >= DBL_MIN)
  {
    const double s = 1.0 / piv;
    return a * s;
  }
  return a / piv;
}

// One step of the residual test of source/matrix_cpu.c:162-185 for row i: 0 (pass), 2, 3.
__device__ __forceinline__ int residual_row_code (double test_val, double b)
{
  if (b > 0.0)
    return (fabs ((test_val - b) / test_val) > 1.e-6) ? 2 : 0;
  return (fabs (test_val - b) > 1.e-6) ? 3 : 0;
}

// Uniform (wave-wide) read of a 64-bit value from lane `src`.
__device__ __forceinline__ double wave_bcast_f64 (double v, int src)
{
  int lo = __builtin_amdgcn_readlane (__double2loint (v), src);
  int hi = __builtin_amdgcn_readlane (__double2hiint (v), src);
  return __hiloint2double (hi, lo);
}

// ----------------------------------------------------------------------------------------
// DPP wave reductions (no LDS traffic).  quad_perm [1,0,3,2] and [2,3,0,1], row_half_mirror and
// row_mirror leave every lane of a 16-lane row holding the row's result; the four rows are then
// combined from lanes 0, 16, 32, 48 with readlane, so the result is wave-uniform.
// ----------------------------------------------------------------------------------------
template <int CTRL>
__device__ __forceinline__ int dpp_i32 (int v)
{
  return __builtin_amdgcn_update_dpp (0, v, CTRL, 0xF, 0xF, false);
}

template <int CTRL>
__device__ __forceinline__ unsigned long long dpp_max_u64_step (unsigned long long v)
{
  const int lo = dpp_i32<CTRL> ((int) (unsigned) v);
  const int hi = dpp_i32<CTRL> ((int) (unsigned) (v >> 32));
  const unsigned long long o = ((unsigned long long) (unsigned) hi << 32) | (unsigned) lo;
  return o > v ? o : v;
}

__device__ __forceinline__ unsigned long long rl_u64 (unsigned long long v, int lane)
{
  const unsigned lo = (unsigned) __builtin_amdgcn_readlane ((int) (unsigned) v, lane);
  const unsigned hi = (unsigned) __builtin_amdgcn_readlane ((int) (unsigned) (v >> 32), lane);
  return ((unsigned long long) hi << 32) | lo;
}

// Wave-uniform maximum of an unsigned 64-bit key.
__device__ __forceinline__ unsigned long long wave_max_u64_dpp (unsigned long long v)
{
  v = dpp_max_u64_step<0xB1> (v);
  v = dpp_max_u64_step<0x4E> (v);
  v = dpp_max_u64_step<0x141> (v);
  v = dpp_max_u64_step<0x140> (v);
  unsigned long long r = rl_u64 (v, 0);
  const unsigned long long r1 = rl_u64 (v, 16), r2 = rl_u64 (v, 32), r3 = rl_u64 (v, 48);
  r = r1 > r ? r1 : r;
  r = r2 > r ? r2 : r;
  r = r3 > r ? r3 : r;
  return r;
}

// Wave-uniform minimum of a 32-bit int.
__device__ __forceinline__ int wave_min_i32_dpp (int v)
{
  v = min (v, dpp_i32<0xB1> (v));
  v = min (v, dpp_i32<0x4E> (v));
  v = min (v, dpp_i32<0x141> (v));
  v = min (v, dpp_i32<0x140> (v));
  int r = __builtin_amdgcn_readlane (v, 0);
  r = min (r, __builtin_amdgcn_readlane (v, 16));
  r = min (r, __builtin_amdgcn_readlane (v, 32));
  r = min (r, __builtin_amdgcn_readlane (v, 48));
  return r;
}

// Wave-uniform maximum of an unsigned 32-bit value (all 64 lanes active): within each row of 16
// lanes by quad and mirror exchanges, then across rows with the row_bcast15 / row_bcast31
// broadcasts, which leave the total in lane 63.
__device__ __forceinline__ unsigned wave_max_u32 (unsigned v)
{
  v = max (v, (unsigned) __builtin_amdgcn_update_dpp (0, (int) v, 0xB1, 0xF, 0xF, false));
  v = max (v, (unsigned) __builtin_amdgcn_update_dpp (0, (int) v, 0x4E, 0xF, 0xF, false));
  v = max (v, (unsigned) __builtin_amdgcn_update_dpp (0, (int) v, 0x141, 0xF, 0xF, false));
  v = max (v, (unsigned) __builtin_amdgcn_update_dpp (0, (int) v, 0x140, 0xF, 0xF, false));
  v = max (v, (unsigned) __builtin_amdgcn_update_dpp ((int) v, (int) v, 0x142, 0xA, 0xF, false));
  v = max (v, (unsigned) __builtin_amdgcn_update_dpp ((int) v, (int) v, 0x143, 0xC, 0xF, false));
  return (unsigned) __builtin_amdgcn_readlane ((int) v, 63);
}

// GSL pivot choice within a wave: among lanes with `alive`, the largest pivot_key, ties to the
// smallest `pos`.  Returns the winning lane (uniform), or -1 when no lane is alive.  The high 32
// bits of the key are reduced first: when a single lane holds their maximum, its key is the
// largest outright; otherwise the full keys (and positions, for ties) decide.
__device__ __forceinline__ int wave_pivot_lane (bool alive, long long key, int pos)
{
  const unsigned long long k1 = alive ? (unsigned long long) key + 1ull : 0ull;
  const unsigned hi = (unsigned) (k1 >> 32);
  const unsigned hm = wave_max_u32 (hi);
  const unsigned long long c0 = __ballot (k1 != 0ull && hi == hm);
  if (c0 == 0ull)
    return -1;
  if (__popcll (c0) == 1)
    return __ffsll ((long long) c0) - 1;
  const unsigned long long mk = wave_max_u64_dpp (k1);
  if (mk == 0ull)
    return -1;
  unsigned long long cand = __ballot (k1 == mk);
  if (__popcll (cand) == 1)
    return __ffsll ((long long) cand) - 1;
  const int mp = wave_min_i32_dpp ((k1 == mk) ? pos : 0x7fffffff);
  cand = __ballot (k1 == mk && pos == mp);
  return __ffsll ((long long) cand) - 1;
}

// Combine of per-wave pivot candidates (key -1: none; ties to the smallest position): lane w
// reads candidate w and the wave reduces them, so the combine costs one LDS read per lane and
// one DPP reduction instead of a chain of dependent reads.  Returns the winner's position and
// row (wave-uniform); nw <= 64.
__device__ __forceinline__ void combine_pivot_candidates (const long long *rkey, const int *rpos, const int *rrow,
                                                          int nw, int lane, int &pos, int &row)
{
  const bool has = lane < nw;
  const long long key = has ? rkey[lane] : -1;
  const int q = has ? rpos[lane] : 0x7fffffff;
  const int r = has ? rrow[lane] : 0;
  const int wl = wave_pivot_lane (key >= 0, key, q);
  const int src = wl >= 0 ? wl : 0;
  pos = __builtin_amdgcn_readlane (q, src);
  row = __builtin_amdgcn_readlane (r, src);
}

// v[slot] for a run-time slot, by bit masks: a select chain over a register array would be turned
// back into a dynamically indexed array, which the compiler then keeps in scratch memory.
template <int R>
__device__ __forceinline__ double pick_slot (const double (&v)[R], int slot)
{
  unsigned long long acc = 0ull;
#pragma unroll
  for (int r = 0; r < R; r++)
    acc |= (unsigned long long) __double_as_longlong (v[r]) & (0ull - (unsigned long long) (slot == r));
  return __longlong_as_double ((long long) acc);
}

// Uniform read of a double held by lane `src`.
__device__ __forceinline__ double rl_f64 (double v, int src)
{
  return __longlong_as_double ((long long) rl_u64 ((unsigned long long) __double_as_longlong (v), src));
}
