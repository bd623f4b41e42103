// k_gj.hip -- batched blocked Gauss-Jordan inverse on the matrix cores (FP64 MFMA), for the
// dense macro-atom matrices that invert_matrix receives (source/matrix_cpu.c:223-257, called per
// cell at source/macro_accelerate.c:315 on N = I - Q, order nlevels_macro + 1).
//
// The row-major input A, read as a column-major matrix, is M = A^T; the inverse of M written back
// in the same place, read row-major again, is A^-1.  M is inverted in place by the blocked
// Gauss-Jordan sweep with 64 x 64 blocks, every cell of a size class in lock step.  For the
// diagonal block S of step s (O = every other block index):
//   1. M[S,S] <- D = M[S,S]^-1        k_gj_diag: unblocked sweep of the 64 x 64 block in LDS
//   2. M[S,O] <- D M[S,O]              k_gj_tile<2>
//   3. M[O,O] <- M[O,O] - M[O,S] M[S,O] k_gj_tile<4>  (M[O,S] still the old values)
//   4. M[O,S] <- -M[O,S] D             k_gj_tile<5>
// After the last step M holds M^-1.  Every product is a 64 x 64 x 64 block product on
// v_mfma_f64_16x16x4f64 (four waves, a 32 x 32 quarter each, operands staged through LDS in two
// K halves; lane l holds A[l % 16][l / 16] and B[l / 16][l % 16] of a 16 x 4 x 16 step and
// element e of its result at row l / 16 + 4 e, column l % 16).  The sweep does not pivot: it runs on cells whose A is row diagonally dominant
// (|a_ii| >= sum_j!=i |a_ij|, which N = I - Q with substochastic Q is).  M is then column
// diagonally dominant, every Schur complement stays so, and partial pivoting would never swap a
// row; k_gj_prepare hands every other cell to the GSL-order pivoting kernel (k_invert_f64).  An
// exact zero pivot (a zero column of M: A singular) gives code 1 (GSL_EDOM), as the reference.
// The inverse is not in GSL's operation order: parity is tolerance-based (relative Frobenius
// error against the FP64 oracle).
#include "hipla_dev.h"
#include "hipla_kernels.h"
#include "hipla_plan.h"

#include <algorithm>
#include <cstdlib>

#define GJ_NB 64                  // block order
#define GJ_KC 32                  // K rows staged in LDS per half
#define GJ_PA (GJ_NB + 4)         // LDS pitch of the A half (k-major: [k][row])
#define GJ_PB (GJ_KC + 2)         // LDS pitch of the B half (column-major: [col][k])

typedef double gj_d4 __attribute__ ((ext_vector_type (4)));

// ---- prepare: copy A into the output, row dominance test, hand-off list ----
__global__ __launch_bounds__ (256) void k_gj_prepare (GjArgs g)
{
  const int cell = g.list[blockIdx.x];
  const int n = g.nsz;
  const double *a = g.A + g.offA[cell];
  double *m = g.M + g.offA[cell];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __shared__ int s_bad;
  if (threadIdx.x == 0)
    s_bad = 0;
  __syncthreads ();
  int bad = 0;
  for (int r = wv; r < n; r += 4)
  {
    double off = 0.0, dg = 0.0;
    const double *ar = a + (size_t) r * n;
    double *mr = m + (size_t) r * n;
    for (int c = lane; c < n; c += 64)
    {
      const double v = ar[c];
      mr[c] = v;
      if (c == r)
        dg = fabs (v);
      else
        off += fabs (v);
    }
    for (int o = 32; o > 0; o >>= 1)
    {
      off += __shfl_xor (off, o);
      dg += __shfl_xor (dg, o);
    }
    bad |= !(dg >= off);          // NaN counts as not dominant
  }
  if (lane == 0 && bad)
    atomicOr (&s_bad, 1);
  __syncthreads ();
  if (threadIdx.x == 0)
  {
    g.codes[cell] = 0;
    g.flag[cell] = s_bad ? 1 : 0;
    if (s_bad)
    {
      const int pos = atomicAdd (g.fb, 1);
      g.fb[1 + pos] = cell;
    }
  }
}

// ---- step 1: D = M[S,S]^-1 by the unblocked sweep (no pivoting) in LDS ----
__global__ __launch_bounds__ (256) void k_gj_diag (GjArgs g)
{
  const int cell = g.list[blockIdx.x];
  if (g.flag[cell])
    return;
  const int n = g.nsz;
  const int k0 = g.sb * GJ_NB;
  double *m = g.M + g.offA[cell];
  __shared__ double B[GJ_NB][GJ_NB + 1];   // B[col][row]
  const int tid = threadIdx.x;
  for (int idx = tid; idx < GJ_NB * GJ_NB; idx += 256)
  {
    const int c = idx >> 6, r = idx & 63;
    B[c][r] = m[(size_t) (k0 + c) * n + k0 + r];
  }
  __syncthreads ();
  const int j = tid & 63;            // this thread's column
  const int r0 = (tid >> 6) * 16;    // and its 16 rows
  for (int k = 0; k < GJ_NB; k++)
  {
    const double piv = B[k][k];
    if (piv == 0.0)
    {
      // an exact zero pivot: M has a zero column (the block's Schur complement of a column
      // diagonally dominant matrix is zero only there); every thread sees the same value
      if (tid == 0)
      {
        g.flag[cell] = 2;
        g.codes[cell] = 1;
      }
      return;
    }
    const double inv = 1.0 / piv;
    const double u = B[j][k];        // row k, column j
    double f[16];
#pragma unroll
    for (int i = 0; i < 16; i++)
      f[i] = B[k][r0 + i];           // column k, this thread's rows
    __syncthreads ();
    if (j == k)
    {
#pragma unroll
      for (int i = 0; i < 16; i++)
        B[k][r0 + i] = (r0 + i == k) ? inv : -f[i] * inv;
    }
    else
    {
      const double uj = u * inv;
#pragma unroll
      for (int i = 0; i < 16; i++)
        B[j][r0 + i] = (r0 + i == k) ? uj : fma (-f[i], uj, B[j][r0 + i]);
    }
    __syncthreads ();
  }
  for (int idx = tid; idx < GJ_NB * GJ_NB; idx += 256)
  {
    const int c = idx >> 6, r = idx & 63;
    m[(size_t) (k0 + c) * n + k0 + r] = B[c][r];
  }
}

// Workgroups are dispatched to the eight XCDs round robin; number them so that consecutive tiles
// (the same cell's, which share the step's row and column panels) run on the same XCD and its L2.
__device__ __forceinline__ int xcd_item (int bid, int G)
{
  const int x = bid & 7, i = bid >> 3;
  const int q0 = G >> 3, rem = G & 7;
  return x < rem ? x * (q0 + 1) + i : rem * (q0 + 1) + (x - rem) * q0 + i;
}

// ---- steps 2-4: one 64 x 64 block product per workgroup ----
//   MODE 2: C(s, tj) = D C(s, tj)           MODE 4: C(ti, tj) -= M(ti, s) M(s, tj)
//   MODE 5: C(ti, s) = -C(ti, s) D
// The two K halves of the operands pass through registers: the second half's global loads are
// in flight while the first runs on the matrix cores.  Four workgroups share a CU (35 KB of LDS
// and under 128 VGPRs each), so one's loads overlap the others' products.
template <int MODE>
__global__ __launch_bounds__ (256, 4) void k_gj_tile (GjArgs g)
{
  const int nb = g.nsz / GJ_NB;
  const int tpc = MODE == 4 ? (nb - 1) * (nb - 1) : nb - 1;
  const int item = xcd_item (blockIdx.x, gridDim.x);
  const int ci = item / tpc, t = item - ci * tpc;
  const int cell = g.list[ci];
  if (g.flag[cell])
    return;
  const int n = g.nsz, sb = g.sb;
  int ti, tj;
  if (MODE == 4)
  {
    ti = t / (nb - 1);
    tj = t - ti * (nb - 1);
    ti += ti >= sb;
    tj += tj >= sb;
  }
  else if (MODE == 2)
  {
    ti = sb;
    tj = t + (t >= sb);
  }
  else
  {
    ti = t + (t >= sb);
    tj = sb;
  }
  const int ai = MODE == 2 ? sb : ti;     // block row of the left operand (its block column is sb)
  const int bj = MODE == 5 ? sb : tj;     // block column of the right operand (block row sb)
  const double sgn = MODE == 2 ? 1.0 : -1.0;
  double *m = g.M + g.offA[cell];
  __shared__ double As[GJ_KC][GJ_PA];     // [k][row] of the left operand, sign applied
  __shared__ double Bs[GJ_NB][GJ_PB];     // [col][k] of the right operand
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int rw = (wv >> 1) * 32, cw = (wv & 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;

  // this thread's eight elements of each operand half: A (k = idx >> 6, row = idx & 63) and
  // B (col = idx >> 5, k = idx & 31) for idx = tid + 256 e
  const double *abase = m + (size_t) (sb * GJ_NB) * n + ai * GJ_NB;
  const double *bbase = m + (size_t) (bj * GJ_NB) * n + sb * GJ_NB;
  double ra[8], rb[8];
#pragma unroll
  for (int e = 0; e < 8; e++)
  {
    const int ia = tid + 256 * e, ib = tid + 256 * e;
    ra[e] = abase[(size_t) (ia >> 6) * n + (ia & 63)];
    rb[e] = bbase[(size_t) (ib >> 5) * n + (ib & 31)];
  }

  gj_d4 acc[2][2];
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bq = 0; bq < 2; bq++)
    {
      if (MODE == 4)
      {
        // v_mfma_f64_16x16x4f64 keeps element e of lane l at row l/16 + 4 e, column l % 16
        const double *src = m + (size_t) (tj * GJ_NB + cw + 16 * bq + lr) * n + ti * GJ_NB + rw + 16 * bi + lk;
        acc[bi][bq] = gj_d4{ src[0], src[4], src[8], src[12] };
      }
      else
        acc[bi][bq] = gj_d4{ 0.0, 0.0, 0.0, 0.0 };
    }

#pragma unroll
  for (int kh = 0; kh < GJ_NB; kh += GJ_KC)
  {
#pragma unroll
    for (int e = 0; e < 8; e++)
    {
      const int ia = tid + 256 * e, ib = tid + 256 * e;
      As[ia >> 6][ia & 63] = sgn * ra[e];
      Bs[ib >> 5][ib & 31] = rb[e];
    }
    __syncthreads ();
    if (kh == 0)
    {
      // the second half's loads, in flight during the first half's products
#pragma unroll
      for (int e = 0; e < 8; e++)
      {
        const int ia = tid + 256 * e, ib = tid + 256 * e;
        ra[e] = abase[(size_t) (GJ_KC + (ia >> 6)) * n + (ia & 63)];
        rb[e] = bbase[(size_t) (ib >> 5) * n + GJ_KC + (ib & 31)];
      }
    }
#pragma unroll
    for (int k = 0; k < GJ_KC; k += 4)
    {
      double av[2], bv[2];
#pragma unroll
      for (int bi = 0; bi < 2; bi++)
        av[bi] = As[k + lk][rw + 16 * bi + lr];
#pragma unroll
      for (int bq = 0; bq < 2; bq++)
        bv[bq] = Bs[cw + 16 * bq + lr][k + lk];
#pragma unroll
      for (int bi = 0; bi < 2; bi++)
#pragma unroll
        for (int bq = 0; bq < 2; bq++)
          acc[bi][bq] = __builtin_amdgcn_mfma_f64_16x16x4f64 (av[bi], bv[bq], acc[bi][bq], 0, 0, 0);
    }
    __syncthreads ();
  }
  // every read of this tile (MODES 2 and 5 read the tile they overwrite) is done: write it
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bq = 0; bq < 2; bq++)
    {
      double *dst = m + (size_t) (tj * GJ_NB + cw + 16 * bq + lr) * n + ti * GJ_NB + rw + 16 * bi + lk;
#pragma unroll
      for (int e = 0; e < 4; e++)
        dst[4 * e] = acc[bi][bq][e];
    }
}

__global__ void k_gj_reset (int *fb)
{
  if (threadIdx.x == 0)
    fb[0] = 0;
}

// HIPLA_INVERT_SWEEP=0 in the environment sends every cell to the GSL-order pivoting kernel
bool gj_eligible (int n)
{
  static const bool on = [] {
    const char *e = getenv ("HIPLA_INVERT_SWEEP");
    return !(e && e[0] == '0');
  }();
  return on && n >= 2 * GJ_NB && n % GJ_NB == 0;
}

hipError_t gj_reset (int *fb, hipStream_t st)
{
  hipLaunchKernelGGL (k_gj_reset, dim3 (1), dim3 (64), 0, st, fb);
  return hipGetLastError ();
}

// every launch of one size class: prepare, then the nb steps of the sweep
hipError_t gj_enqueue_class (const GjArgs &g0, hipStream_t st)
{
  if (g0.ncells <= 0)
    return hipSuccess;
  GjArgs g = g0;
  const int nb = g.nsz / GJ_NB;
  hipLaunchKernelGGL (k_gj_prepare, dim3 (g.ncells), dim3 (256), 0, st, g);
  hipError_t e = hipGetLastError ();
  for (int s = 0; s < nb && e == hipSuccess; s++)
  {
    g.sb = s;
    hipLaunchKernelGGL (k_gj_diag, dim3 (g.ncells), dim3 (256), 0, st, g);
    hipLaunchKernelGGL (k_gj_tile<2>, dim3 (g.ncells * (nb - 1)), dim3 (256), 0, st, g);
    hipLaunchKernelGGL (k_gj_tile<4>, dim3 (g.ncells * (nb - 1) * (nb - 1)), dim3 (256), 0, st, g);
    hipLaunchKernelGGL (k_gj_tile<5>, dim3 (g.ncells * (nb - 1)), dim3 (256), 0, st, g);
    e = hipGetLastError ();
  }
  return e;
}
